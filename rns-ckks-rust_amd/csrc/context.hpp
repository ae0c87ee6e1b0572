// Device context: the prime chain (Q primes then P primes), per-prime NTT twiddle tables and
// the basis-conversion constants the keyswitch kernels need, all resident in HBM.
//
// Mirrors RnsCkksContext / RnsContext / NttPlan (src/rns/context.rs:15-29,
// src/rns/basis.rs:24-112, src/math/ntt.rs:4-60): a context is built once per parameter set;
// the level of an operand is the length of its Q-prefix, and the eval-key basis of level l is
// Q[0..l] followed by all of P (src/rns/context/construction.rs:72-90).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace rck {

// Per-prime constants read by every kernel (one entry per prime of Q || P).
struct DevPrime {
    double q;          // modulus
    double qinv;       // fl(1/q)
    double ninv;       // N^{-1} mod q, centred
    uint64_t qu;       // q as an integer (wave-uniform integer arithmetic)
    const double* fw;  // forward twiddles: fw[k] = psi^{bitrev_logN(k)}, centred, k < N
    const double* iv;  // inverse twiddles: iv[k] = psi^{-bitrev_logN(k)}, centred
    const double* fwc; // cyclic forward twiddles: fwc[2^s + g] = omega^{bitrev_s(g) N / 2^{s+1}}, centred
    const double* ivc; // cyclic inverse twiddles: ivc[k] = fwc[k]^{-1}, centred
};

// Decomposition / basis-conversion constants (keyswitching.rs:196-322).
constexpr int kMaxQ = 64;     // Q primes supported by the decomposition kernels
constexpr int kMaxP = 4;      // aux primes supported by mod-down
constexpr int kWords = 64;    // 32-bit words of the multiword scratch (2048 bits)
constexpr int kWords30 = 72;  // 30-bit words of the decomposition's Q_l tables (2160 bits)

struct HostPrime {
    uint64_t q, g, root, root_inv, psi, psi_inv, n_inv;
};

struct Context {
    uint32_t n = 0, logn = 0;
    uint32_t L = 0, K = 0;  // number of Q and P primes
    uint32_t scale_bits = 0;
    std::vector<HostPrime> hp;  // [L + K]
    DevPrime* d_primes = nullptr;
    double* d_tables = nullptr;
    // inv_q[i * (L+K) + j] = q_j^{-1} mod q_i (centred), for i != j over the whole chain
    double* d_invq = nullptr;
    // d_qwords[(l * 2 + h) * kWords30 + w]: 30-bit words of Q_l = q_0 ... q_{l-1} (h = 0) and of
    // floor(Q_l / 2) (h = 1), l = 0..L (the decomposition's base-2^30 digits are these words)
    uint32_t* d_qwords = nullptr;
    std::vector<std::vector<uint32_t>> qwords;  // host copy
    std::vector<uint32_t> qbits;                // bits(Q_l)
    bool garner_lazy = false;  // every Q prime is below 2.5x the smallest (decomposition skips reductions)
    // mod-down constants (aux basis P): pwords = words of P, halfp = floor(P/2)
    uint32_t* d_pwords = nullptr;
    double* d_pmodq = nullptr;     // [L] P mod q_i (centred)
    double* d_pinvq = nullptr;     // [L] P^{-1} mod q_i (centred)
    double* d_pmod_aux = nullptr;  // [K][L] p_k mod q_i (centred)
    // Exact-convolution keyswitch: two auxiliary NTT primes s1, s2 (prime indices L+K, L+K+1 of
    // d_primes / the tables, not part of any operand's basis).  A digit times a key limb is computed as
    // an exact integer negacyclic convolution modulo s1 s2 and then reduced mod each eval prime.
    uint32_t S = 0;                   // auxiliary primes present (0 or 2)
    uint64_t sp[2] = {0, 0};
    double sp_inv12 = 0;              // s1^{-1} mod s2 (centred)
    double* d_spmod = nullptr;        // [L+K][2]: s1 mod p_j, s1 s2 mod p_j (centred), per eval prime
    int ks_method = 2;                // keys created from now on: 0 exact convolution (hybrid, N >= 2^12), 1 digit
                                      // NTT, 2 both forms (each keyswitch picks the faster for its batch size)
    // 64-bit integer arithmetic (wide.hpp): chosen when a modulus is at or above 2^50 + 2^46 (beyond
    // the FP64-exact kernels), or on request.  Per-prime constants, twiddles and basis-conversion
    // factors in Montgomery form (times R = 2^64), built on first use.
    bool wide = false;
    struct WidePrime* d_wp = nullptr;  // [L+K]
    uint64_t* d_wtables = nullptr;     // [L+K][4][N]
    uint64_t* d_winvq = nullptr;       // [L+K][L+K]: q_j^{-1} R mod q_i
    uint64_t* d_wpmodq = nullptr;      // [L]: P mod q_i
    uint64_t* d_wpinvq = nullptr;      // [L]: P^{-1} R mod q_i
    uint64_t* d_wpaux = nullptr;       // [K][L]: p_k R mod q_i
    uint64_t* d_wphalf = nullptr;      // [kMaxP]: floor(P / 2) as 64-bit words
    bool owns_tables = true;  // false for a scratch clone sharing another context's tables
    bool frozen = false;      // workspace may not grow (graph capture in progress)
    // keyswitch launch-shape tuning (0 = automatic): eval limbs per digit-MAC group, digit splits
    int tune_group = 0, tune_dsplit = 0;
    // scratch owned by the context, grown on demand
    std::vector<void*> ws_ptr;
    std::vector<size_t> ws_size;

    uint32_t prime_count() const { return L + K; }
    void* workspace(int slot, size_t bytes);
    // a second (default-priority, non-blocking) stream for overlapping independent launches, and a
    // small pool of timing-free events for ordering work between the two
    hipStream_t aux_stream();
    hipEvent_t event(int i);
    hipStream_t aux = nullptr;
    std::vector<hipEvent_t> events;
    // a context sharing this one's device tables with its own (empty) scratch
    Context* scratch_clone() const;
    ~Context();
};

// switch a context to the 64-bit integer path (builds its tables on first use)
void context_make_wide(Context& c);

// errors
void set_error(const std::string& msg);
const char* last_error();

struct Error {
    int code;
    std::string msg;
};

#define RCK_CHECK_HIP(expr)                                                                  \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            throw ::rck::Error{-2, std::string(#expr) + ": " + hipGetErrorString(e_)};     \
        }                                                                                    \
    } while (0)

#define RCK_REQUIRE(cond, message)                          \
    do {                                                    \
        if (!(cond)) throw ::rck::Error{-1, (message)};     \
    } while (0)

// every FP64 kernel launcher starts with this: a 64-bit integer context must never reach one
inline void require_fp64(const Context& c) {
    RCK_REQUIRE(!c.wide, "this operation has no 64-bit integer path (moduli at or above 2^50 + 2^46)");
}

// Per-kernel timing (diagnostics): when enabled, every launch is bracketed by HIP events on its
// own stream; prof_report() sums the elapsed times per kernel name.
bool prof_enabled();
struct ProfScope {
    hipStream_t s;
    const char* name;
    int slot = -1;
    ProfScope(hipStream_t s_, const char* n_);
    ~ProfScope();
};
void prof_set(bool on);
std::string prof_report();

// host modular helpers (exact, 128-bit)
uint64_t h_mulmod(uint64_t a, uint64_t b, uint64_t q);
uint64_t h_powmod(uint64_t a, uint64_t e, uint64_t q);
bool h_invmod(uint64_t a, uint64_t q, uint64_t* out);
double h_centred(uint64_t x, uint64_t q);

}  // namespace rck
