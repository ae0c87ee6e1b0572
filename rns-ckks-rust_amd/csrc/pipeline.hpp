// Kernel compositions for the crate's operations (pipeline.hip).
#pragma once
#include "ops.hpp"

namespace rck {

// A keyswitching key resident in HBM in NTT form: data[digit][comp][limb][N] (reduced doubles)
// over level+1 Q-limbs followed by aux limbs (hybrid) or none (Q-only).
struct KeyDev {
    double* data = nullptr;
    // exact-convolution form (hybrid keys, N >= 2^12): the NTT forms under the two auxiliary primes of the
    // key's centred limbs, as the matrix-core byte planes of ks_exact.hip (data is null when the key
    // holds this form only)
    uint8_t* mdata = nullptr;
    // 64-bit integer contexts instead: wdata[digit][comp][limb][N], NTT form times R = 2^64 (wide.hpp)
    uint64_t* wdata = nullptr;
    uint32_t digits = 0, level = 0, aux = 0, limbs = 0;
    uint64_t galois = 0;
};

// exact-convolution keyswitch (ks_exact.hip)
bool exact_keyswitch_available(const Context& c, uint32_t aux);
// worst-case digits per group (every digit at 2^29)
int exact_group_size(const Context& c, int nq);
// grouping of `levels` digits the keyswitch runs for ciphertexts within the one-group bound: *G groups
// of *gsize digits (one group when typical digit magnitudes allow it, else the worst-case groups)
void exact_grouping(const Context& c, int nq, int levels, int* gsize, int* G);
// keep_data: the key keeps its digit-NTT form as well (both methods available)
void key_make_exact(Context& c, KeyDev* k, bool keep_data = false);
void key_export_exact_digit(Context& c, const KeyDev* k, uint32_t d, uint64_t* out);
// addr (nullable): addends d0, d1 after the inverse row stages (kMultFoldRows; ciphertext b's d0 at
// addr + b * addr_ps, d1 at + limbs*N), folded in times P before mod-down (add0, add1 then null)
void keyswitch_exact(Context& c, const uint64_t* poly, long long in_ps, int B, int limbs, const KeyDev* key, int levels,
                     const uint64_t* add0, const uint64_t* add1, long long add_ps, const std::vector<uint64_t*>& outs,
                     hipStream_t s, unsigned long long ginv, int spw, const double* addr = nullptr,
                     long long addr_ps = 0);

// Launch shape of the digit MAC: eval limbs per group and digit splits (0 = automatic).
struct KsShape {
    int group = 0;     // eval limbs per column-pass/MAC launch
    int dsplit = 0;    // digit ranges accumulated by separate workgroups
    int rows = 0;      // ring rows per MAC workgroup (4 or 16)
    int prefetch = -1; // load schedule: 0 late keys, 1 early keys, 2 next digit's rows prefetched, 3 keys in groups
    int width = 16;    // elements per thread in the row pass (16, or 8 with rows 8 / 16)
};
// load schedule value of an exact-convolution shape: rows = slots per wave of its multiply-accumulate
// (1, 2 or 4; width is unused and 16)
constexpr int kExactShape = 7;

inline void check_ptr(const void* p, const char* what) {
    RCK_REQUIRE(p != nullptr, std::string("null ") + what);
}

// negacyclic products a[z] * b[z] (b_ps == 0: the same b for every z) of `polys` polynomials; the NTT
// forms go to the context's workspace, or to `scratch` (poly_mul_scratch_bytes bytes) when given
size_t poly_mul_scratch_bytes(const Context& c, long long b_ps, int limbs, int polys);
void poly_mul(Context& c, const uint64_t* a, long long a_ps, const uint64_t* b, long long b_ps, uint64_t* out,
              long long out_ps, int limbs, int polys, PrimeMap pm, hipStream_t s, double* scratch = nullptr);
// What mult_batch leaves of d0 and d1 (two-pass rings): with kMultD2Only out[b] gets d2 only, and d0, d1 stay
// in the returned buffer ([B][3][limbs][N] doubles, d0 at +0 and d1 at +limbs*N of each ciphertext), in NTT
// form (kMultFoldNtt) or after the inverse row stages, before the inverse column pass (kMultFoldRows).
enum MultOut { kMultFull = 0, kMultFoldNtt = 1, kMultFoldRows = 2 };
// B products lhs[b] x rhs[b] (ciphertext stride in_ps) -> quadratic out[b] (stride out_ps), or d2 only (mode)
double* mult_batch(Context& c, const uint64_t* lhs, const uint64_t* rhs, long long in_ps, int B, int limbs,
                   uint64_t* out, long long out_ps, hipStream_t s, int mode = kMultFull);
void mult_impl(Context& c, const uint64_t* lhs, const uint64_t* rhs, int limbs, uint64_t* out, hipStream_t s);
void rescale_impl(Context& c, const uint64_t* in, int limbs, int polys, uint64_t* out, hipStream_t s);
// keyswitch of B polynomials (stride in_ps) with one key, adding add0/add1 (stride add_ps, nullable)
// into outs[b] = [c0][c1].  addn (hybrid keys, two-pass rings; nullable): NTT-form addends
// addn[b * addn_ps + {0, limbs*N}] folded into the accumulators (times P) instead of add0/add1.
// ginv != 0 (hybrid keys, two-pass rings): poly and add0 are read through the automorphism X -> X^g.
// addr (exact-convolution keyswitch only; nullable): d0, d1 addends after the inverse row stages, as in
// keyswitch_exact.
void keyswitch_batch(Context& c, const uint64_t* poly, long long in_ps, int B, int limbs, const KeyDev* key,
                     const uint64_t* add0, const uint64_t* add1, long long add_ps, const std::vector<uint64_t*>& outs,
                     hipStream_t s, KsShape shape = {}, const double* addn = nullptr, long long addn_ps = 0,
                     unsigned long long ginv = 0, const double* addr = nullptr, long long addr_ps = 0);
// rotate (RnsCkksContext::rotate, context/rotation.rs:104-127) of B ciphertexts at `limbs` limbs with one
// rotation key
void rotate_batch(Context& c, const std::vector<const uint64_t*>& as, int limbs, const KeyDev* key,
                  const std::vector<uint64_t*>& outs, hipStream_t s, KsShape shape = {});
// mult + relinearize of B ciphertext pairs (stride in_ps) at `limbs` limbs: outs[b] = [c0][c1]
void mult_relin_batch(Context& c, const uint64_t* lhs, const uint64_t* rhs, long long in_ps, int B, int limbs,
                      const KeyDev* key, const std::vector<uint64_t*>& outs, hipStream_t s, KsShape shape = {});
void keyswitch_impl(Context& c, const uint64_t* poly, int limbs, const KeyDev* key, const uint64_t* add0,
                    const uint64_t* add1, uint64_t* out, hipStream_t s);
// B same-kind operations at one level sharing a key (RCK_OP_* kinds of the C ABI): in0 as[b], in1 bs[b],
// result outs[b].  Operands need not be contiguous (they are staged when they are not equally spaced).
void run_op_group(Context& c, int kind, int limbs, const KeyDev* key, const std::vector<const uint64_t*>& as,
                  const std::vector<const uint64_t*>& bs, const std::vector<uint64_t*>& outs, hipStream_t s,
                  KsShape shape = {});
// Does a keyswitch of B ciphertexts at `limbs` limbs (E eval limbs, D digits) with this key take the
// exact convolution?  Keys with one form take it; for keys with both, an explicit shape decides, then
// the tuned times of both methods, else B >= 4.
bool ks_use_exact(const Context& c, const KeyDev* key, int limbs, int E, int D, int B, const KsShape& shape);
// the automatic digit-MAC shape for B ciphertexts with E eval limbs and D digits: a tuned entry
// when one exists for (N, limbs, E, D, B), otherwise a heuristic
KsShape ks_auto_shape(const Context& c, int limbs, int E, int D, int B, bool exact = false);
// tuned-shape table shared by the per-op calls and every batch of the process (one entry per method)
void ks_tuned_store(const Context& c, int limbs, int E, int D, int B, KsShape s, float ms, bool exact = false);
bool ks_tuned_find(const Context& c, int limbs, int E, int D, int B, KsShape* s, float* ms, bool exact = false);
// The table as text, one entry per line: "logn limbs E D B group dsplit rows prefetch width ms"; import
// validates every line before merging (returns the number of entries merged).
std::string ks_tuned_export();
int ks_tuned_import(const std::string& text);
// candidates the autotuner times for a shape
std::vector<KsShape> ks_candidates(const Context& c, int E, int D, int B, bool exact = false);

}  // namespace rck
