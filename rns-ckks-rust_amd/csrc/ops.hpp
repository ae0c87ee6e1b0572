// Launchers of the elementwise kernels (ops.hip) and the keyswitch pipeline (keyswitch.hip).
#pragma once
#include <vector>

#include "ntt.hpp"

namespace rck {

void launch_addsub(const Context& c, Batch<const uint64_t> a, Batch<const uint64_t> b, Batch<uint64_t> out, int limbs,
                   int polys, int sub, PrimeMap pm, hipStream_t s);
void launch_galois(const Context& c, Batch<const uint64_t> in, Batch<uint64_t> out, int limbs, int polys,
                   uint64_t galois, PrimeMap pm, hipStream_t s);
// g^{-1} mod 2N for a galois element g (odd, reduced, invertible; the crate's error texts otherwise)
uint64_t galois_inverse(const Context& c, uint64_t galois);
void launch_rescale(const Context& c, Batch<const uint64_t> in, Batch<uint64_t> out, int limbs_in, int polys,
                    hipStream_t s);
void launch_mul_add_u64(const Context& c, Batch<const uint64_t> a, Batch<const uint64_t> b, Batch<const uint64_t> cc,
                        bool has_c, Batch<uint64_t> out, int limbs, int polys, PrimeMap pm, hipStream_t s);
void launch_mul_f64(const Context& c, Batch<const double> a, Batch<const double> b, Batch<double> out, int limbs,
                    int polys, PrimeMap pm, hipStream_t s);
void launch_tensor(const Context& c, const double* x, double* d, int limbs, PrimeMap pm, hipStream_t s);
void launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);

// keyswitch pipeline
// Balanced base-2^30 decomposition of a coefficient-form polynomial with nq Q-limbs into
// `levels` signed digits per coefficient: digits[i * N + c] (keyswitching.rs:196-236, 381-467).
// B polynomials at poly stride `ps` elements -> digits [B][levels][N].
// ginv != 0: decompose sigma_g(poly) instead (coefficients gathered through X -> X^g, g * ginv = 1 mod 2N).
// l1 (nullable): [B][N / 64] partial sums of |digit| (each over 64 coefficients and every digit).
void launch_decompose(const Context& c, const uint64_t* poly, int nq, int levels, int32_t* digits, hipStream_t s,
                      int B = 1, long long ps = 0, unsigned long long ginv = 0, double* l1 = nullptr);
// Forward column pass of `polys` digit rows (int32 [polys][N]) under the g primes pm(0 .. g-1), each row read
// once for all of them: out[z][jl][N] (reduced doubles, the row pass's input).
void launch_digit_cols(const Context& c, const int32_t* digits, double* out, int g, int polys, PrimeMap pm,
                       hipStream_t s);
// For B ciphertexts sharing one key: acc_{0,1}[j] = sum_i NTT_j(digit_i) * key[i][{0,1}][j] for every eval
// prime j (Q-prefix nq, then naux aux primes), NTT form.  digits: [B][D][N]; key: [>=D][2][nq+naux][N]
// NTT-form doubles; part: [dsplit][B][2][nq+naux][N] partial sums over digit ranges.  Eval limbs
// are processed `group` at a time (column-pass scratch of B*D*group limbs).
// rows: ring rows per MAC workgroup (4 or 16; 8 or 16 with width 8); prefetch: load schedule (0 late
// keys, 1 early keys, 2 next digit's rows prefetched, 3 keys in groups of four positions, 4 keys
// issued before the second radix-16 phase, 5 four ciphertexts per workgroup with the next digit's
// rows copied into LDS while the current one is transformed, 6 the same with four digit ranges of
// one ciphertext per workgroup; 5 and 6 with rows 4 only).
// width: elements per thread in the MAC's row pass (16, or 8: half the registers, more resident waves).
// inv (load schedule 5 or 6, rows 4, one digit range only): the MAC's epilogue also runs the inverse row
// stages of the accumulators, first adding P * addn on the Q limbs when inv->addn is set; part is then
// in coefficient row order, ready for the inverse column pass.
struct MacFold {
    const double* addn = nullptr;  // NTT-form addends: addn + b * bs + comp * cs + limb * N
    long long bs = 0, cs = 0;
    int nq = 0;
    const double* pmodq = nullptr;  // centred P mod q_i
};
void launch_digit_mac(const Context& c, const int32_t* digits, int B, int D, const double* key, int nq, int naux,
                      double* part, int dsplit, int group, int rows, int prefetch, int width, hipStream_t s,
                      const MacFold* inv = nullptr);
// mod_down_eval_to_q (keyswitching.rs:278-322) on coefficient-form accumulators over Q-prefix
// nq || P (2 polys, u64), optionally adding add0 / add1 (nq limbs each) to the two results.
void launch_mod_down(const Context& c, const uint64_t* acc, int nq, const uint64_t* add0, const uint64_t* add1,
                     uint64_t* out, long long out_ps, hipStream_t s);

// Inverse column pass of the Q-limbs of B x 2 keyswitch accumulators (acc: [B][2][nq+K][N] after the
// inverse row pass) fused with mod-down by P (aux limbs already in coefficient form in
// accP: [B][2][K][N]) and the optional adds add0/add1 (+ b * add_ps); writes outs[b] = [c0][c1].
// ginv != 0: add0 is read through the automorphism X -> X^g (g * ginv = 1 mod 2N).
void launch_inv_cols_moddown(const Context& c, const double* acc, int nq, int B, const uint64_t* accP,
                             const uint64_t* add0, const uint64_t* add1, long long add_ps,
                             const std::vector<uint64_t*>& outs, hipStream_t s, unsigned long long ginv = 0);

}  // namespace rck
