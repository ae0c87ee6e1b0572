// 64-bit integer arithmetic path for contexts whose moduli do not fit the FP64-exact kernels
// (any modulus at or above 2^50 + 2^46, up to the crate's full u64 range).
//
// The crate reduces every product with u128 arithmetic (src/math/ntt.rs:132-150), so it accepts
// any odd prime below 2^64.  This path keeps residues canonical in [0, q) as u64 and multiplies in
// Montgomery form with R = 2^64: mont(a, b) = a b R^{-1} mod q.  Constants that multiply data
// (twiddles, key limbs, basis-conversion factors) are stored pre-multiplied by R, so one mont()
// gives the plain product.  The reduction carries its 65th bit explicitly, so it is exact for every
// odd q < 2^64.
#pragma once
#include "ntt.hpp"

namespace rck {

struct WidePrime {
    uint64_t q;       // modulus
    uint64_t qn;      // -q^{-1} mod 2^64
    uint64_t r2;      // R^2 mod q (mont(x, r2) = x R mod q for any 64-bit x)
    uint64_t ninv;    // N^{-1} R mod q
    const uint64_t* fw;   // psi^{bitrev_logN(k)} R mod q, k < N
    const uint64_t* iv;   // psi^{-bitrev_logN(k)} R mod q
    const uint64_t* tw;   // psi^{-j} R mod q, natural order
    const uint64_t* twi;  // psi^{j} R mod q, natural order
};

// a b R^{-1} mod q, canonical, for a b < q 2^64 (in particular any a < 2^64 with b < q)
__device__ __forceinline__ uint64_t w_mul(uint64_t a, uint64_t b, uint64_t q, uint64_t qn) {
    const uint64_t lo = a * b, hi = __umul64hi(a, b);
    const uint64_t m = lo * qn;
    const uint64_t mh = __umul64hi(m, q);
    // (a b + m q) / 2^64 = hi + mh + carry, where the low words lo + (m q mod 2^64) sum to 0 or 2^64
    const uint64_t s = hi + mh;
    bool ov = s < hi;
    const uint64_t s2 = s + (lo != 0 ? 1u : 0u);
    ov |= s2 < s;
    // the true sum is below 2q; with the 65th bit set it exceeds q, and s2 - q wraps to the result
    return (ov || s2 >= q) ? s2 - q : s2;
}

__device__ __forceinline__ uint64_t w_add(uint64_t a, uint64_t b, uint64_t q) {
    const uint64_t s = a + b;
    return (s < a || s >= q) ? s - q : s;
}

__device__ __forceinline__ uint64_t w_sub(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a - b + q; }

// x mod q for any 64-bit x
__device__ __forceinline__ uint64_t w_red(uint64_t x, const WidePrime& P) { return w_mul(w_mul(x, P.r2, P.q, P.qn), 1, P.q, P.qn); }

// ---------------------------------------------------------------------------------------
// Host side (wide.hip)
// ---------------------------------------------------------------------------------------
// Builds the wide tables of a context (every prime of Q || P).  Called by context_create.
void wide_tables_create(Context& c, const std::vector<uint64_t>& all, uint32_t T);

// negacyclic NTT of `polys` x `limbs` u64 limbs in place (merged psi; forward: natural order in,
// bit-reversed out; inverse: bit-reversed in, natural out, times N^{-1} when scale)
void wide_ntt(const Context& c, bool inverse, uint64_t* data, long long ls, long long ps, int limbs, int polys,
              PrimeMap pm, hipStream_t s, bool scale = true);
// the crate's cyclic NTT (natural order in and out), in place
void wide_ntt_cyclic(Context& c, bool inverse, uint64_t* data, int limbs, PrimeMap pm, hipStream_t s);

void wide_addsub(const Context& c, Batch<const uint64_t> a, Batch<const uint64_t> b, Batch<uint64_t> out, int limbs,
                 int polys, int sub, PrimeMap pm, hipStream_t s);
void wide_galois(const Context& c, Batch<const uint64_t> in, Batch<uint64_t> out, int limbs, int polys,
                 unsigned long long ginv, PrimeMap pm, hipStream_t s);
void wide_rescale(const Context& c, Batch<const uint64_t> in, Batch<uint64_t> out, int limbs_in, int polys,
                  hipStream_t s);
void wide_mul_add(const Context& c, Batch<const uint64_t> a, Batch<const uint64_t> b, Batch<const uint64_t> cc,
                  bool has_c, Batch<uint64_t> out, int limbs, int polys, PrimeMap pm, hipStream_t s);
void wide_poly_mul(Context& c, const uint64_t* a, long long a_ps, const uint64_t* b, long long b_ps, uint64_t* out,
                   long long out_ps, int limbs, int polys, PrimeMap pm, hipStream_t s);
void wide_mult(Context& c, const uint64_t* lhs, const uint64_t* rhs, int limbs, uint64_t* out, hipStream_t s);

struct KeyDev;
// key: [digits][2][limbs][N] coefficient form (device) -> k->wdata (NTT form, times R)
void wide_key_create(Context& c, KeyDev* k, const uint64_t* key_dev);
void wide_key_export(Context& c, const KeyDev* k, uint64_t* out_host);
// keyswitch of one polynomial (limbs Q-limbs) with the optional adds -> out = [c0][c1]
void wide_keyswitch(Context& c, const uint64_t* poly, int limbs, const KeyDev* key, const uint64_t* add0,
                    const uint64_t* add1, uint64_t* out, hipStream_t s);

}  // namespace rck
