// Negacyclic NTT over every RNS limb, as LDS-staged radix-16 register passes.
//
// Transform: the merged (psi-twisted) Cooley-Tukey NTT with bit-reversed twiddles
//   fw[k] = psi^{bitrev_logN(k)};  output slot i holds a(psi^{2 bitrev(i) + 1}).
// A pointwise product of two such transforms, followed by the Gentleman-Sande inverse with
// iv[k] = psi^{-bitrev(k)} and N^{-1}, is the negacyclic product in Z_q[X]/(X^N + 1) — the value
// ModularPolynomial::multiply_ntt (src/poly/mod.rs:134-161) computes with a psi twist, a
// cyclic NTT, pointwise product, inverse NTT and an untwist.  The product is unique, so the
// two agree exactly.
//
// Layout: N = N1 * N2 split into two passes over a limb held limb-major in HBM.
//   column pass: the first log2(N1) stages are N2 independent N1-point transforms on the
//                columns x = r + N2*c (consecutive r are adjacent: coalesced loads);
//   row pass:    the remaining log2(N2) stages are N1 independent N2-point transforms on the
//                contiguous rows b*N2 .. b*N2+N2-1, with twiddles fw[(N1 + b) << s + i].
// Each pass keeps 16 elements per thread in registers (radix-16: 4 stages per register phase)
// and exchanges once through LDS between its two phases.  256 threads per workgroup; a
// workgroup owns 4096 elements = 32 KiB of LDS.
//
// NTT-domain storage order (two-pass rings): after the row pass thread tau holds the row
// positions x = 16 tau + e; they are stored at (x % 16) * (N2 / 16) + x / 16 within the row, so
// that for every e the lanes of a wave write (and the key multiply-accumulate later reads)
// consecutive addresses.  Every NTT-domain operand (keys, accumulators, pointwise products)
// uses this order; pointwise operations do not care, and the cyclic-mode permutation kernels
// map through ntt_phys().  Rings N <= 2^11 keep the natural merged-NTT order.
#pragma once
#include "context.hpp"
#include "modarith.hpp"

namespace rck {

// Prime index of limb l of an operand with nq Q-limbs: Q primes first, then the aux primes.
__device__ __forceinline__ int prime_of(int l, int nq, int qmax) { return l < nq ? l : qmax + (l - nq); }

// Cooley-Tukey butterfly (forward): (u, v) -> (u + v w, u - v w)
__device__ __forceinline__ void ct_bfly(double& a, double& b, double w, double q, double qinv) {
    double t = mulmod(b, w, q, qinv);
    double u = a;
    a = u + t;
    b = u - t;
}

// Gentleman-Sande butterfly (inverse): (u, v) -> (u + v, (u - v) w)
__device__ __forceinline__ void gs_bfly(double& a, double& b, double w, double q, double qinv) {
    double u = a, v = b;
    a = u + v;
    b = mulmod(u - v, w, q, qinv);
}

// Twiddle-table reads.  The tables are device (global) memory.  kTwGlobal names that address
// space, so the compiler counts these loads precisely and can keep other loads in flight across
// them; kTwScalar (the index is the same in every lane: column passes, base 1) reads through the
// scalar cache; kTwFlat is a generic load, after which the compiler waits for every outstanding
// memory operation (measured faster in the digit MAC, whose loads are all consumed at once).
enum TwMode { kTwGlobal = 0, kTwScalar = 1, kTwFlat = 2 };

template <int MODE>
__device__ __forceinline__ double tw_load(const double* tbl, int i) {
    if constexpr (MODE == kTwScalar) {
        return ((const __attribute__((address_space(4))) double*)tbl)[i];
    } else if constexpr (MODE == kTwGlobal) {
        return ((const __attribute__((address_space(1))) double*)tbl)[i];
    } else {
        return tbl[i];
    }
}

// Workgroup barrier ordering LDS only.  __syncthreads() also waits for every outstanding global
// load and store of the wave (a release fence), which would serialise one operand's or prime's
// memory traffic with the next one's transform.  Values loaded from global memory still get their
// own waits before use.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// LDS index of a row tile with one pad word per 16 (conflict-free for the x = 16 tau + e reads).
// The row pitch R + R/16 is 16, 8, 4 doubles past a multiple of 32 for R = 256, 128, 64, so the
// rows that one half-wave touches (2, 4, 8 of them) land on disjoint banks as well.
template <int R>
constexpr int row_pitch() {
    return R + R / 16;
}
template <int R>
__device__ __forceinline__ int row_lds(int row, int x) {
    return row * row_pitch<R>() + x + (x >> 4);
}

template <int E>
__device__ __forceinline__ void reduce_all(double (&v)[E], double q, double qinv) {
#pragma unroll
    for (int e = 0; e < E; e++) v[e] = reduce(v[e], q, qinv);
}

// Phase A of a 2^LOGR-point sub-transform: the thread holds x = tau + (R/16) e, e < 16, and
// runs stages 0..3 (distance bits LOGR-1 .. LOGR-4).  Twiddle of stage s, group i:
// tbl[(base << s) + i].  Inputs |v| <= 0.51q; outputs reduced (REDUCE) or, without the final
// reduction, |v| <= 3.26q (2.44q for inputs far below q).  MODE: how the twiddles are read.
template <int MODE = kTwGlobal, bool REDUCE = true>
__device__ __forceinline__ void fwd_phase_a(double (&v)[16], const double* __restrict__ tbl, int base, double q,
                                            double qinv) {
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int half = 8 >> s;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            if ((e & half) == 0) {
                double w = tw_load<MODE>(tbl, (base << s) + (e >> (4 - s)));
                ct_bfly(v[e], v[e + half], w, q, qinv);
            }
        }
    }
    if constexpr (REDUCE) reduce_all(v, q, qinv);
}

// Phase B: the thread holds x = 16 tau + e and runs stages 4..LOGR-1.
template <int LOGR, int MODE = kTwGlobal>
__device__ __forceinline__ void fwd_phase_b(double (&v)[16], const double* __restrict__ tbl, int base, int tau,
                                            double q, double qinv) {
#pragma unroll
    for (int s = 4; s < LOGR; s++) {
        const int half = (1 << LOGR) >> (s + 1);
#pragma unroll
        for (int e = 0; e < 16; e++) {
            if ((e & half) == 0) {
                double w = tw_load<MODE>(tbl, (base << s) + ((16 * tau + e) >> (LOGR - s)));
                ct_bfly(v[e], v[e + half], w, q, qinv);
            }
        }
    }
    reduce_all(v, q, qinv);
}

// Inverse Gentleman-Sande stages on 16 registers take the pair distances 1, 2, 4, 8 in that order.
// From reduced inputs (|v| <= 0.51q) a sum chain doubles per stage and a product is
// <= (0.5 + 0.133 |u - v| / q) q (modarith.hpp), so only two sets of values need reducing for every
// product input to stay <= 3.75q: before distance 4 the values v[e], e % 4 == 0 (each the sum of
// four inputs, <= 2q), and before distance 8 the values v[e], e % 8 == 1 (<= 2.27q); every value
// then stays <= 3.07q, and the phase ends with one full reduction (outputs reduced, as before).
__device__ __forceinline__ void gs_lazy_reduce(double (&v)[16], int half, double q, double qinv) {
    if (half == 4) {
#pragma unroll
        for (int e = 0; e < 16; e += 4) v[e] = reduce(v[e], q, qinv);
    } else if (half == 8) {
        v[1] = reduce(v[1], q, qinv);
        v[9] = reduce(v[9], q, qinv);
    }
}

// Inverse phase B: stages LOGR-1 .. 4 (pair distances 1, 2, .. in increasing order).
template <int LOGR, int MODE = kTwGlobal>
__device__ __forceinline__ void inv_phase_b(double (&v)[16], const double* __restrict__ tbl, int base, int tau,
                                            double q, double qinv) {
#pragma unroll
    for (int s = LOGR - 1; s >= 4; s--) {
        const int half = (1 << LOGR) >> (s + 1);
        gs_lazy_reduce(v, half, q, qinv);
#pragma unroll
        for (int e = 0; e < 16; e++) {
            if ((e & half) == 0) {
                double w = tw_load<MODE>(tbl, (base << s) + ((16 * tau + e) >> (LOGR - s)));
                gs_bfly(v[e], v[e + half], w, q, qinv);
            }
        }
    }
    reduce_all(v, q, qinv);
}

// Inverse phase A: stages 3 .. 0 on x = tau + (R/16) e (pair distances 1, 2, 4, 8; reductions as in
// inv_phase_b).
template <int MODE = kTwGlobal>
__device__ __forceinline__ void inv_phase_a(double (&v)[16], const double* __restrict__ tbl, int base, double q,
                                            double qinv) {
#pragma unroll
    for (int s = 3; s >= 0; s--) {
        const int half = 8 >> s;
        gs_lazy_reduce(v, half, q, qinv);
#pragma unroll
        for (int e = 0; e < 16; e++) {
            if ((e & half) == 0) {
                double w = tw_load<MODE>(tbl, (base << s) + (e >> (4 - s)));
                gs_bfly(v[e], v[e + half], w, q, qinv);
            }
        }
    }
    reduce_all(v, q, qinv);
}

// Operand addressing for a batch of limbs: element (z, y, idx) at base + z*ps + y*ls + idx.
template <typename T>
struct Batch {
    T* base;
    long long ls;  // limb stride (elements)
    long long ps;  // poly stride (elements)
    __device__ __forceinline__ T* at(int z, int y) const { return base + (long long)z * ps + (long long)y * ls; }
};

struct PrimeMap {
    int nq;    // Q-limbs in the operand
    int qmax;  // Q primes in the context (aux primes start there)
    int yoff;  // limb index of blockIdx.y == 0
    __device__ __forceinline__ int operator()(int y) const { return prime_of(y + yoff, nq, qmax); }
};

enum InKind { kInU64 = 0, kInI32 = 1, kInF64 = 2 };
enum OutKind { kOutU64 = 0, kOutF64 = 1 };

// Sizes of the two passes for a ring degree 2^logn (2^12 .. 2^16).
__host__ __device__ inline void split_logn(int logn, int* n1, int* n2) {
    *n1 = logn / 2;
    *n2 = logn - *n1;
}

// Physical position of NTT-output slot i (logical merged-NTT order) in an NTT-domain limb.
__host__ __device__ inline unsigned ntt_phys(unsigned i, int logn) {
    if (logn <= 11) return i;
    int n1, n2;
    split_logn(logn, &n1, &n2);
    const unsigned x = i & ((1u << n2) - 1);
    return (i & ~((1u << n2) - 1)) | ((x & 15u) << (n2 - 4)) | (x >> 4);
}

// ---------------------------------------------------------------------------------------
// Host launchers (ntt.hip)
// ---------------------------------------------------------------------------------------
// forward transform of a batch (coefficient form -> NTT form, doubles, reduced)
void launch_fwd(const Context& c, int in_kind, const void* in, long long in_ls, long long in_ps, double* out,
                long long out_ls, long long out_ps, int limbs, int polys, PrimeMap pm, hipStream_t s);
// forward row pass only / column pass only (used by fused pipelines); in2 (nullable): polys
// zsplit .. polys-1 are read from in2 (poly zsplit + k = in2's poly k, same strides as in)
void launch_fwd_cols(const Context& c, int in_kind, const void* in, long long in_ls, long long in_ps, double* out,
                     long long out_ls, long long out_ps, int limbs, int polys, PrimeMap pm, hipStream_t s,
                     const void* in2 = nullptr, int zsplit = 0);
void launch_fwd_rows(const Context& c, double* data, long long ls, long long ps, int limbs, int polys, PrimeMap pm,
                     hipStream_t s);
// inverse transform (NTT form doubles -> coefficient form u64 canonical, or reduced doubles)
// nsum > 1: the input of every (poly, limb) is the sum of nsum partial operands, sstride elements apart
void launch_inv(const Context& c, double* in, long long in_ls, long long in_ps, int out_kind, void* out,
                long long out_ls, long long out_ps, int limbs, int polys, PrimeMap pm, hipStream_t s, int nsum = 1,
                long long sstride = 0);
// addn (nullable): for polys = 2B keyswitch accumulators [B][2][limbs][N], adds P * addn (NTT form, ciphertext
// stride addn_bs, component stride addn_cs, limb stride N) to the first nq limbs
// gate (nullable, device): the launch does nothing unless *gate != 0
void launch_inv_rows(const Context& c, double* data, long long ls, long long ps, int limbs, int polys, PrimeMap pm,
                     hipStream_t s, int nsum = 1, long long sstride = 0, const double* addn = nullptr,
                     long long addn_bs = 0, long long addn_cs = 0, int nq = 0, const int* gate = nullptr);
// xl: B x [a0 a1], xr: B x [b0 b1] after the forward column pass (limbs limbs each) -> d: B x [d0 d1 d2]
// after the inverse row pass (tensor fused between the row passes); fold: d0, d1 stay in NTT form
void launch_mult_rows(const Context& c, const double* xl, const double* xr, double* d, int limbs, int B, PrimeMap pm,
                      hipStream_t s, bool fold = false);
void launch_inv_cols(const Context& c, const double* in, long long in_ls, long long in_ps, int out_kind, void* out,
                     long long out_ls, long long out_ps, int limbs, int polys, PrimeMap pm, hipStream_t s);

// The crate's cyclic NTT (natural order in and out) of `limbs` canonical u64 limbs, in place; t: a
// workspace of limbs * N doubles (unused for N <= 2^11).
void launch_ntt_cyclic(const Context& c, bool inverse, uint64_t* data, double* t, int limbs, PrimeMap pm,
                       hipStream_t s);

}  // namespace rck
