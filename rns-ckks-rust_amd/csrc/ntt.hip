// NTT pass kernels and their launchers (see ntt.hpp for the transform and the layout).
#include "ntt.hpp"

namespace rck {

template <int IN>
__device__ __forceinline__ double load_in(const void* p, long long idx, double q) {
    if constexpr (IN == kInU64) {
        return load_centred(static_cast<const uint64_t*>(p)[idx], q);
    } else if constexpr (IN == kInI32) {
        return (double)static_cast<const int32_t*>(p)[idx];
    } else {
        return static_cast<const double*>(p)[idx];
    }
}

// ---------------------------------------------------------------------------------------
// Column pass, forward: stages 0..LOGR-1 on the columns x = r + N2 c.
// Thread (tau, col): tid = tau * S + col, so consecutive lanes touch consecutive columns.
// ---------------------------------------------------------------------------------------
template <int LOGR, int IN, bool CYC = false>
__global__ __launch_bounds__(256) void k_fwd_cols(Batch<const void> in, int in_elem, Batch<double> out, int logn,
                                                  PrimeMap pm, const DevPrime* __restrict__ primes,
                                                  const void* in2, int zsplit) {
    constexpr int R = 1 << LOGR, TPS = R / 16, S = 256 / TPS;
    __shared__ double lds[R * S];
    const int N2 = 1 << (logn - LOGR);
    const int tid = threadIdx.x, col = tid % S, tau = tid / S;
    const int r = blockIdx.x * S + col;
    const int y = blockIdx.y, z = blockIdx.z;
    const DevPrime P = primes[pm(y)];
    const double q = P.q, qinv = P.qinv;
    // polys z >= zsplit come from in2 (same strides): both operands of a product in one launch
    const bool second = in2 != nullptr && z >= zsplit;
    const char* src = static_cast<const char*>(second ? in2 : in.base) +
                      ((long long)(second ? z - zsplit : z) * in.ps + (long long)y * in.ls) * in_elem;
    double v[16];
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = load_in<IN>(src, r + (long long)N2 * (tau + TPS * e), q);
    const double* fw = CYC ? P.fwc : P.fw;
    fwd_phase_a<kTwScalar>(v, fw, 1, q, qinv);
#pragma unroll
    for (int e = 0; e < 16; e++) lds[(tau + TPS * e) * S + col] = v[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = lds[(16 * tau + e) * S + col];
    fwd_phase_b<LOGR>(v, fw, 1, tau, q, qinv);
    double* dst = out.at(z, y);
#pragma unroll
    for (int e = 0; e < 16; e++) dst[r + (long long)N2 * (16 * tau + e)] = v[e];
}

// ---------------------------------------------------------------------------------------
// Row pass, forward: stages log2(N1) .. logN-1 on the contiguous rows (in place).
// Thread (row, tau): tid = row * TPS + tau.
// ---------------------------------------------------------------------------------------
template <int LOGR>
__global__ __launch_bounds__(256) void k_fwd_rows(Batch<double> data, int logn, PrimeMap pm,
                                                  const DevPrime* __restrict__ primes) {
    constexpr int R = 1 << LOGR, TPS = R / 16, S = 256 / TPS;
    __shared__ double lds[S * row_pitch<R>()];
    const int N1 = 1 << (logn - LOGR);
    const int tid = threadIdx.x, tau = tid % TPS, row = tid / TPS;
    const int b = blockIdx.x * S + row;
    const int y = blockIdx.y, z = blockIdx.z;
    const DevPrime P = primes[pm(y)];
    const double q = P.q, qinv = P.qinv;
    double* p = data.at(z, y) + (long long)b * R;
    double v[16];
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = p[tau + TPS * e];
    fwd_phase_a(v, P.fw, N1 + b, q, qinv);
#pragma unroll
    for (int e = 0; e < 16; e++) lds[row_lds<R>(row, tau + TPS * e)] = v[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = lds[row_lds<R>(row, 16 * tau + e)];
    fwd_phase_b<LOGR>(v, P.fw, N1 + b, tau, q, qinv);
#pragma unroll
    for (int e = 0; e < 16; e++) p[e * TPS + tau] = v[e];  // NTT-domain row order (ntt.hpp)
}

// Row pass, inverse (in place).
// nsum > 1: the input is the sum of nsum partial operands at element stride sstride (digit-split
// keyswitch accumulators); the result is written over the first one.
// addn (nullable): for the first nq limbs, P * addn[z >> 1][z & 1] (NTT form, P mod q_y = pmodq[y]) is
// added too: the relinearization terms folded into keyswitch accumulators [B][2][E][N] (mod-down by P
// then returns keyswitch + addend exactly).
// gate (nullable): the launch does nothing unless *gate != 0 (a recomputation chosen on the device).
template <int LOGR>
__global__ __launch_bounds__(256) void k_inv_rows(Batch<double> data, int logn, PrimeMap pm,
                                                  const DevPrime* __restrict__ primes, int nsum, long long sstride,
                                                  const double* __restrict__ addn, long long addn_bs,
                                                  long long addn_cs, int nq, const double* __restrict__ pmodq,
                                                  const int* __restrict__ gate) {
    constexpr int R = 1 << LOGR, TPS = R / 16, S = 256 / TPS;
    __shared__ double lds[S * row_pitch<R>()];
    if (gate != nullptr && *gate == 0) return;
    const int N1 = 1 << (logn - LOGR);
    const int tid = threadIdx.x, tau = tid % TPS, row = tid / TPS;
    const int b = blockIdx.x * S + row;
    const int y = blockIdx.y, z = blockIdx.z;
    const DevPrime P = primes[pm(y)];
    const double q = P.q, qinv = P.qinv;
    double* p = data.at(z, y) + (long long)b * R;
    double v[16];
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = p[e * TPS + tau];  // NTT-domain row order (ntt.hpp)
    const bool fold = addn != nullptr && y + pm.yoff < nq;
    if (nsum > 1 || fold) {
        for (int k = 1; k < nsum; k++) {
#pragma unroll
            for (int e = 0; e < 16; e++) v[e] += p[k * sstride + e * TPS + tau];
        }
        if (fold) {
            const double* an = addn + (long long)(z >> 1) * addn_bs + (long long)(z & 1) * addn_cs +
                               (long long)(y + pm.yoff) * (1LL << logn) + (long long)b * R;
            const double pm_q = pmodq[y + pm.yoff];
#pragma unroll
            for (int e = 0; e < 16; e++) v[e] += mulmod(an[e * TPS + tau], pm_q, q, qinv);
        }
        reduce_all(v, q, qinv);  // |sum| <= nsum * 0.51q + 0.97q < 2^53 for nsum <= 8
    }
    inv_phase_b<LOGR>(v, P.iv, N1 + b, tau, q, qinv);
#pragma unroll
    for (int e = 0; e < 16; e++) lds[row_lds<R>(row, 16 * tau + e)] = v[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = lds[row_lds<R>(row, tau + TPS * e)];
    inv_phase_a(v, P.iv, N1 + b, q, qinv);
#pragma unroll
    for (int e = 0; e < 16; e++) p[tau + TPS * e] = v[e];
}

// ---------------------------------------------------------------------------------------
// ct x ct tensor fused between the passes (multiply_ciphertexts_rns, src/rns/ops.rs:84-100):
// xl[z] = [a0 a1], xr[z] = [b0 b1] after the forward column pass; this kernel finishes the forward row
// stages of all four, forms d0 = a0 b0, d1 = a0 b1 + a1 b0, d2 = a1 b1 in registers, runs the
// inverse row stages of d0..d2 and writes them to d[z] = [d0 d1 d2].
// ---------------------------------------------------------------------------------------
template <int LOGR>
__device__ __forceinline__ void row_fwd(double (&v)[16], const double* p, double* lds, int row, int tau, int base,
                                        const double* __restrict__ fw, double q, double qinv) {
    constexpr int R = 1 << LOGR, TPS = R / 16;
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = p[tau + TPS * e];
    fwd_phase_a(v, fw, base, q, qinv);
    lds_barrier();
#pragma unroll
    for (int e = 0; e < 16; e++) lds[row_lds<R>(row, tau + TPS * e)] = v[e];
    lds_barrier();
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = lds[row_lds<R>(row, 16 * tau + e)];
    fwd_phase_b<LOGR>(v, fw, base, tau, q, qinv);
}

template <int LOGR>
__device__ __forceinline__ void row_inv_store(double (&v)[16], double* p, double* lds, int row, int tau, int base,
                                              const double* __restrict__ iv, double q, double qinv) {
    constexpr int R = 1 << LOGR, TPS = R / 16;
    inv_phase_b<LOGR>(v, iv, base, tau, q, qinv);
    lds_barrier();
#pragma unroll
    for (int e = 0; e < 16; e++) lds[row_lds<R>(row, 16 * tau + e)] = v[e];
    lds_barrier();
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = lds[row_lds<R>(row, tau + TPS * e)];
    inv_phase_a(v, iv, base, q, qinv);
#pragma unroll
    for (int e = 0; e < 16; e++) p[tau + TPS * e] = v[e];
}

// FOLD: d0 and d1 are stored in NTT form (NTT-domain row order) for a keyswitch that folds them into
// its accumulators; only d2 gets the inverse row stages.  NT threads per workgroup (NT / (R/16) rows):
// 128 for small batches, so that the grid still covers every CU.
// Work items (row block, limb, ciphertext) are numbered ciphertext fastest, and each XCD (the 8 take
// consecutive workgroups in turn) runs a contiguous range of them: the B ciphertexts of a row block
// share its twiddles in one L2 instead of every XCD fetching them again for each ciphertext.
template <int LOGR, bool FOLD, int NT>
__global__ __launch_bounds__(NT) void k_mult_rows(const double* __restrict__ xl, const double* __restrict__ xr,
                                                  double* __restrict__ d, long long P, int logn, PrimeMap pm,
                                                  const DevPrime* __restrict__ primes, int B, int nrb,
                                                  int items) {
    constexpr int R = 1 << LOGR, TPS = R / 16, S = NT / TPS;
    __shared__ double lds[S * row_pitch<R>()];
    const int w = blockIdx.x, per = gridDim.x >> 3;
    const int item = (w & 7) * per + (w >> 3);
    if (item >= items) return;
    const int z = item % B, rb = (item / B) % nrb, y = item / B / nrb;
    const int n = 1 << logn, N1 = n >> LOGR;
    const int tid = threadIdx.x, tau = tid % TPS, row = tid / TPS;
    const int b = rb * S + row;
    const DevPrime Pr = primes[pm(y)];
    const double q = Pr.q, qinv = Pr.qinv;
    const long long off = (long long)y * n + (long long)b * R;
    const double* lz = xl + (long long)z * 2 * P + off;
    const double* rz = xr + (long long)z * 2 * P + off;
    double* dz = d + (long long)z * 3 * P + off;
    double a0[16], a1[16], b0[16], b1[16];
    row_fwd<LOGR>(a0, lz, lds, row, tau, N1 + b, Pr.fw, q, qinv);
    row_fwd<LOGR>(a1, lz + P, lds, row, tau, N1 + b, Pr.fw, q, qinv);
    row_fwd<LOGR>(b0, rz, lds, row, tau, N1 + b, Pr.fw, q, qinv);
    row_fwd<LOGR>(b1, rz + P, lds, row, tau, N1 + b, Pr.fw, q, qinv);
    double t[16];
#pragma unroll
    for (int e = 0; e < 16; e++) {
        t[e] = reduce(mulmod(a0[e], b1[e], q, qinv) + mulmod(a1[e], b0[e], q, qinv), q, qinv);
        a0[e] = reduce(mulmod(a0[e], b0[e], q, qinv), q, qinv);
        a1[e] = reduce(mulmod(a1[e], b1[e], q, qinv), q, qinv);
    }
    if constexpr (FOLD) {
#pragma unroll
        for (int e = 0; e < 16; e++) {
            dz[e * TPS + tau] = a0[e];
            dz[P + e * TPS + tau] = t[e];
        }
    } else {
        row_inv_store<LOGR>(a0, dz, lds, row, tau, N1 + b, Pr.iv, q, qinv);
        row_inv_store<LOGR>(t, dz + P, lds, row, tau, N1 + b, Pr.iv, q, qinv);
    }
    row_inv_store<LOGR>(a1, dz + 2 * P, lds, row, tau, N1 + b, Pr.iv, q, qinv);
}

// Column pass, inverse, with the N^{-1} scaling; writes canonical u64 or reduced doubles.
template <int LOGR, int OUT, bool CYC = false>
__global__ __launch_bounds__(256) void k_inv_cols(Batch<const double> in, Batch<void> out, int out_elem, int logn,
                                                  PrimeMap pm, const DevPrime* __restrict__ primes) {
    constexpr int R = 1 << LOGR, TPS = R / 16, S = 256 / TPS;
    __shared__ double lds[R * S];
    const int N2 = 1 << (logn - LOGR);
    const int tid = threadIdx.x, col = tid % S, tau = tid / S;
    const int r = blockIdx.x * S + col;
    const int y = blockIdx.y, z = blockIdx.z;
    const DevPrime P = primes[pm(y)];
    const double q = P.q, qinv = P.qinv;
    const double* src = in.at(z, y);
    double v[16];
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = src[r + (long long)N2 * (16 * tau + e)];
    const double* iv = CYC ? P.ivc : P.iv;
    inv_phase_b<LOGR>(v, iv, 1, tau, q, qinv);
#pragma unroll
    for (int e = 0; e < 16; e++) lds[(16 * tau + e) * S + col] = v[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = lds[(tau + TPS * e) * S + col];
    inv_phase_a<kTwScalar>(v, iv, 1, q, qinv);
    char* dst = static_cast<char*>(out.base) + ((long long)z * out.ps + (long long)y * out.ls) * out_elem;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const long long idx = r + (long long)N2 * (tau + TPS * e);
        double x = mulmod(v[e], P.ninv, q, qinv);
        if constexpr (OUT == kOutU64) {
            reinterpret_cast<uint64_t*>(dst)[idx] = f64_to_u64(canon(x, q, qinv));
        } else {
            reinterpret_cast<double*>(dst)[idx] = reduce(x, q, qinv);
        }
    }
}

// ---------------------------------------------------------------------------------------
// Cyclic NTT, natural order in and out (NttPlan::forward / inverse, src/math/ntt.rs:78-89).
// The column and row passes above run unchanged with the cyclic tables fwc / ivc (context.hip),
// which makes slot i of the row-major result hold X_{brv_logN(i)}.  For slot i = b N2 + x that is
// X_k with k = brv(x) N1 + brv(b) (brv over log2 N2 and log2 N1 bits).  A workgroup therefore takes
// the S rows b = brv(beta), beta = beta0 .. beta0 + S - 1, so that for every x its S outputs are the
// consecutive k = brv(x) N1 + beta: after the row stages the tile goes through LDS once more and
// each store instruction writes S-element runs.  The inverse row pass gathers the same way.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ int brv_bits(int x, int bits) { return (int)(__brev((unsigned)x) >> (32 - bits)); }

template <int LOGR>
__global__ __launch_bounds__(256) void k_cyc_fwd_rows(const double* __restrict__ in, uint64_t* __restrict__ out,
                                                      int logn, PrimeMap pm, const DevPrime* __restrict__ primes) {
    constexpr int R = 1 << LOGR, TPS = R / 16, S = 256 / TPS;
    __shared__ double lds[S * row_pitch<R>()];
    const int n1b = logn - LOGR, N1 = 1 << n1b;
    const int tid = threadIdx.x, tau = tid % TPS, row = tid / TPS;
    const int beta0 = blockIdx.x * S, b = brv_bits(beta0 + row, n1b);
    const int y = blockIdx.y;
    const DevPrime P = primes[pm(y)];
    const double q = P.q, qinv = P.qinv;
    const long long limb = (long long)y << logn;
    const double* p = in + limb + (long long)b * R;
    double v[16];
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = p[tau + TPS * e];
    fwd_phase_a(v, P.fwc, N1 + b, q, qinv);
#pragma unroll
    for (int e = 0; e < 16; e++) lds[row_lds<R>(row, tau + TPS * e)] = v[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = lds[row_lds<R>(row, 16 * tau + e)];
    fwd_phase_b<LOGR>(v, P.fwc, N1 + b, tau, q, qinv);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; e++) lds[row_lds<R>(row, 16 * tau + e)] = v[e];
    __syncthreads();
    // output run: lane -> (row offset bo = tid % S, m = tid / S + TPS e), k = m N1 + beta0 + bo
    const int bo = tid % S, m0 = tid / S;
    uint64_t* dst = out + limb + beta0 + bo;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int m = m0 + TPS * e;
        dst[(long long)m * N1] = f64_to_u64(canon(lds[row_lds<R>(bo, brv_bits(m, LOGR))], q, qinv));
    }
}

template <int LOGR>
__global__ __launch_bounds__(256) void k_cyc_inv_rows(const uint64_t* __restrict__ in, double* __restrict__ out,
                                                      int logn, PrimeMap pm, const DevPrime* __restrict__ primes) {
    constexpr int R = 1 << LOGR, TPS = R / 16, S = 256 / TPS;
    __shared__ double lds[S * row_pitch<R>()];
    const int n1b = logn - LOGR, N1 = 1 << n1b;
    const int tid = threadIdx.x, tau = tid % TPS, row = tid / TPS;
    const int beta0 = blockIdx.x * S, b = brv_bits(beta0 + row, n1b);
    const int y = blockIdx.y;
    const DevPrime P = primes[pm(y)];
    const double q = P.q, qinv = P.qinv;
    const long long limb = (long long)y << logn;
    const int bo = tid % S, m0 = tid / S;
    const uint64_t* src = in + limb + beta0 + bo;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int m = m0 + TPS * e;
        lds[row_lds<R>(bo, brv_bits(m, LOGR))] = load_centred(src[(long long)m * N1], q);
    }
    __syncthreads();
    double v[16];
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = lds[row_lds<R>(row, 16 * tau + e)];
    inv_phase_b<LOGR>(v, P.ivc, N1 + b, tau, q, qinv);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; e++) lds[row_lds<R>(row, 16 * tau + e)] = v[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = lds[row_lds<R>(row, tau + TPS * e)];
    inv_phase_a(v, P.ivc, N1 + b, q, qinv);
    double* p = out + limb + (long long)b * R;
#pragma unroll
    for (int e = 0; e < 16; e++) p[tau + TPS * e] = v[e];
}

// Small rings: the whole limb in LDS, in place, the bit reversal folded into the load (inverse) or the
// store (forward).
template <bool INV>
__global__ __launch_bounds__(256) void k_small_cyc(uint64_t* data, int logn, PrimeMap pm,
                                                   const DevPrime* __restrict__ primes) {
    __shared__ double a[2048];
    const int n = 1 << logn, y = blockIdx.y;
    const DevPrime P = primes[pm(y)];
    const double q = P.q, qinv = P.qinv;
    uint64_t* d = data + ((long long)y << logn);
    for (int i = threadIdx.x; i < n; i += blockDim.x) a[INV ? brv_bits(i, logn) : i] = load_centred(d[i], q);
    __syncthreads();
    if constexpr (!INV) {
        for (int s = 0; s < logn; s++) {
            const int t = n >> (s + 1);
            for (int k = threadIdx.x; k < n / 2; k += blockDim.x) {
                const int g = k / t, j = k % t, x = 2 * g * t + j;
                double u = a[x], v = a[x + t];
                ct_bfly(u, v, P.fwc[(1 << s) + g], q, qinv);
                a[x] = reduce(u, q, qinv);
                a[x + t] = reduce(v, q, qinv);
            }
            __syncthreads();
        }
        for (int i = threadIdx.x; i < n; i += blockDim.x) d[i] = f64_to_u64(canon(a[brv_bits(i, logn)], q, qinv));
    } else {
        for (int s = logn - 1; s >= 0; s--) {
            const int t = n >> (s + 1);
            for (int k = threadIdx.x; k < n / 2; k += blockDim.x) {
                const int g = k / t, j = k % t, x = 2 * g * t + j;
                double u = a[x], v = a[x + t];
                gs_bfly(u, v, P.ivc[(1 << s) + g], q, qinv);
                a[x] = reduce(u, q, qinv);
                a[x + t] = reduce(v, q, qinv);
            }
            __syncthreads();
        }
        for (int i = threadIdx.x; i < n; i += blockDim.x)
            d[i] = f64_to_u64(canon(mulmod(a[i], P.ninv, q, qinv), q, qinv));
    }
}

// ---------------------------------------------------------------------------------------
// Small rings (N <= 2^11): one workgroup per limb, every stage through LDS.
// ---------------------------------------------------------------------------------------
template <int IN>
__global__ __launch_bounds__(256) void k_small_fwd(Batch<const void> in, int in_elem, Batch<double> out, int logn,
                                                   PrimeMap pm, const DevPrime* __restrict__ primes) {
    __shared__ double a[2048];
    const int n = 1 << logn, y = blockIdx.y, z = blockIdx.z;
    const DevPrime P = primes[pm(y)];
    const double q = P.q, qinv = P.qinv;
    const char* src = static_cast<const char*>(in.base) + ((long long)z * in.ps + (long long)y * in.ls) * in_elem;
    for (int i = threadIdx.x; i < n; i += blockDim.x) a[i] = load_in<IN>(src, i, q);
    __syncthreads();
    for (int s = 0; s < logn; s++) {
        const int t = n >> (s + 1);
        for (int k = threadIdx.x; k < n / 2; k += blockDim.x) {
            const int g = k / t, j = k % t, x = 2 * g * t + j;
            double u = a[x], v = a[x + t];
            ct_bfly(u, v, P.fw[(1 << s) + g], q, qinv);
            a[x] = reduce(u, q, qinv);
            a[x + t] = reduce(v, q, qinv);
        }
        __syncthreads();
    }
    double* dst = out.at(z, y);
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = a[i];
}

template <int OUT>
__global__ __launch_bounds__(256) void k_small_inv(Batch<const double> in, Batch<void> out, int out_elem, int logn,
                                                   PrimeMap pm, const DevPrime* __restrict__ primes) {
    __shared__ double a[2048];
    const int n = 1 << logn, y = blockIdx.y, z = blockIdx.z;
    const DevPrime P = primes[pm(y)];
    const double q = P.q, qinv = P.qinv;
    const double* src = in.at(z, y);
    for (int i = threadIdx.x; i < n; i += blockDim.x) a[i] = src[i];
    __syncthreads();
    for (int s = logn - 1; s >= 0; s--) {
        const int t = n >> (s + 1);
        for (int k = threadIdx.x; k < n / 2; k += blockDim.x) {
            const int g = k / t, j = k % t, x = 2 * g * t + j;
            double u = a[x], v = a[x + t];
            gs_bfly(u, v, P.iv[(1 << s) + g], q, qinv);
            a[x] = reduce(u, q, qinv);
            a[x + t] = reduce(v, q, qinv);
        }
        __syncthreads();
    }
    char* dst = static_cast<char*>(out.base) + ((long long)z * out.ps + (long long)y * out.ls) * out_elem;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double x = mulmod(a[i], P.ninv, q, qinv);
        if constexpr (OUT == kOutU64)
            reinterpret_cast<uint64_t*>(dst)[i] = f64_to_u64(canon(x, q, qinv));
        else
            reinterpret_cast<double*>(dst)[i] = reduce(x, q, qinv);
    }
}

// ---------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------
static int elem_size(int kind) { return kind == kInI32 ? 4 : 8; }

static void check_dims(const Context& c, int limbs, int polys) {
    RCK_REQUIRE(c.logn >= 1 && c.logn <= 16, "ring degree must be in 2^1 .. 2^16");
    RCK_REQUIRE(limbs > 0 && limbs <= 65535 && polys > 0 && polys <= 65535, "bad launch batch");
}

#define RCK_LAUNCH_CHECK() RCK_CHECK_HIP(hipGetLastError())

template <int LOGR>
static void fwd_cols_t(const Context& c, int in_kind, Batch<const void> in, Batch<double> out, int limbs, int polys,
                       PrimeMap pm, hipStream_t s, const void* in2, int zsplit) {
    constexpr int S = 256 / ((1 << LOGR) / 16);
    const int n2 = 1 << (c.logn - LOGR);
    dim3 grid(n2 / S, limbs, polys);
    ProfScope ps(s, "ntt_fwd_cols");
    if (in_kind == kInU64)
        k_fwd_cols<LOGR, kInU64><<<grid, 256, 0, s>>>(in, 8, out, c.logn, pm, c.d_primes, in2, zsplit);
    else if (in_kind == kInI32)
        k_fwd_cols<LOGR, kInI32><<<grid, 256, 0, s>>>(in, 4, out, c.logn, pm, c.d_primes, in2, zsplit);
    else
        k_fwd_cols<LOGR, kInF64><<<grid, 256, 0, s>>>(in, 8, out, c.logn, pm, c.d_primes, in2, zsplit);
    RCK_LAUNCH_CHECK();
}

template <int LOGR>
static void fwd_rows_t(const Context& c, Batch<double> d, int limbs, int polys, PrimeMap pm, hipStream_t s) {
    constexpr int S = 256 / ((1 << LOGR) / 16);
    const int n1 = 1 << (c.logn - LOGR);
    ProfScope ps(s, "ntt_fwd_rows");
    k_fwd_rows<LOGR><<<dim3(n1 / S, limbs, polys), 256, 0, s>>>(d, c.logn, pm, c.d_primes);
    RCK_LAUNCH_CHECK();
}

template <int LOGR>
static void inv_rows_t(const Context& c, Batch<double> d, int limbs, int polys, PrimeMap pm, hipStream_t s, int nsum,
                       long long sstride, const double* addn, long long addn_bs, long long addn_cs, int nq,
                       const int* gate) {
    constexpr int S = 256 / ((1 << LOGR) / 16);
    const int n1 = 1 << (c.logn - LOGR);
    ProfScope ps(s, gate ? "ntt_inv_rows_gated" : "ntt_inv_rows");
    k_inv_rows<LOGR><<<dim3(n1 / S, limbs, polys), 256, 0, s>>>(d, c.logn, pm, c.d_primes, nsum, sstride, addn,
                                                                addn_bs, addn_cs, nq, c.d_pmodq, gate);
    RCK_LAUNCH_CHECK();
}

template <int LOGR>
static void inv_cols_t(const Context& c, Batch<const double> in, int out_kind, Batch<void> out, int limbs, int polys,
                       PrimeMap pm, hipStream_t s) {
    constexpr int S = 256 / ((1 << LOGR) / 16);
    const int n2 = 1 << (c.logn - LOGR);
    dim3 grid(n2 / S, limbs, polys);
    ProfScope ps(s, "ntt_inv_cols");
    if (out_kind == kOutU64)
        k_inv_cols<LOGR, kOutU64><<<grid, 256, 0, s>>>(in, out, 8, c.logn, pm, c.d_primes);
    else
        k_inv_cols<LOGR, kOutF64><<<grid, 256, 0, s>>>(in, out, 8, c.logn, pm, c.d_primes);
    RCK_LAUNCH_CHECK();
}

static bool small_ring(const Context& c) { return c.logn <= 11; }

template <int LOGR, int NT>
static void mult_rows_nt(const Context& c, const double* xl, const double* xr, double* d, int limbs, int B,
                         PrimeMap pm, hipStream_t s, bool fold) {
    constexpr int S = NT / ((1 << LOGR) / 16);
    const int n1 = 1 << (c.logn - LOGR);
    const int nrb = n1 / S;
    const long long items = (long long)nrb * limbs * B;
    RCK_REQUIRE(items < (1LL << 30), "tensor launch too large");
    const dim3 grid((unsigned)((items + 7) / 8 * 8));
    if (fold)
        k_mult_rows<LOGR, true, NT><<<grid, NT, 0, s>>>(xl, xr, d, (long long)limbs * c.n, c.logn, pm, c.d_primes, B,
                                                        nrb, (int)items);
    else
        k_mult_rows<LOGR, false, NT><<<grid, NT, 0, s>>>(xl, xr, d, (long long)limbs * c.n, c.logn, pm, c.d_primes,
                                                         B, nrb, (int)items);
}

template <int LOGR>
static void mult_rows_t(const Context& c, const double* xl, const double* xr, double* d, int limbs, int B, PrimeMap pm,
                        hipStream_t s, bool fold) {
    constexpr int S = 256 / ((1 << LOGR) / 16);
    const int n1 = 1 << (c.logn - LOGR);
    ProfScope ps(s, "mult_rows");
    // two 256-thread workgroups fit a CU (registers): below 4 per CU, halve the workgroups instead
    if ((long long)(n1 / S) * limbs * B < 4LL * 256)
        mult_rows_nt<LOGR, 128>(c, xl, xr, d, limbs, B, pm, s, fold);
    else
        mult_rows_nt<LOGR, 256>(c, xl, xr, d, limbs, B, pm, s, fold);
    RCK_LAUNCH_CHECK();
}

void launch_mult_rows(const Context& c, const double* xl, const double* xr, double* d, int limbs, int B, PrimeMap pm,
                      hipStream_t s, bool fold) {
    require_fp64(c);
    check_dims(c, limbs, B);
    RCK_REQUIRE(!small_ring(c), "fused tensor rows need N >= 2^12");
    int n1, n2;
    split_logn(c.logn, &n1, &n2);
    switch (n2) {
        case 6: mult_rows_t<6>(c, xl, xr, d, limbs, B, pm, s, fold); break;
        case 7: mult_rows_t<7>(c, xl, xr, d, limbs, B, pm, s, fold); break;
        case 8: mult_rows_t<8>(c, xl, xr, d, limbs, B, pm, s, fold); break;
        default: RCK_REQUIRE(false, "unsupported ring degree");
    }
}

void launch_fwd_cols(const Context& c, int in_kind, const void* in, long long in_ls, long long in_ps, double* out,
                     long long out_ls, long long out_ps, int limbs, int polys, PrimeMap pm, hipStream_t s,
                     const void* in2, int zsplit) {
    require_fp64(c);
    check_dims(c, limbs, polys);
    RCK_REQUIRE(in2 == nullptr || (zsplit > 0 && zsplit < polys), "bad second operand split");
    RCK_REQUIRE(!small_ring(c), "two-pass NTT needs N >= 2^12");
    Batch<const void> bi{in, in_ls, in_ps};
    Batch<double> bo{out, out_ls, out_ps};
    int n1, n2;
    split_logn(c.logn, &n1, &n2);
    switch (n1) {
        case 6: fwd_cols_t<6>(c, in_kind, bi, bo, limbs, polys, pm, s, in2, zsplit); break;
        case 7: fwd_cols_t<7>(c, in_kind, bi, bo, limbs, polys, pm, s, in2, zsplit); break;
        case 8: fwd_cols_t<8>(c, in_kind, bi, bo, limbs, polys, pm, s, in2, zsplit); break;
        default: RCK_REQUIRE(false, "unsupported ring degree");
    }
}

void launch_fwd_rows(const Context& c, double* data, long long ls, long long ps, int limbs, int polys, PrimeMap pm,
                     hipStream_t s) {
    require_fp64(c);
    check_dims(c, limbs, polys);
    Batch<double> b{data, ls, ps};
    int n1, n2;
    split_logn(c.logn, &n1, &n2);
    switch (n2) {
        case 6: fwd_rows_t<6>(c, b, limbs, polys, pm, s); break;
        case 7: fwd_rows_t<7>(c, b, limbs, polys, pm, s); break;
        case 8: fwd_rows_t<8>(c, b, limbs, polys, pm, s); break;
        default: RCK_REQUIRE(false, "unsupported ring degree");
    }
}

void launch_inv_rows(const Context& c, double* data, long long ls, long long ps, int limbs, int polys, PrimeMap pm,
                     hipStream_t s, int nsum, long long sstride, const double* addn, long long addn_bs,
                     long long addn_cs, int nq, const int* gate) {
    require_fp64(c);
    check_dims(c, limbs, polys);
    RCK_REQUIRE(nsum >= 1 && nsum <= 8, "at most 8 partial sums");
    RCK_REQUIRE(addn == nullptr || (polys % 2 == 0 && nq <= limbs && nq <= (int)c.L), "bad folded addend");
    Batch<double> b{data, ls, ps};
    int n1, n2;
    split_logn(c.logn, &n1, &n2);
    switch (n2) {
        case 6: inv_rows_t<6>(c, b, limbs, polys, pm, s, nsum, sstride, addn, addn_bs, addn_cs, nq, gate); break;
        case 7: inv_rows_t<7>(c, b, limbs, polys, pm, s, nsum, sstride, addn, addn_bs, addn_cs, nq, gate); break;
        case 8: inv_rows_t<8>(c, b, limbs, polys, pm, s, nsum, sstride, addn, addn_bs, addn_cs, nq, gate); break;
        default: RCK_REQUIRE(false, "unsupported ring degree");
    }
}

void launch_inv_cols(const Context& c, const double* in, long long in_ls, long long in_ps, int out_kind, void* out,
                     long long out_ls, long long out_ps, int limbs, int polys, PrimeMap pm, hipStream_t s) {
    require_fp64(c);
    check_dims(c, limbs, polys);
    Batch<const double> bi{in, in_ls, in_ps};
    Batch<void> bo{out, out_ls, out_ps};
    int n1, n2;
    split_logn(c.logn, &n1, &n2);
    switch (n1) {
        case 6: inv_cols_t<6>(c, bi, out_kind, bo, limbs, polys, pm, s); break;
        case 7: inv_cols_t<7>(c, bi, out_kind, bo, limbs, polys, pm, s); break;
        case 8: inv_cols_t<8>(c, bi, out_kind, bo, limbs, polys, pm, s); break;
        default: RCK_REQUIRE(false, "unsupported ring degree");
    }
}

void launch_fwd(const Context& c, int in_kind, const void* in, long long in_ls, long long in_ps, double* out,
                long long out_ls, long long out_ps, int limbs, int polys, PrimeMap pm, hipStream_t s) {
    require_fp64(c);
    check_dims(c, limbs, polys);
    if (small_ring(c)) {
        Batch<const void> bi{in, in_ls, in_ps};
        Batch<double> bo{out, out_ls, out_ps};
        dim3 grid(1, limbs, polys);
        ProfScope ps(s, "ntt_small_fwd");
        if (in_kind == kInU64)
            k_small_fwd<kInU64><<<grid, 256, 0, s>>>(bi, 8, bo, c.logn, pm, c.d_primes);
        else if (in_kind == kInI32)
            k_small_fwd<kInI32><<<grid, 256, 0, s>>>(bi, 4, bo, c.logn, pm, c.d_primes);
        else
            k_small_fwd<kInF64><<<grid, 256, 0, s>>>(bi, 8, bo, c.logn, pm, c.d_primes);
        RCK_LAUNCH_CHECK();
        return;
    }
    launch_fwd_cols(c, in_kind, in, in_ls, in_ps, out, out_ls, out_ps, limbs, polys, pm, s);
    launch_fwd_rows(c, out, out_ls, out_ps, limbs, polys, pm, s);
}

void launch_inv(const Context& c, double* in, long long in_ls, long long in_ps, int out_kind, void* out,
                long long out_ls, long long out_ps, int limbs, int polys, PrimeMap pm, hipStream_t s, int nsum,
                long long sstride) {
    require_fp64(c);
    check_dims(c, limbs, polys);
    if (small_ring(c)) {
        RCK_REQUIRE(nsum == 1, "partial sums need the two-pass NTT");
        Batch<const double> bi{in, in_ls, in_ps};
        Batch<void> bo{out, out_ls, out_ps};
        dim3 grid(1, limbs, polys);
        ProfScope ps(s, "ntt_small_inv");
        if (out_kind == kOutU64)
            k_small_inv<kOutU64><<<grid, 256, 0, s>>>(bi, bo, 8, c.logn, pm, c.d_primes);
        else
            k_small_inv<kOutF64><<<grid, 256, 0, s>>>(bi, bo, 8, c.logn, pm, c.d_primes);
        RCK_LAUNCH_CHECK();
        return;
    }
    launch_inv_rows(c, in, in_ls, in_ps, limbs, polys, pm, s, nsum, sstride);
    launch_inv_cols(c, in, in_ls, in_ps, out_kind, out, out_ls, out_ps, limbs, polys, pm, s);
}

template <int N1B, int N2B>
static void cyclic_t(const Context& c, bool inverse, uint64_t* data, double* t, int limbs, PrimeMap pm,
                     hipStream_t s) {
    constexpr int Sc = 256 / ((1 << N1B) / 16), Sr = 256 / ((1 << N2B) / 16);
    const dim3 gc((1 << N2B) / Sc, limbs, 1), gr((1 << N1B) / Sr, limbs, 1);
    Batch<const void> bi{data, c.n, 0};
    Batch<double> bt{t, c.n, 0};
    if (!inverse) {
        {
            ProfScope ps(s, "ntt_cyc_fwd_cols");
            k_fwd_cols<N1B, kInU64, true><<<gc, 256, 0, s>>>(bi, 8, bt, c.logn, pm, c.d_primes, nullptr, 0);
            RCK_LAUNCH_CHECK();
        }
        ProfScope ps(s, "ntt_cyc_fwd_rows");
        k_cyc_fwd_rows<N2B><<<gr, 256, 0, s>>>(t, data, c.logn, pm, c.d_primes);
        RCK_LAUNCH_CHECK();
    } else {
        {
            ProfScope ps(s, "ntt_cyc_inv_rows");
            k_cyc_inv_rows<N2B><<<gr, 256, 0, s>>>(data, t, c.logn, pm, c.d_primes);
            RCK_LAUNCH_CHECK();
        }
        ProfScope ps(s, "ntt_cyc_inv_cols");
        k_inv_cols<N1B, kOutU64, true><<<gc, 256, 0, s>>>({t, c.n, 0}, {data, c.n, 0}, 8, c.logn, pm, c.d_primes);
        RCK_LAUNCH_CHECK();
    }
}

void launch_ntt_cyclic(const Context& c, bool inverse, uint64_t* data, double* t, int limbs, PrimeMap pm,
                       hipStream_t s) {
    require_fp64(c);
    check_dims(c, limbs, 1);
    if (small_ring(c)) {
        ProfScope ps(s, inverse ? "ntt_small_cyc_inv" : "ntt_small_cyc_fwd");
        if (inverse)
            k_small_cyc<true><<<dim3(1, limbs, 1), 256, 0, s>>>(data, c.logn, pm, c.d_primes);
        else
            k_small_cyc<false><<<dim3(1, limbs, 1), 256, 0, s>>>(data, c.logn, pm, c.d_primes);
        RCK_LAUNCH_CHECK();
        return;
    }
    RCK_REQUIRE(t != nullptr, "cyclic NTT needs a workspace");
    switch (c.logn) {
        case 12: cyclic_t<6, 6>(c, inverse, data, t, limbs, pm, s); break;
        case 13: cyclic_t<6, 7>(c, inverse, data, t, limbs, pm, s); break;
        case 14: cyclic_t<7, 7>(c, inverse, data, t, limbs, pm, s); break;
        case 15: cyclic_t<7, 8>(c, inverse, data, t, limbs, pm, s); break;
        case 16: cyclic_t<8, 8>(c, inverse, data, t, limbs, pm, s); break;
        default: RCK_REQUIRE(false, "unsupported ring degree");
    }
}

}  // namespace rck
