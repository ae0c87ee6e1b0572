// Elementwise kernels: coefficient-domain add/sub/automorphism/rescale/multiply-add on
// canonical u64 limbs and NTT-domain pointwise products on doubles.  All are HBM-bound
// streaming kernels: one thread per
// coefficient, grid.x over coefficients, grid.y over limbs, grid.z over polynomials.
#include "ops.hpp"
#include "wide.hpp"

namespace rck {

__device__ __forceinline__ long long gidx() { return (long long)blockIdx.x * blockDim.x + threadIdx.x; }

// RnsPolynomial::add / sub (src/rns/polynomial.rs:99-119 over poly/mod.rs:38-59)
__global__ void k_addsub(Batch<const uint64_t> a, Batch<const uint64_t> b, Batch<uint64_t> out, int n, int sub,
                         PrimeMap pm, const DevPrime* __restrict__ primes) {
    const long long i = gidx();
    if (i >= n) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const uint64_t q = (uint64_t)primes[pm(y)].q;
    const uint64_t x = a.at(z, y)[i], v = b.at(z, y)[i];
    uint64_t r;
    if (sub)
        r = x >= v ? x - v : q - (v - x);
    else {
        r = x + v;
        r = r >= q ? r - q : r;
    }
    out.at(z, y)[i] = r;
}

// apply_galois_automorphism (src/poly/mod.rs:110-130) as a gather: out[j] = +-in[j g^{-1} mod 2N]
__global__ void k_galois(Batch<const uint64_t> in, Batch<uint64_t> out, int n, unsigned long long ginv, PrimeMap pm,
                         const DevPrime* __restrict__ primes) {
    const long long j = gidx();
    if (j >= n) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const uint64_t q = (uint64_t)primes[pm(y)].q;
    const unsigned long long two_n = 2ull * n;
    const unsigned long long k = ((unsigned long long)j * ginv) & (two_n - 1);
    uint64_t r;
    if (k < (unsigned long long)n) {
        r = in.at(z, y)[k];
    } else {
        const uint64_t x = in.at(z, y)[k - n];
        r = x == 0 ? 0 : q - x;
    }
    out.at(z, y)[j] = r;
}

// rescale_poly_rns (src/rns/ops.rs:135-186): (x_i - centred(x_last)) * q_last^{-1} mod q_i
__global__ void k_rescale(Batch<const uint64_t> in, Batch<uint64_t> out, int n, int last, int T,
                          const DevPrime* __restrict__ primes, const double* __restrict__ invq) {
    const long long c = gidx();
    if (c >= n) return;
    const int i = blockIdx.y, z = blockIdx.z;
    const DevPrime Pl = primes[last], Pi = primes[i];
    const uint64_t xl = in.at(z, last)[c];
    const uint64_t ql = (uint64_t)Pl.q;
    const double cen = xl <= (ql >> 1) ? (double)xl : (double)xl - Pl.q;
    const double cm = reduce(cen, Pi.q, Pi.qinv);
    const double num = (double)in.at(z, i)[c] - cm;
    const double r = mulmod(num, invq[(long long)i * T + last], Pi.q, Pi.qinv);
    out.at(z, i)[c] = f64_to_u64(canon(r, Pi.q, Pi.qinv));
}

// pointwise modular multiply-add on canonical limbs: out = a * b + c (c optional)
__global__ void k_mul_add_u64(Batch<const uint64_t> a, Batch<const uint64_t> b, Batch<const uint64_t> cc, int has_c,
                              Batch<uint64_t> out, int n, PrimeMap pm, const DevPrime* __restrict__ primes) {
    const long long i = gidx();
    if (i >= n) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const DevPrime P = primes[pm(y)];
    double r = mulmod(load_centred(a.at(z, y)[i], P.q), load_centred(b.at(z, y)[i], P.q), P.q, P.qinv);
    if (has_c) r += load_centred(cc.at(z, y)[i], P.q);
    out.at(z, y)[i] = f64_to_u64(canon(r, P.q, P.qinv));
}

// NTT-domain pointwise product (reduced doubles in and out): out = a * b (+ c)
__global__ void k_mul_f64(Batch<const double> a, Batch<const double> b, Batch<double> out, int n, PrimeMap pm,
                          const DevPrime* __restrict__ primes) {
    const long long i = gidx();
    if (i >= n) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const DevPrime P = primes[pm(y)];
    out.at(z, y)[i] = reduce(mulmod(a.at(z, y)[i], b.at(z, y)[i], P.q, P.qinv), P.q, P.qinv);
}

// ct x ct tensor in the NTT domain (multiply_ciphertexts_rns, src/rns/ops.rs:84-100):
// x = [a0 a1 b0 b1] (4 polys, reduced doubles) -> d = [a0 b0, a0 b1 + a1 b0, a1 b1]
__global__ void k_tensor(const double* __restrict__ x, double* __restrict__ d, long long poly, int n, PrimeMap pm,
                         const DevPrime* __restrict__ primes) {
    const long long i = gidx();
    if (i >= n) return;
    const int y = blockIdx.y;
    const DevPrime P = primes[pm(y)];
    const long long o = (long long)y * n + i;
    const double a0 = x[o], a1 = x[poly + o], b0 = x[2 * poly + o], b1 = x[3 * poly + o];
    d[o] = reduce(mulmod(a0, b0, P.q, P.qinv), P.q, P.qinv);
    d[poly + o] = reduce(mulmod(a0, b1, P.q, P.qinv) + mulmod(a1, b0, P.q, P.qinv), P.q, P.qinv);
    d[2 * poly + o] = reduce(mulmod(a1, b1, P.q, P.qinv), P.q, P.qinv);
}

// plain streaming copy (diagnostics: calibrates hardware byte counters against a known volume)
__global__ void k_copy(const double4* __restrict__ src, double4* __restrict__ dst, long long n4) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

void launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
    RCK_REQUIRE(bytes % 32 == 0, "copy size must be a multiple of 32 bytes");
    ProfScope ps(s, "copy");
    k_copy<<<dim3(4096), 256, 0, s>>>(static_cast<const double4*>(src), static_cast<double4*>(dst),
                                       (long long)(bytes / 32));
    RCK_CHECK_HIP(hipGetLastError());
}

static dim3 grid_for(int n, int limbs, int polys) { return dim3((n + 255) / 256, limbs, polys); }

#define CHK() RCK_CHECK_HIP(hipGetLastError())

void launch_addsub(const Context& c, Batch<const uint64_t> a, Batch<const uint64_t> b, Batch<uint64_t> out, int limbs,
                   int polys, int sub, PrimeMap pm, hipStream_t s) {
    if (c.wide) return wide_addsub(c, a, b, out, limbs, polys, sub, pm, s);
    ProfScope ps(s, "addsub");
    k_addsub<<<grid_for(c.n, limbs, polys), 256, 0, s>>>(a, b, out, c.n, sub, pm, c.d_primes);
    CHK();
}

uint64_t galois_inverse(const Context& c, uint64_t galois) {
    uint64_t ginv;
    RCK_REQUIRE(galois % 2 == 1, "galois automorphisms require an odd exponent");
    RCK_REQUIRE(galois < 2ull * c.n, "galois element must be reduced modulo 2N");
    RCK_REQUIRE(h_invmod(galois, 2ull * c.n, &ginv), "galois element must be invertible modulo 2N");
    return ginv;
}

void launch_galois(const Context& c, Batch<const uint64_t> in, Batch<uint64_t> out, int limbs, int polys,
                   uint64_t galois, PrimeMap pm, hipStream_t s) {
    const uint64_t ginv = galois_inverse(c, galois);
    if (c.wide) return wide_galois(c, in, out, limbs, polys, ginv, pm, s);
    ProfScope ps(s, "galois");
    k_galois<<<grid_for(c.n, limbs, polys), 256, 0, s>>>(in, out, c.n, ginv, pm, c.d_primes);
    CHK();
}

void launch_rescale(const Context& c, Batch<const uint64_t> in, Batch<uint64_t> out, int limbs_in, int polys,
                    hipStream_t s) {
    RCK_REQUIRE(limbs_in > 1, "rescale requires at least two levels");
    if (c.wide) return wide_rescale(c, in, out, limbs_in, polys, s);
    ProfScope ps(s, "rescale");
    k_rescale<<<grid_for(c.n, limbs_in - 1, polys), 256, 0, s>>>(in, out, c.n, limbs_in - 1, c.L + c.K, c.d_primes,
                                                                   c.d_invq);
    CHK();
}

void launch_mul_add_u64(const Context& c, Batch<const uint64_t> a, Batch<const uint64_t> b, Batch<const uint64_t> cc,
                        bool has_c, Batch<uint64_t> out, int limbs, int polys, PrimeMap pm, hipStream_t s) {
    if (c.wide) return wide_mul_add(c, a, b, cc, has_c, out, limbs, polys, pm, s);
    ProfScope ps(s, "mul_add_u64");
    k_mul_add_u64<<<grid_for(c.n, limbs, polys), 256, 0, s>>>(a, b, cc, has_c ? 1 : 0, out, c.n, pm, c.d_primes);
    CHK();
}

void launch_mul_f64(const Context& c, Batch<const double> a, Batch<const double> b, Batch<double> out, int limbs,
                    int polys, PrimeMap pm, hipStream_t s) {
    require_fp64(c);
    ProfScope ps(s, "mul_f64");
    k_mul_f64<<<grid_for(c.n, limbs, polys), 256, 0, s>>>(a, b, out, c.n, pm, c.d_primes);
    CHK();
}

void launch_tensor(const Context& c, const double* x, double* d, int limbs, PrimeMap pm, hipStream_t s) {
    require_fp64(c);
    ProfScope ps(s, "tensor");
    k_tensor<<<grid_for(c.n, limbs, 1), 256, 0, s>>>(x, d, (long long)limbs * c.n, c.n, pm, c.d_primes);
    CHK();
}

}  // namespace rck
