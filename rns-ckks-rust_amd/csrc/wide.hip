// 64-bit integer path: NTT, elementwise operations, ct x ct product, rescale and the hybrid /
// Q-only keyswitch for contexts whose moduli exceed the FP64-exact range (up to the crate's full
// u64 range).  Same operations and results as the FP64 kernels; the arithmetic is Montgomery
// multiplication on canonical u64 residues (wide.hpp).
//
// NTT: the merged-psi Cooley-Tukey forward transform (natural order in, bit-reversed out) and the
// Gentleman-Sande inverse.  Stages whose butterfly distance is at least the LDS tile (4096 elements,
// 32 KiB) run up to four at a time in registers (one streaming pass over the limb per four stages);
// the remaining stages of each contiguous tile run in LDS in one kernel.  Negacyclic products, keys and accumulators use this
// transform; the crate's cyclic NTT (src/math/ntt.rs:78-89) is a psi^{-j} twist, this transform and
// a bit-reversal (and the reverse for the inverse).
#include <algorithm>
#include <cstring>

#include "pipeline.hpp"
#include "wide.hpp"

namespace rck {

namespace {

constexpr int kTileLog = 12;  // 4096 elements per LDS tile

__device__ __forceinline__ long long w_gidx() { return (long long)blockIdx.x * blockDim.x + threadIdx.x; }

dim3 w_grid(long long n, int y, int z) { return dim3((unsigned)((n + 255) / 256), (unsigned)y, (unsigned)z); }

void w_check() { RCK_CHECK_HIP(hipGetLastError()); }

uint64_t h_neg_inv64(uint64_t q) {
    uint64_t x = q;  // q q = 1 mod 8; each Newton step doubles the correct low bits
    for (int i = 0; i < 6; i++) x *= 2 - q * x;
    return 0 - x;
}

uint64_t h_r_mod(uint64_t q) { return (uint64_t)(((unsigned __int128)1 << 64) % q); }

uint32_t h_bitrev(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

template <typename T>
T* w_upload(const std::vector<T>& v) {
    T* d = nullptr;
    RCK_CHECK_HIP(hipMalloc(&d, std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) RCK_CHECK_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------------------
void wide_tables_create(Context& c, const std::vector<uint64_t>& all, uint32_t T) {
    const uint32_t n = c.n;
    std::vector<uint64_t> tables((size_t)T * 4 * n);
    std::vector<uint64_t> pw(n), pwi(n);
    for (uint32_t i = 0; i < T; i++) {
        const HostPrime& h = c.hp[i];
        const uint64_t q = h.q, R = h_r_mod(q);
        uint64_t a = 1, b = 1;
        for (uint32_t k = 0; k < n; k++) {
            pw[k] = a;
            pwi[k] = b;
            a = h_mulmod(a, h.psi, q);
            b = h_mulmod(b, h.psi_inv, q);
        }
        uint64_t* fw = &tables[(size_t)i * 4 * n];
        uint64_t *iv = fw + n, *tw = iv + n, *twi = tw + n;
        for (uint32_t k = 0; k < n; k++) {
            const uint32_t r = h_bitrev(k, (int)c.logn);
            fw[k] = h_mulmod(pw[r], R, q);
            iv[k] = h_mulmod(pwi[r], R, q);
            tw[k] = h_mulmod(pwi[k], R, q);
            twi[k] = h_mulmod(pw[k], R, q);
        }
    }
    c.d_wtables = w_upload(tables);
    std::vector<WidePrime> wp(T);
    for (uint32_t i = 0; i < T; i++) {
        const uint64_t q = c.hp[i].q, R = h_r_mod(q);
        const uint64_t* base = c.d_wtables + (size_t)i * 4 * n;
        wp[i] = WidePrime{q, h_neg_inv64(q), h_mulmod(R, R, q), h_mulmod(c.hp[i].n_inv, R, q), base, base + n,
                          base + 2 * n, base + 3 * n};
    }
    c.d_wp = w_upload(wp);
    // q_j^{-1} R mod q_i
    std::vector<uint64_t> invq((size_t)T * T, 0);
    for (uint32_t i = 0; i < T; i++)
        for (uint32_t j = 0; j < T; j++) {
            if (i == j) continue;
            uint64_t v;
            RCK_REQUIRE(h_invmod(all[j] % all[i], all[i], &v), "RNS moduli must be pairwise coprime");
            invq[(size_t)i * T + j] = h_mulmod(v, h_r_mod(all[i]), all[i]);
        }
    c.d_winvq = w_upload(invq);
    const uint32_t L = c.L, K = c.K;
    if (K > 0) {
        std::vector<uint64_t> pmodq(L), pinvq(L), paux((size_t)K * L);
        for (uint32_t i = 0; i < L; i++) {
            const uint64_t qi = all[i], R = h_r_mod(qi);
            uint64_t m = 1 % qi;
            for (uint32_t k = 0; k < K; k++) {
                paux[(size_t)k * L + i] = h_mulmod(all[L + k] % qi, R, qi);
                m = h_mulmod(m, all[L + k] % qi, qi);
            }
            uint64_t inv;
            RCK_REQUIRE(h_invmod(m, qi, &inv), "P must be invertible modulo each q_i");
            pmodq[i] = m;
            pinvq[i] = h_mulmod(inv, R, qi);
        }
        // floor(P / 2) as kMaxP 64-bit words
        std::vector<uint64_t> pword(kMaxP + 1, 0);
        pword[0] = 1;
        for (uint32_t k = 0; k < K; k++) {
            unsigned __int128 carry = 0;
            for (int w = 0; w <= kMaxP; w++) {
                const unsigned __int128 t = (unsigned __int128)pword[w] * all[L + k] + carry;
                pword[w] = (uint64_t)t;
                carry = t >> 64;
            }
        }
        std::vector<uint64_t> half(kMaxP, 0);
        for (int w = 0; w < kMaxP; w++) half[w] = (pword[w] >> 1) | (w + 1 <= kMaxP ? pword[w + 1] << 63 : 0);
        c.d_wpmodq = w_upload(pmodq);
        c.d_wpinvq = w_upload(pinvq);
        c.d_wpaux = w_upload(paux);
        c.d_wphalf = w_upload(half);
    }
}

// ---------------------------------------------------------------------------------------
// NTT
// ---------------------------------------------------------------------------------------
// S consecutive radix-2 stages with butterfly distances from t = 2^lt (>= the tile) down to
// t / 2^(S-1) (forward CT) or up from t / 2^(S-1) to t (inverse GS), in registers: a thread holds the
// 2^S elements j = blk 2t + r + m delta (delta = 2t / 2^S, m < 2^S) that these stages combine.  The
// stage of distance d uses table entry N / 2d + (j >> log2 2d) for element j.
template <bool INV, int S>
__global__ __launch_bounds__(256) void k_w_ntt_pass(Batch<uint64_t> data, int logn, int lt, PrimeMap pm,
                                                    const WidePrime* __restrict__ wp, int scale) {
    constexpr int M = 1 << S;
    const long long gid = w_gidx();
    if (gid >= (1LL << (logn - S))) return;
    const WidePrime P = wp[pm(blockIdx.y)];
    uint64_t* a = data.at(blockIdx.z, blockIdx.y);
    const int ld = lt + 1 - S;  // log2 delta
    const long long blk = gid >> ld, r = gid & ((1LL << ld) - 1);
    const long long j0 = (blk << (lt + 1)) + r;
    uint64_t v[M];
#pragma unroll
    for (int m = 0; m < M; m++) v[m] = a[j0 + ((long long)m << ld)];
    if constexpr (!INV) {
#pragma unroll
        for (int k = 0; k < S; k++) {  // distance t >> k = 2^(ld + S - 1 - k) elements, 2^(S-1-k) registers
            const int h = M >> (k + 1);
            const long long base = (1LL << (logn - lt - 1 + k)) + (blk << k);
#pragma unroll
            for (int m = 0; m < M; m++) {
                if (m & h) continue;
                const uint64_t W = w_mul(v[m + h], P.fw[base + (m >> (S - k))], P.q, P.qn);
                const uint64_t U = v[m];
                v[m] = w_add(U, W, P.q);
                v[m + h] = w_sub(U, W, P.q);
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < S; k++) {  // distance delta 2^k
            const int h = 1 << k;
            const long long base = (1LL << (logn - ld - k - 1)) + (blk << (S - 1 - k));
#pragma unroll
            for (int m = 0; m < M; m++) {
                if (m & h) continue;
                const uint64_t U = v[m], V = v[m + h];
                v[m] = w_add(U, V, P.q);
                v[m + h] = w_mul(w_sub(U, V, P.q), P.iv[base + (m >> (k + 1))], P.q, P.qn);
            }
        }
        if (scale) {
#pragma unroll
            for (int m = 0; m < M; m++) v[m] = w_mul(v[m], P.ninv, P.q, P.qn);
        }
    }
#pragma unroll
    for (int m = 0; m < M; m++) a[j0 + ((long long)m << ld)] = v[m];
}

// The stages of butterfly distance below the tile (2^tb elements, contiguous) in LDS: forward
// distances 2^(tb-1) .. 1, inverse 1 .. 2^(tb-1) then the optional N^{-1} scaling.
template <bool INV>
__global__ __launch_bounds__(256) void k_w_ntt_tile(Batch<uint64_t> data, int logn, int tb, PrimeMap pm,
                                                    const WidePrime* __restrict__ wp, int scale) {
    __shared__ uint64_t a[1 << kTileLog];
    const int tile = 1 << tb, half = tile >> 1;
    const WidePrime P = wp[pm(blockIdx.y)];
    const long long g0 = (long long)blockIdx.x * tile;
    uint64_t* src = data.at(blockIdx.z, blockIdx.y) + g0;
    for (int i = threadIdx.x; i < tile; i += blockDim.x) a[i] = src[i];
    __syncthreads();
    for (int s = 0; s < tb; s++) {
        const int lt = INV ? s : tb - 1 - s;
        const int t = 1 << lt;
        const long long m = 1LL << (logn - 1 - lt);
        for (int b = threadIdx.x; b < half; b += blockDim.x) {
            const int il = b >> lt, j = (il << (lt + 1)) + (b & (t - 1));
            const long long i = (g0 >> (lt + 1)) + il;
            const uint64_t U = a[j], V = a[j + t];
            if constexpr (!INV) {
                const uint64_t W = w_mul(V, P.fw[m + i], P.q, P.qn);
                a[j] = w_add(U, W, P.q);
                a[j + t] = w_sub(U, W, P.q);
            } else {
                a[j] = w_add(U, V, P.q);
                a[j + t] = w_mul(w_sub(U, V, P.q), P.iv[m + i], P.q, P.qn);
            }
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < tile; i += blockDim.x) src[i] = scale ? w_mul(a[i], P.ninv, P.q, P.qn) : a[i];
}

static void w_pass(bool inv, int st, Batch<uint64_t> d, int logn, int lt, int limbs, int polys, PrimeMap pm,
                   const WidePrime* wp, int scale, hipStream_t s) {
    const dim3 g = w_grid(1LL << (logn - st), limbs, polys);
#define RCK_WP(I, S_) k_w_ntt_pass<I, S_><<<g, 256, 0, s>>>(d, logn, lt, pm, wp, scale)
    switch (st * 2 + (inv ? 1 : 0)) {
        case 2: RCK_WP(false, 1); break;
        case 3: RCK_WP(true, 1); break;
        case 4: RCK_WP(false, 2); break;
        case 5: RCK_WP(true, 2); break;
        case 6: RCK_WP(false, 3); break;
        case 7: RCK_WP(true, 3); break;
        case 8: RCK_WP(false, 4); break;
        default: RCK_WP(true, 4); break;
    }
#undef RCK_WP
}

void wide_ntt(const Context& c, bool inverse, uint64_t* data, long long ls, long long ps, int limbs, int polys,
              PrimeMap pm, hipStream_t s, bool scale) {
    if (limbs <= 0 || polys <= 0) return;
    const int logn = (int)c.logn, tb = std::min(logn, kTileLog);
    const long long n = c.n;
    Batch<uint64_t> d{data, ls, ps};
    const dim3 gt((unsigned)(n >> tb), (unsigned)limbs, (unsigned)polys);
    ProfScope prof(s, inverse ? "wide_ntt_inv" : "wide_ntt_fwd");
    if (!inverse) {
        // the stages above the tile, up to four per pass (largest distances first)
        for (int lt = logn - 1; lt >= tb;) {
            const int st = std::min(4, lt - tb + 1);
            w_pass(false, st, d, logn, lt, limbs, polys, pm, c.d_wp, 0, s);
            lt -= st;
        }
        k_w_ntt_tile<false><<<gt, 256, 0, s>>>(d, logn, tb, pm, c.d_wp, 0);
    } else {
        k_w_ntt_tile<true><<<gt, 256, 0, s>>>(d, logn, tb, pm, c.d_wp, scale && tb == logn ? 1 : 0);
        for (int lo = tb; lo < logn;) {  // distances 2^lo .. 2^(lo + st - 1)
            const int st = std::min(4, logn - lo);
            w_pass(true, st, d, logn, lo + st - 1, limbs, polys, pm, c.d_wp, scale && lo + st == logn ? 1 : 0, s);
            lo += st;
        }
    }
    w_check();
}

// out[j] = in[j] * table[j] (table: psi^{-j} R (tw) or psi^{j} R (twi))
__global__ void k_w_twist(Batch<const uint64_t> in, Batch<uint64_t> out, int n, int inverse_table, PrimeMap pm,
                          const WidePrime* __restrict__ wp) {
    const long long j = w_gidx();
    if (j >= n) return;
    const WidePrime P = wp[pm(blockIdx.y)];
    const uint64_t* tbl = inverse_table ? P.twi : P.tw;
    out.at(blockIdx.z, blockIdx.y)[j] = w_mul(in.at(blockIdx.z, blockIdx.y)[j], tbl[j], P.q, P.qn);
}

// out[k] = in[bitrev(k)]
__global__ void k_w_brv(Batch<const uint64_t> in, Batch<uint64_t> out, int n, int logn) {
    const long long k = w_gidx();
    if (k >= n) return;
    const unsigned r = __brev((unsigned)k) >> (32 - logn);
    out.at(blockIdx.z, blockIdx.y)[k] = in.at(blockIdx.z, blockIdx.y)[r];
}

void wide_ntt_cyclic(Context& c, bool inverse, uint64_t* data, int limbs, PrimeMap pm, hipStream_t s) {
    const long long n = c.n;
    uint64_t* t = static_cast<uint64_t*>(c.workspace(8, (size_t)limbs * n * 8));
    const dim3 g = w_grid(n, limbs, 1);
    ProfScope prof(s, inverse ? "wide_cyclic_inv" : "wide_cyclic_fwd");
    if (!inverse) {
        // X_k = sum_j x_j w^{jk}: the merged transform of x_j psi^{-j} holds X_k at slot bitrev(k)
        k_w_twist<<<g, 256, 0, s>>>({data, n, 0}, {t, n, 0}, (int)n, 0, pm, c.d_wp);
        wide_ntt(c, false, t, n, 0, limbs, 1, pm, s);
        k_w_brv<<<g, 256, 0, s>>>({t, n, 0}, {data, n, 0}, (int)n, (int)c.logn);
    } else {
        k_w_brv<<<g, 256, 0, s>>>({data, n, 0}, {t, n, 0}, (int)n, (int)c.logn);
        wide_ntt(c, true, t, n, 0, limbs, 1, pm, s);
        k_w_twist<<<g, 256, 0, s>>>({t, n, 0}, {data, n, 0}, (int)n, 1, pm, c.d_wp);
    }
    w_check();
}

// ---------------------------------------------------------------------------------------
// elementwise
// ---------------------------------------------------------------------------------------
__global__ void k_w_addsub(Batch<const uint64_t> a, Batch<const uint64_t> b, Batch<uint64_t> out, int n, int sub,
                           PrimeMap pm, const WidePrime* __restrict__ wp) {
    const long long i = w_gidx();
    if (i >= n) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const uint64_t q = wp[pm(y)].q;
    const uint64_t x = a.at(z, y)[i], v = b.at(z, y)[i];
    out.at(z, y)[i] = sub ? w_sub(x, v, q) : w_add(x, v, q);
}

void wide_addsub(const Context& c, Batch<const uint64_t> a, Batch<const uint64_t> b, Batch<uint64_t> out, int limbs,
                 int polys, int sub, PrimeMap pm, hipStream_t s) {
    ProfScope prof(s, "wide_addsub");
    k_w_addsub<<<w_grid(c.n, limbs, polys), 256, 0, s>>>(a, b, out, (int)c.n, sub, pm, c.d_wp);
    w_check();
}

__global__ void k_w_galois(Batch<const uint64_t> in, Batch<uint64_t> out, int n, unsigned long long ginv, PrimeMap pm,
                           const WidePrime* __restrict__ wp) {
    const long long j = w_gidx();
    if (j >= n) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const uint64_t q = wp[pm(y)].q;
    const unsigned long long k = ((unsigned long long)j * ginv) & (2ull * n - 1);
    uint64_t r;
    if (k < (unsigned long long)n) {
        r = in.at(z, y)[k];
    } else {
        const uint64_t x = in.at(z, y)[k - n];
        r = x == 0 ? 0 : q - x;
    }
    out.at(z, y)[j] = r;
}

void wide_galois(const Context& c, Batch<const uint64_t> in, Batch<uint64_t> out, int limbs, int polys,
                 unsigned long long ginv, PrimeMap pm, hipStream_t s) {
    ProfScope prof(s, "wide_galois");
    k_w_galois<<<w_grid(c.n, limbs, polys), 256, 0, s>>>(in, out, (int)c.n, ginv, pm, c.d_wp);
    w_check();
}

// rescale_poly_rns (src/rns/ops.rs:135-186): (x_i - centred(x_last)) q_last^{-1} mod q_i
__global__ void k_w_rescale(Batch<const uint64_t> in, Batch<uint64_t> out, int n, int last, int T,
                            const WidePrime* __restrict__ wp, const uint64_t* __restrict__ invq) {
    const long long c = w_gidx();
    if (c >= n) return;
    const int i = blockIdx.y, z = blockIdx.z;
    const WidePrime Pi = wp[i];
    const uint64_t ql = wp[last].q, xl = in.at(z, last)[c];
    const uint64_t cm = xl <= (ql >> 1) ? w_red(xl, Pi) : w_sub(0, w_red(ql - xl, Pi), Pi.q);
    const uint64_t num = w_sub(in.at(z, i)[c], cm, Pi.q);
    out.at(z, i)[c] = w_mul(num, invq[(long long)i * T + last], Pi.q, Pi.qn);
}

void wide_rescale(const Context& c, Batch<const uint64_t> in, Batch<uint64_t> out, int limbs_in, int polys,
                  hipStream_t s) {
    RCK_REQUIRE(limbs_in > 1, "rescale requires at least two levels");
    ProfScope prof(s, "wide_rescale");
    k_w_rescale<<<w_grid(c.n, limbs_in - 1, polys), 256, 0, s>>>(in, out, (int)c.n, limbs_in - 1, (int)(c.L + c.K),
                                                                   c.d_wp, c.d_winvq);
    w_check();
}

// out = a b (+ c), plain residues: mont(mont(a, b), R^2) = a b
__global__ void k_w_mul_add(Batch<const uint64_t> a, Batch<const uint64_t> b, Batch<const uint64_t> cc, int has_c,
                            Batch<uint64_t> out, int n, PrimeMap pm, const WidePrime* __restrict__ wp) {
    const long long i = w_gidx();
    if (i >= n) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const WidePrime P = wp[pm(y)];
    uint64_t r = w_mul(w_mul(a.at(z, y)[i], b.at(z, y)[i], P.q, P.qn), P.r2, P.q, P.qn);
    if (has_c) r = w_add(r, cc.at(z, y)[i], P.q);
    out.at(z, y)[i] = r;
}

void wide_mul_add(const Context& c, Batch<const uint64_t> a, Batch<const uint64_t> b, Batch<const uint64_t> cc,
                  bool has_c, Batch<uint64_t> out, int limbs, int polys, PrimeMap pm, hipStream_t s) {
    ProfScope prof(s, "wide_mul_add");
    k_w_mul_add<<<w_grid(c.n, limbs, polys), 256, 0, s>>>(a, b, cc, has_c ? 1 : 0, out, (int)c.n, pm, c.d_wp);
    w_check();
}

// copy `polys` polynomials of P elements between strided layouts
static void w_copy_polys(uint64_t* dst, long long dst_ps, const uint64_t* src, long long src_ps, long long P, int polys,
                         hipStream_t s) {
    if (polys == 1 || (dst_ps == P && src_ps == P)) {
        RCK_CHECK_HIP(hipMemcpyAsync(dst, src, (size_t)P * polys * 8, hipMemcpyDeviceToDevice, s));
        return;
    }
    RCK_CHECK_HIP(hipMemcpy2DAsync(dst, (size_t)dst_ps * 8, src, (size_t)src_ps * 8, (size_t)P * 8, (size_t)polys,
                                   hipMemcpyDeviceToDevice, s));
}

void wide_poly_mul(Context& c, const uint64_t* a, long long a_ps, const uint64_t* b, long long b_ps, uint64_t* out,
                   long long out_ps, int limbs, int polys, PrimeMap pm, hipStream_t s) {
    const long long n = c.n, P = (long long)limbs * n;
    const int bpolys = b_ps ? polys : 1;
    uint64_t* t = static_cast<uint64_t*>(c.workspace(0, (size_t)(polys + bpolys) * P * 8));
    uint64_t* tb = t + (size_t)polys * P;
    w_copy_polys(t, P, a, a_ps, P, polys, s);
    w_copy_polys(tb, P, b, b_ps ? b_ps : P, P, bpolys, s);
    wide_ntt(c, false, t, n, P, limbs, polys, pm, s);
    wide_ntt(c, false, tb, n, P, limbs, bpolys, pm, s);
    wide_mul_add(c, {t, n, P}, {tb, n, b_ps ? P : 0}, {nullptr, 0, 0}, false, {t, n, P}, limbs, polys, pm, s);
    wide_ntt(c, true, t, n, P, limbs, polys, pm, s);
    w_copy_polys(out, out_ps ? out_ps : P, t, P, P, polys, s);
}

// d0 = a0 b0, d1 = a0 b1 + a1 b0, d2 = a1 b1 on NTT forms x = [a0 a1 b0 b1]
__global__ void k_w_tensor(const uint64_t* __restrict__ x, uint64_t* __restrict__ d, long long P, int n, PrimeMap pm,
                           const WidePrime* __restrict__ wp) {
    const long long c = w_gidx();
    if (c >= n) return;
    const int y = blockIdx.y;
    const WidePrime W = wp[pm(y)];
    const long long o = (long long)y * n + c;
    const uint64_t a0 = w_mul(x[o], W.r2, W.q, W.qn), a1 = w_mul(x[P + o], W.r2, W.q, W.qn);  // times R
    const uint64_t b0 = x[2 * P + o], b1 = x[3 * P + o];
    d[o] = w_mul(a0, b0, W.q, W.qn);
    d[P + o] = w_add(w_mul(a0, b1, W.q, W.qn), w_mul(a1, b0, W.q, W.qn), W.q);
    d[2 * P + o] = w_mul(a1, b1, W.q, W.qn);
}

// multiply_ciphertexts (src/rns/ops.rs:21-57): (a0 b0, a0 b1 + a1 b0, a1 b1), each a negacyclic product
void wide_mult(Context& c, const uint64_t* lhs, const uint64_t* rhs, int limbs, uint64_t* out, hipStream_t s) {
    const long long n = c.n, P = (long long)limbs * n;
    PrimeMap pm{limbs, (int)c.L, 0};
    uint64_t* x = static_cast<uint64_t*>(c.workspace(0, (size_t)4 * P * 8));
    uint64_t* d = static_cast<uint64_t*>(c.workspace(1, (size_t)3 * P * 8));
    RCK_CHECK_HIP(hipMemcpyAsync(x, lhs, (size_t)2 * P * 8, hipMemcpyDeviceToDevice, s));
    RCK_CHECK_HIP(hipMemcpyAsync(x + 2 * P, rhs, (size_t)2 * P * 8, hipMemcpyDeviceToDevice, s));
    wide_ntt(c, false, x, n, P, limbs, 4, pm, s);
    {
        ProfScope prof(s, "wide_tensor");
        k_w_tensor<<<w_grid(n, limbs, 1), 256, 0, s>>>(x, d, P, (int)n, pm, c.d_wp);
        w_check();
    }
    wide_ntt(c, true, d, n, P, limbs, 3, pm, s);
    RCK_CHECK_HIP(hipMemcpyAsync(out, d, (size_t)3 * P * 8, hipMemcpyDeviceToDevice, s));
}

// ---------------------------------------------------------------------------------------
// keys
// ---------------------------------------------------------------------------------------
// x -> x R (Montgomery form) or back
__global__ void k_w_mont(uint64_t* __restrict__ data, long long ls, long long ps, int n, int to, PrimeMap pm,
                         const WidePrime* __restrict__ wp) {
    const long long i = w_gidx();
    if (i >= n) return;
    const WidePrime P = wp[pm(blockIdx.y)];
    uint64_t* p = data + (long long)blockIdx.z * ps + (long long)blockIdx.y * ls + i;
    *p = w_mul(*p, to ? P.r2 : 1, P.q, P.qn);
}

void wide_key_create(Context& c, KeyDev* k, const uint64_t* key_dev) {
    const long long n = c.n, P = (long long)k->limbs * n;
    const size_t bytes = (size_t)k->digits * 2 * P * 8;
    RCK_CHECK_HIP(hipMalloc(&k->wdata, bytes));
    RCK_CHECK_HIP(hipMemcpy(k->wdata, key_dev, bytes, hipMemcpyDeviceToDevice));
    PrimeMap pm{(int)(k->level + 1), (int)c.L, 0};
    const int polys = (int)k->digits * 2;
    wide_ntt(c, false, k->wdata, n, P, (int)k->limbs, polys, pm, nullptr);
    k_w_mont<<<w_grid(n, (int)k->limbs, polys), 256>>>(k->wdata, n, P, (int)n, 1, pm, c.d_wp);
    w_check();
    RCK_CHECK_HIP(hipDeviceSynchronize());
}

void wide_key_export(Context& c, const KeyDev* k, uint64_t* out_host) {
    const long long n = c.n, P = (long long)k->limbs * n;
    PrimeMap pm{(int)(k->level + 1), (int)c.L, 0};
    uint64_t* t = static_cast<uint64_t*>(c.workspace(8, (size_t)2 * P * 8));
    for (uint32_t d = 0; d < k->digits; d++) {
        RCK_CHECK_HIP(hipMemcpy(t, k->wdata + (size_t)d * 2 * P, (size_t)2 * P * 8, hipMemcpyDeviceToDevice));
        k_w_mont<<<w_grid(n, (int)k->limbs, 2), 256>>>(t, n, P, (int)n, 0, pm, c.d_wp);
        w_check();
        wide_ntt(c, true, t, n, P, (int)k->limbs, 2, pm, nullptr);
        RCK_CHECK_HIP(hipMemcpy(out_host + (size_t)d * 2 * P, t, (size_t)2 * P * 8, hipMemcpyDeviceToHost));
    }
}

// ---------------------------------------------------------------------------------------
// keyswitch
// ---------------------------------------------------------------------------------------
// Balanced base-2^30 decomposition (keyswitching.rs:196-236, 381-467), one thread per coefficient,
// as k_decompose with 64-bit moduli: Garner digits m_i (LDS), Horner into 30-bit words with q and m_i
// split into three words each, centring, balanced digits with the final carry dropped.
constexpr int kWDecBS = 64;

template <int W>
__global__ __launch_bounds__(kWDecBS) void k_w_decompose(const uint64_t* __restrict__ poly, int n, int nq, int levels,
                                                         const WidePrime* __restrict__ wp,
                                                         const uint64_t* __restrict__ invq, int T,
                                                         const uint32_t* __restrict__ qw,
                                                         const uint32_t* __restrict__ hw, int32_t* __restrict__ digits) {
    extern __shared__ uint64_t lds_wm[];  // [nq][kWDecBS]
    const int tid = threadIdx.x;
    const long long c = (long long)blockIdx.x * kWDecBS + tid;
    if (c >= n) return;
    // Garner in the reference's order: m_i = (((r_i - m_0) q_0^{-1} - m_1) q_1^{-1} - ...) mod q_i
    for (int i = 0; i < nq; i++) {
        const WidePrime P = wp[i];
        uint64_t v = poly[(long long)i * n + c];
        for (int j = 0; j < i; j++) {
            const uint64_t inv = invq[(long long)i * T + j];
            // (v - m_j) q_j^{-1} = v q_j^{-1} - m_j q_j^{-1}: m_j < q_j need not be below q_i
            v = w_sub(w_mul(v, inv, P.q, P.qn), w_mul(lds_wm[j * kWDecBS + tid], inv, P.q, P.qn), P.q);
        }
        lds_wm[i * kWDecBS + tid] = v;
    }
    // Horner over 30-bit words: X = X q_i + m_i, i = nq-1 .. 0.  With q = q0 + q1 2^30 + q2 2^60 and
    // m = m0 + m1 2^30 + m2 2^60, word k is the low 30 bits of
    //   S_k = x_k q0 + x_{k-1} q1 + x_{k-2} q2 + m_k + C_k,  C_{k+1} = S_k >> 30,
    // and S_k < 2^60 + 2^60 + 2^34 + 2^30 + 2^32 never overflows.  After step i, X < 2^{64 (nq - i)}.
    uint32_t x[W];
#pragma unroll
    for (int k = 0; k < W; k++) x[k] = 0;
    for (int i = nq - 1; i >= 0; i--) {
        const uint64_t q = wp[i].q, mi = lds_wm[i * kWDecBS + tid];
        const uint64_t q0 = q & 0x3fffffffu, q1 = (q >> 30) & 0x3fffffffu, q2 = q >> 60;
        const uint64_t mw0 = mi & 0x3fffffffu, mw1 = (mi >> 30) & 0x3fffffffu, mw2 = mi >> 60;
        const int k_len = min(W, (64 * (nq - i) + 29) / 30 + 1);
        uint64_t C = 0;
        uint32_t p1 = 0, p2 = 0;  // x_{k-1}, x_{k-2} before this step
#pragma unroll
        for (int k = 0; k < W; k++) {
            if (k < k_len) {
                const uint64_t mk = k == 0 ? mw0 : (k == 1 ? mw1 : (k == 2 ? mw2 : 0));
                const uint64_t S = (uint64_t)x[k] * q0 + (uint64_t)p1 * q1 + (uint64_t)p2 * q2 + mk + C;
                p2 = p1;
                p1 = x[k];
                x[k] = (uint32_t)S & 0x3fffffffu;
                C = S >> 30;
            }
        }
    }
    // centring: X > floor(Q/2) -> X - Q (two's complement over the W words)
    bool gt = false, decided = false;
#pragma unroll
    for (int k = W - 1; k >= 0; k--) {
        const uint32_t h = hw[k];
        if (!decided && x[k] != h) {
            gt = x[k] > h;
            decided = true;
        }
    }
    if (gt) {
        uint32_t borrow = 0;
#pragma unroll
        for (int k = 0; k < W; k++) {
            const uint32_t t = x[k] - qw[k] - borrow;
            x[k] = t & 0x3fffffffu;
            borrow = t >> 31;
        }
    }
    // balanced digits, final carry dropped (keyswitching.rs:434-454)
    const uint32_t sx = (x[W - 1] >> 29) ? 0x3fffffffu : 0u;
    int carry = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        if (i < levels) {
            int d = (int)x[i] + carry;
            carry = d >= (1 << 29);
            digits[(long long)i * n + c] = carry ? d - (1 << 30) : d;
        }
    }
    for (int i = W; i < levels; i++) {
        int d = (int)sx + carry;
        carry = d >= (1 << 29);
        digits[(long long)i * n + c] = carry ? d - (1 << 30) : d;
    }
}

template <int W>
static void w_dec_t(const Context& c, const uint64_t* poly, int nq, int levels, int32_t* digits, hipStream_t s) {
    const uint32_t* qw = c.d_qwords + (size_t)nq * 2 * kWords30;
    ProfScope prof(s, "wide_decompose");
    k_w_decompose<W><<<(c.n + kWDecBS - 1) / kWDecBS, kWDecBS, (size_t)nq * kWDecBS * 8, s>>>(
        poly, (int)c.n, nq, levels, c.d_wp, c.d_winvq, (int)(c.L + c.K), qw, qw + kWords30, digits);
    w_check();
}

static void wide_decompose(const Context& c, const uint64_t* poly, int nq, int levels, int32_t* digits, hipStream_t s) {
    const int need = std::max(1, (int)(c.qbits[nq] + 29) / 30);
    if (need <= 8)
        w_dec_t<8>(c, poly, nq, levels, digits, s);
    else if (need <= 16)
        w_dec_t<16>(c, poly, nq, levels, digits, s);
    else if (need <= 32)
        w_dec_t<32>(c, poly, nq, levels, digits, s);
    else if (need <= 48)
        w_dec_t<48>(c, poly, nq, levels, digits, s);
    else if (need <= kWords30)
        w_dec_t<kWords30>(c, poly, nq, levels, digits, s);
    else
        RCK_REQUIRE(false, "decomposition needs more than 2160 bits");
}

// digit d (|d| < 2^29) as a residue mod every eval prime (its own centred lift, keyswitching.rs:242-276):
// out[i][j][x] for digit i, eval limb j
__global__ void k_w_digits(const int32_t* __restrict__ digits, uint64_t* __restrict__ out, int n, int E, PrimeMap pm,
                           const WidePrime* __restrict__ wp) {
    const long long x = w_gidx();
    if (x >= n) return;
    const int j = blockIdx.y, i = blockIdx.z;
    const uint64_t q = wp[pm(j)].q;
    const int32_t d = digits[(long long)i * n + x];
    out[((long long)i * E + j) * n + x] = d >= 0 ? (uint64_t)d : q - (uint64_t)(-(long long)d);
}

// acc[comp][j][x] = sum_i dn[i][j][x] key[i][comp][j][x] (key times R)
__global__ void k_w_mac(const uint64_t* __restrict__ dn, const uint64_t* __restrict__ key, uint64_t* __restrict__ acc,
                        int n, int D, int E, PrimeMap pm, const WidePrime* __restrict__ wp) {
    const long long x = w_gidx();
    if (x >= n) return;
    const int j = blockIdx.y;
    const WidePrime P = wp[pm(j)];
    const long long EN = (long long)E * n, o = (long long)j * n + x;
    uint64_t a0 = 0, a1 = 0;
    for (int i = 0; i < D; i++) {
        const uint64_t d = dn[(long long)i * EN + o];
        a0 = w_add(a0, w_mul(d, key[(long long)i * 2 * EN + o], P.q, P.qn), P.q);
        a1 = w_add(a1, w_mul(d, key[(long long)i * 2 * EN + EN + o], P.q, P.qn), P.q);
    }
    acc[o] = a0;
    acc[EN + o] = a1;
}

// mod_down_eval_to_q (keyswitching.rs:278-322) with the optional adds, as k_mod_down:
//   x_P = Garner(aux residues), c = centred(x_P) mod q_i, out = (x_i - c) P^{-1} (+ add) mod q_i.
__global__ void k_w_mod_down(const uint64_t* __restrict__ acc, int n, int nq, int K, int L, int T,
                             const uint64_t* __restrict__ add0, const uint64_t* __restrict__ add1,
                             uint64_t* __restrict__ out, long long out_ps, const WidePrime* __restrict__ wp,
                             const uint64_t* __restrict__ invq, const uint64_t* __restrict__ pmodq,
                             const uint64_t* __restrict__ pinvq, const uint64_t* __restrict__ paux,
                             const uint64_t* __restrict__ phalf) {
    const long long c = w_gidx();
    if (c >= n) return;
    const int z = blockIdx.y;
    const long long E = nq + K;
    const uint64_t* a = acc + (long long)z * E * n;
    uint64_t m[kMaxP];
    for (int k = 0; k < K; k++) {
        const WidePrime Pk = wp[L + k];
        uint64_t v = a[(long long)(nq + k) * n + c];
        for (int j = 0; j < k; j++) {
            const uint64_t inv = invq[(long long)(L + k) * T + (L + j)];
            v = w_sub(w_mul(v, inv, Pk.q, Pk.qn), w_mul(m[j], inv, Pk.q, Pk.qn), Pk.q);
        }
        m[k] = v;
    }
    // X_P = m_0 + m_1 p_0 + ... as kMaxP 64-bit words (Horner from the top), compared with floor(P/2)
    uint64_t xw[kMaxP];
#pragma unroll
    for (int w = 0; w < kMaxP; w++) xw[w] = 0;
    xw[0] = m[K - 1];
    for (int k = K - 2; k >= 0; k--) {
        const uint64_t p = wp[L + k].q;
        uint64_t carry = m[k];
#pragma unroll
        for (int w = 0; w < kMaxP; w++) {
            uint64_t lo = xw[w] * p, hi = __umul64hi(xw[w], p);
            lo += carry;
            hi += lo < carry;
            xw[w] = lo;
            carry = hi;
        }
    }
    bool gt = false, decided = false;
#pragma unroll
    for (int w = kMaxP - 1; w >= 0; w--) {
        if (!decided && xw[w] != phalf[w]) {
            gt = xw[w] > phalf[w];
            decided = true;
        }
    }
    const uint64_t* add = z == 0 ? add0 : add1;
    uint64_t* o = out + (long long)z * out_ps;
    for (int i = 0; i < nq; i++) {
        const WidePrime Pi = wp[i];
        // centred(x_P) mod q_i by Horner in Montgomery form (evaluate_mixed_radix_mod, keyswitching.rs:412-423)
        uint64_t cv = w_mul(m[K - 1], Pi.r2, Pi.q, Pi.qn);
        for (int k = K - 2; k >= 0; k--)
            cv = w_add(w_mul(cv, paux[(long long)k * L + i], Pi.q, Pi.qn), w_mul(m[k], Pi.r2, Pi.q, Pi.qn), Pi.q);
        cv = w_mul(cv, 1, Pi.q, Pi.qn);
        if (gt) cv = w_sub(cv, pmodq[i], Pi.q);
        uint64_t r = w_mul(w_sub(a[(long long)i * n + c], cv, Pi.q), pinvq[i], Pi.q, Pi.qn);
        if (add) r = w_add(r, add[(long long)i * n + c], Pi.q);
        o[(long long)i * n + c] = r;
    }
}

void wide_keyswitch(Context& c, const uint64_t* poly, int limbs, const KeyDev* key, const uint64_t* add0,
                    const uint64_t* add1, uint64_t* out, hipStream_t s) {
    RCK_REQUIRE(key->wdata != nullptr, "keyswitch key was created for another arithmetic of this context");
    const long long n = c.n, P = (long long)limbs * n;
    const int naux = (int)key->aux, E = limbs + naux;
    const int levels = naux ? (int)((c.qbits[limbs] + 29) / 30) : (int)key->digits;
    const long long EN = (long long)E * n;
    PrimeMap pm{limbs, (int)c.L, 0};
    int32_t* digits = static_cast<int32_t*>(c.workspace(5, (size_t)levels * n * 4));
    uint64_t* dn = static_cast<uint64_t*>(c.workspace(6, (size_t)levels * EN * 8));
    uint64_t* acc = static_cast<uint64_t*>(c.workspace(4, (size_t)2 * EN * 8));
    wide_decompose(c, poly, limbs, levels, digits, s);
    {
        ProfScope prof(s, "wide_digits");
        k_w_digits<<<w_grid(n, E, levels), 256, 0, s>>>(digits, dn, (int)n, E, pm, c.d_wp);
        w_check();
    }
    wide_ntt(c, false, dn, n, EN, E, levels, pm, s);
    {
        ProfScope prof(s, "wide_mac");
        k_w_mac<<<w_grid(n, E, 1), 256, 0, s>>>(dn, key->wdata, acc, (int)n, levels, E, pm, c.d_wp);
        w_check();
    }
    wide_ntt(c, true, acc, n, EN, E, 2, pm, s);
    if (naux) {
        ProfScope prof(s, "wide_mod_down");
        k_w_mod_down<<<w_grid(n, 2, 1), 256, 0, s>>>(acc, (int)n, limbs, naux, (int)c.L, (int)(c.L + c.K), add0, add1,
                                                     out, P, c.d_wp, c.d_winvq, c.d_wpmodq, c.d_wpinvq, c.d_wpaux,
                                                     c.d_wphalf);
        w_check();
        return;
    }
    // Q-only (keyswitching.rs:130-157): the accumulators are the result, plus the adds
    const uint64_t* adds[2] = {add0, add1};
    for (int z = 0; z < 2; z++) {
        if (adds[z])
            wide_addsub(c, {adds[z], n, 0}, {acc + z * EN, n, 0}, {out + z * P, n, 0}, limbs, 1, 0, pm, s);
        else
            RCK_CHECK_HIP(hipMemcpyAsync(out + z * P, acc + z * EN, (size_t)P * 8, hipMemcpyDeviceToDevice, s));
    }
}

}  // namespace rck
