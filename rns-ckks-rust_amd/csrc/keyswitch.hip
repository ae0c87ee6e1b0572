// Keyswitch pipeline: balanced gadget decomposition, digit NTT fused with the key
// multiply-accumulate, and mod-down by the aux modulus P.
//
// keyswitch_rns_hybrid (src/rns/keyswitching.rs:159-194) computes, over Q_l || P,
//   acc_k = sum_i lift(d_i) * ksk_i.c_k  and returns  mod_down(acc_0), mod_down(acc_1).
// Here the digits d_i are produced once per coefficient as signed 32-bit integers (they are
// prime-independent), every (digit, prime) pair is NTT'd straight from those integers (a digit
// |d| < 2^29 < Q_l/2 is its own centred lift, which is exactly what lift_poly_to_context gives
// it on the aux primes), and the products with the NTT-form key are summed in registers.
#include "ops.hpp"

namespace rck {

// ---------------------------------------------------------------------------------------
// Decomposition: one thread per coefficient.
//   Garner mixed radix of the residues (keyswitching.rs:397-410; m_j live in LDS),
//   Horner into a W-word binary integer X in [0, Q_l), centring X > floor(Q_l/2) -> X - Q_l
//   (two's complement over W words), then balanced base-2^30 digits of X, low digit first,
//   with the reference's carry rule and the final carry dropped (keyswitching.rs:434-454).
// The balanced digits of a value mod 2^{30 levels} are unique, so this equals the reference's
// base-2^30 digit arithmetic bit for bit.
// ---------------------------------------------------------------------------------------
constexpr int kDecBS = 128;

template <int W>
__global__ __launch_bounds__(kDecBS) void k_decompose(const uint64_t* __restrict__ poly, int n, int nq, int levels,
                                                      const DevPrime* __restrict__ primes,
                                                      const double* __restrict__ invq, int T,
                                                      const uint32_t* __restrict__ qw, const uint32_t* __restrict__ hw,
                                                      int32_t* __restrict__ digits) {
    extern __shared__ double m[];  // [nq][kDecBS]
    const int tid = threadIdx.x;
    const long long c = (long long)blockIdx.x * kDecBS + tid;
    if (c >= n) return;
    // Garner (row order as in the reference: m_i from r_i and m_0 .. m_{i-1})
    for (int i = 0; i < nq; i++) {
        const double q = primes[i].q, qinv = primes[i].qinv;
        double v = (double)poly[(long long)i * n + c];
        for (int j = 0; j < i; j++) {
            double a = reduce(v - m[j * kDecBS + tid], q, qinv);
            v = mulmod(a, invq[(long long)i * T + j], q, qinv);
        }
        m[i * kDecBS + tid] = i == 0 ? v : canon(v, q, qinv);
    }
    // Horner: X = m_{nq-1}; X = X * q_i + m_i
    uint32_t x[W];
#pragma unroll
    for (int k = 0; k < W; k++) x[k] = 0;
    {
        uint64_t top = (uint64_t)m[(nq - 1) * kDecBS + tid];
        x[0] = (uint32_t)top;
        x[1] = (uint32_t)(top >> 32);
    }
    for (int i = nq - 2; i >= 0; i--) {
        const uint64_t q = (uint64_t)primes[i].q;
        const uint32_t qlo = (uint32_t)q, qhi = (uint32_t)(q >> 32);
        uint64_t carry = 0;
        uint32_t y[W];
#pragma unroll
        for (int k = 0; k < W; k++) {  // Y = X * qlo  (t < 2^64 - 2^32)
            uint64_t t = (uint64_t)x[k] * qlo + carry;
            y[k] = (uint32_t)t;
            carry = t >> 32;
        }
        carry = (uint64_t)m[i * kDecBS + tid];  // + m_i (t < 2^53 below)
        uint32_t prev = 0;
#pragma unroll
        for (int k = 0; k < W; k++) {  // X = Y + (X * qhi) << 32 + m_i
            uint64_t t = (uint64_t)y[k] + (uint64_t)prev * qhi + carry;
            prev = x[k];
            x[k] = (uint32_t)t;
            carry = t >> 32;
        }
    }
    // centring: X > floor(Q/2)  ->  X - Q
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
        uint64_t borrow = 0;
#pragma unroll
        for (int k = 0; k < W; k++) {
            uint64_t t = (uint64_t)x[k] - qw[k] - borrow;
            x[k] = (uint32_t)t;
            borrow = (t >> 32) & 1;
        }
    }
    // balanced digits of X (two's complement, sign-extended over W words)
    int carry = 0;
    constexpr int DMAX = (32 * W) / 30;
#pragma unroll
    for (int i = 0; i < DMAX; i++) {
        if (i < levels) {
            const int bit = 30 * i, w = bit >> 5, sh = bit & 31;
            uint64_t lo = x[w];
            uint64_t hi = (w + 1 < W) ? x[w + 1] : (x[W - 1] >> 31 ? 0xffffffffu : 0u);
            int raw = (int)(((lo | (hi << 32)) >> sh) & 0x3fffffffu);
            int d = raw + carry;
            if (d >= (1 << 29)) {
                d -= (1 << 30);
                carry = 1;
            } else {
                carry = 0;
            }
            digits[(long long)i * n + c] = d;
        }
    }
}

template <int W>
static void dec_t(const Context& c, const uint64_t* poly, int nq, int levels, int32_t* digits, hipStream_t s) {
    const uint32_t* qw = c.d_qwords + (size_t)nq * 2 * kWords;
    ProfScope ps(s, "decompose");
    k_decompose<W><<<dim3((c.n + kDecBS - 1) / kDecBS), kDecBS, (size_t)nq * kDecBS * 8, s>>>(poly, c.n, nq, levels, c.d_primes, c.d_invq,
                                                                        c.L + c.K, qw, qw + kWords, digits);
    RCK_CHECK_HIP(hipGetLastError());
}

void launch_decompose(const Context& c, const uint64_t* poly, int nq, int levels, int32_t* digits, hipStream_t s) {
    RCK_REQUIRE(nq >= 1 && nq <= (int)c.L, "bad limb count for decomposition");
    const int need = std::max((int)(c.qbits[nq] + 32) / 32 + 1, (30 * levels + 31) / 32 + 1);
    if (need <= 8)
        dec_t<8>(c, poly, nq, levels, digits, s);
    else if (need <= 16)
        dec_t<16>(c, poly, nq, levels, digits, s);
    else if (need <= 24)
        dec_t<24>(c, poly, nq, levels, digits, s);
    else if (need <= 32)
        dec_t<32>(c, poly, nq, levels, digits, s);
    else if (need <= 40)
        dec_t<40>(c, poly, nq, levels, digits, s);
    else if (need <= 48)
        dec_t<48>(c, poly, nq, levels, digits, s);
    else if (need <= 64)
        dec_t<64>(c, poly, nq, levels, digits, s);
    else
        RCK_REQUIRE(false, "decomposition needs more than 2048 bits");
}

// ---------------------------------------------------------------------------------------
// Digit MAC: row pass of the forward NTT of every digit for one eval prime, fused with the
// multiply-accumulate against the NTT-form key.  The column pass ran before into `cols`
// ([levels][N] doubles for this prime).
// ---------------------------------------------------------------------------------------
template <int LOGR>
__global__ __launch_bounds__(256) void k_digit_mac_rows(const double* __restrict__ cols, int levels,
                                                        const double* __restrict__ key, long long key_comp_stride,
                                                        long long key_digit_stride, double* __restrict__ acc0,
                                                        double* __restrict__ acc1, int logn,
                                                        const DevPrime* __restrict__ primes, int prime) {
    constexpr int R = 1 << LOGR, TPS = R / 16, S = 256 / TPS;
    __shared__ double lds[S * (R + R / 16 + 1)];
    const int n = 1 << logn, N1 = n >> LOGR;
    const int tid = threadIdx.x, tau = tid % TPS, row = tid / TPS;
    const int b = blockIdx.x * S + row;
    const DevPrime P = primes[prime];
    const double q = P.q, qinv = P.qinv;
    double a0[16], a1[16];
#pragma unroll
    for (int e = 0; e < 16; e++) a0[e] = a1[e] = 0.0;
    const long long off = (long long)b * R;
    for (int i = 0; i < levels; i++) {
        const double* p = cols + (long long)i * n + off;
        double v[16];
#pragma unroll
        for (int e = 0; e < 16; e++) v[e] = p[tau + TPS * e];
        fwd_phase_a(v, P.fw, N1 + b, q, qinv);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 16; e++) lds[row_lds<R>(row, tau + TPS * e)] = v[e];
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 16; e++) v[e] = lds[row_lds<R>(row, 16 * tau + e)];
        fwd_phase_b<LOGR>(v, P.fw, N1 + b, tau, q, qinv);
        const double* k0 = key + (long long)i * key_digit_stride + off + 16 * tau;
        const double* k1 = k0 + key_comp_stride;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            a0[e] += mulmod(v[e], k0[e], q, qinv);
            a1[e] += mulmod(v[e], k1[e], q, qinv);
        }
        if ((i & 3) == 3) {
            reduce_all(a0, q, qinv);
            reduce_all(a1, q, qinv);
        }
    }
    reduce_all(a0, q, qinv);
    reduce_all(a1, q, qinv);
#pragma unroll
    for (int e = 0; e < 16; e++) {
        acc0[off + 16 * tau + e] = a0[e];
        acc1[off + 16 * tau + e] = a1[e];
    }
}

template <int LOGR>
static void mac_rows_t(const Context& c, const double* cols, int levels, const double* key, long long comp_stride,
                       long long digit_stride, double* acc0, double* acc1, int prime, hipStream_t s) {
    constexpr int S = 256 / ((1 << LOGR) / 16);
    const int n1 = 1 << (c.logn - LOGR);
    ProfScope ps(s, "digit_mac_rows");
    k_digit_mac_rows<LOGR><<<dim3(n1 / S), 256, 0, s>>>(cols, levels, key, comp_stride, digit_stride, acc0, acc1,
                                                         c.logn, c.d_primes, prime);
    RCK_CHECK_HIP(hipGetLastError());
}

// Small rings: plain forward NTT of every digit for every prime, then pointwise MAC.
__global__ void k_mac_small(const double* __restrict__ dn, int levels, const double* __restrict__ key, int limbs,
                            int n, double* __restrict__ acc, PrimeMap pm, const DevPrime* __restrict__ primes) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int y = blockIdx.y;
    const DevPrime P = primes[pm(y)];
    const long long poly = (long long)limbs * n;
    double s0 = 0.0, s1 = 0.0;
    for (int d = 0; d < levels; d++) {
        const double v = dn[d * poly + (long long)y * n + i];
        s0 += mulmod(v, key[(2LL * d) * poly + (long long)y * n + i], P.q, P.qinv);
        s1 += mulmod(v, key[(2LL * d + 1) * poly + (long long)y * n + i], P.q, P.qinv);
        if ((d & 3) == 3) {
            s0 = reduce(s0, P.q, P.qinv);
            s1 = reduce(s1, P.q, P.qinv);
        }
    }
    acc[(long long)y * n + i] = reduce(s0, P.q, P.qinv);
    acc[poly + (long long)y * n + i] = reduce(s1, P.q, P.qinv);
}

void launch_digit_mac(const Context& c, const int32_t* digits, int levels, const double* key, int nq, int naux,
                      double* acc, hipStream_t s) {
    const int limbs = nq + naux;
    const long long n = c.n, poly = (long long)limbs * n;
    PrimeMap pm{nq, (int)c.L, 0};
    if (c.logn <= 11) {
        double* dn = static_cast<double*>(const_cast<Context&>(c).workspace(7, (size_t)levels * poly * 8));
        // every digit row is NTT'd for every prime: in limb stride 0, poly stride n
        launch_fwd(c, kInI32, digits, 0, n, dn, n, poly, limbs, levels, pm, s);
        ProfScope ps(s, "mac_small");
        k_mac_small<<<dim3((c.n + 255) / 256, limbs), 256, 0, s>>>(dn, levels, key, limbs, c.n, acc, pm, c.d_primes);
        RCK_CHECK_HIP(hipGetLastError());
        return;
    }
    double* cols = static_cast<double*>(const_cast<Context&>(c).workspace(7, (size_t)levels * n * 8));
    int n1, n2;
    split_logn(c.logn, &n1, &n2);
    for (int j = 0; j < limbs; j++) {
        PrimeMap pj{nq, (int)c.L, j};
        launch_fwd_cols(c, kInI32, digits, 0, n, cols, 0, n, 1, levels, pj, s);
        const int prime = j < nq ? j : (int)c.L + (j - nq);
        const double* kj = key + (long long)j * n;
        double* a0 = acc + (long long)j * n;
        double* a1 = acc + poly + (long long)j * n;
        switch (n2) {
            case 6: mac_rows_t<6>(c, cols, levels, kj, poly, 2 * poly, a0, a1, prime, s); break;
            case 7: mac_rows_t<7>(c, cols, levels, kj, poly, 2 * poly, a0, a1, prime, s); break;
            case 8: mac_rows_t<8>(c, cols, levels, kj, poly, 2 * poly, a0, a1, prime, s); break;
            default: RCK_REQUIRE(false, "unsupported ring degree");
        }
    }
}

// ---------------------------------------------------------------------------------------
// mod-down (keyswitching.rs:278-322) with the optional relinearization add.
//   x_P = Garner(acc_P) in [0, P);  c = centred(x_P) mod q_i;  out = (x_i - c) P^{-1} mod q_i.
// One thread per (coefficient, Q-limb, accumulator).
// ---------------------------------------------------------------------------------------
__global__ void k_mod_down(const uint64_t* __restrict__ acc, int n, int nq, int K, int L, int T,
                           const uint64_t* __restrict__ add0, const uint64_t* __restrict__ add1,
                           uint64_t* __restrict__ out,
                           long long out_ps, const DevPrime* __restrict__ primes, const double* __restrict__ invq,
                           const uint32_t* __restrict__ pwords, const double* __restrict__ pmodq,
                           const double* __restrict__ pinvq, const double* __restrict__ pmod_aux) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int i = blockIdx.y, z = blockIdx.z;
    const long long limbs = nq + K;
    const uint64_t* a = acc + (long long)z * limbs * n;
    // Garner over the aux primes (mixed_radix_digits, keyswitching.rs:397-410)
    double m[kMaxP];
    for (int k = 0; k < K; k++) {
        const DevPrime Pk = primes[L + k];
        double v = (double)a[(long long)(nq + k) * n + c];
        for (int j = 0; j < k; j++) {
            double d = reduce(v - m[j], Pk.q, Pk.qinv);
            v = mulmod(d, invq[(long long)(L + k) * T + (L + j)], Pk.q, Pk.qinv);
        }
        m[k] = k == 0 ? v : canon(v, Pk.q, Pk.qinv);
    }
    // X_P as words (<= 8 x 32 bits), compared with floor(P/2)
    uint32_t xw[8];
#pragma unroll
    for (int w = 0; w < 8; w++) xw[w] = 0;
    for (int k = K - 1; k >= 0; k--) {
        uint64_t pk = k == K - 1 ? 0 : (uint64_t)primes[L + k].q;
        if (k == K - 1) {
            uint64_t t = (uint64_t)m[k];
            xw[0] = (uint32_t)t;
            xw[1] = (uint32_t)(t >> 32);
            continue;
        }
        const uint32_t plo = (uint32_t)pk, phi = (uint32_t)(pk >> 32);
        uint64_t carry = 0;
        uint32_t y[8];
#pragma unroll
        for (int w = 0; w < 8; w++) {
            uint64_t t = (uint64_t)xw[w] * plo + carry;
            y[w] = (uint32_t)t;
            carry = t >> 32;
        }
        carry = (uint64_t)m[k];
        uint32_t prev = 0;
#pragma unroll
        for (int w = 0; w < 8; w++) {
            uint64_t t = (uint64_t)y[w] + (uint64_t)prev * phi + carry;
            prev = xw[w];
            xw[w] = (uint32_t)t;
            carry = t >> 32;
        }
    }
    bool gt = false, decided = false;
#pragma unroll
    for (int w = 7; w >= 0; w--) {
        const uint32_t h = pwords[8 + w];
        if (!decided && xw[w] != h) {
            gt = xw[w] > h;
            decided = true;
        }
    }
    // centred(x_P) mod q_i by Horner (evaluate_mixed_radix_mod, keyswitching.rs:412-423)
    const DevPrime Pi = primes[i];
    const double q = Pi.q, qinv = Pi.qinv;
    double cv = reduce(m[K - 1], q, qinv);
    for (int k = K - 2; k >= 0; k--) cv = reduce(mulmod(cv, pmod_aux[(long long)k * L + i], q, qinv) + m[k], q, qinv);
    if (gt) cv -= pmodq[i];
    const double xi = (double)a[(long long)i * n + c];
    double r = mulmod(reduce(xi - cv, q, qinv), pinvq[i], q, qinv);
    const uint64_t* add = z == 0 ? add0 : add1;
    if (add) r += (double)add[(long long)i * n + c];
    out[(long long)z * out_ps + (long long)i * n + c] = f64_to_u64(canon(r, q, qinv));
}

void launch_mod_down(const Context& c, const uint64_t* acc, int nq, const uint64_t* add0, const uint64_t* add1,
                     uint64_t* out, long long out_ps, hipStream_t s) {
    RCK_REQUIRE(c.K > 0, "mod-down requires an auxiliary basis");
    ProfScope ps(s, "mod_down");
    k_mod_down<<<dim3((c.n + 255) / 256, nq, 2), 256, 0, s>>>(acc, c.n, nq, c.K, c.L, c.L + c.K, add0, add1, out, out_ps, c.d_primes, c.d_invq, c.d_pwords,
                                                              c.d_pmodq, c.d_pinvq, c.d_pmod_aux);
    RCK_CHECK_HIP(hipGetLastError());
}

}  // namespace rck
