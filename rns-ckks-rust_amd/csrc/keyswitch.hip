// Keyswitch pipeline: balanced gadget decomposition, digit NTT fused with the key
// multiply-accumulate, and mod-down by the aux modulus P.
//
// keyswitch_rns_hybrid (src/rns/keyswitching.rs:159-194) computes, over Q_l || P,
//   acc_k = sum_i lift(d_i) * ksk_i.c_k  and returns  mod_down(acc_0), mod_down(acc_1).
// Here the digits d_i are produced once per coefficient as signed 32-bit integers (they are
// prime-independent), every (digit, prime) pair is NTT'd straight from those integers (a digit
// |d| < 2^29 < Q_l/2 is its own centred lift, which is exactly what lift_poly_to_context gives
// it on the aux primes), and the products with the NTT-form key are summed in registers.
#include "moddown.hpp"
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
constexpr int kDecBS = 64;

// NQM (>= nq, a multiple of 8): bound of the limb count; the Garner digits m_0 .. m_{nq-1} live in
// registers (loops fully unrolled over NQM with limb guards), so the workgroup's LDS holds only the
// parked words of X and more waves stay resident.  NQM = 0: the digits live in LDS (any nq).
// ginv != 0: the input is read through the automorphism X -> X^g (g * ginv = 1 mod 2N), i.e. the
// decomposition of sigma_g(poly): coefficient c is (+/-) poly[c * ginv mod 2N] (as k_galois).
// l1 (nullable; needs n % kDecBS == 0): l1[b][block] = sum of |d| over the block's coefficients and
// every digit, an integer below 2^41 (exact as a double).
template <int W, int NQM>
__global__ __launch_bounds__(kDecBS) void k_decompose(const uint64_t* __restrict__ poly, long long ps, int n, int nq,
                                                      int levels,
                                                      const DevPrime* __restrict__ primes,
                                                      const double* __restrict__ invq, int T,
                                                      const uint32_t* __restrict__ qw, const uint32_t* __restrict__ hw,
                                                      int32_t* __restrict__ digits, unsigned long long ginv,
                                                      double* __restrict__ l1) {
    // LDS (NQM == 0 only): [nq][kDecBS] Garner digits
    extern __shared__ double lds_m[];
    const int tid = threadIdx.x;
    const long long c = (long long)blockIdx.x * kDecBS + tid;
    if (c >= n) return;
    poly += (long long)blockIdx.y * ps;
    digits += (long long)blockIdx.y * levels * n;
    long long src = c;
    bool neg = false;
    if (ginv) {
        const unsigned long long k = ((unsigned long long)c * ginv) & (2ull * n - 1);
        neg = k >= (unsigned long long)n;
        src = neg ? (long long)k - n : (long long)k;
    }
    // residue of limb i of the (rotated) coefficient, as a double in [0, q_i)
    auto rd = [&](int i, double q) -> double {
        const uint64_t x = poly[(long long)i * n + src];
        return neg && x != 0 ? q - (double)x : (double)x;
    };
    // Garner (row order as in the reference: m_i from r_i and m_0 .. m_{i-1})
    double m[NQM > 0 ? NQM : 1];
    if constexpr (NQM > 0) {
#pragma unroll
        for (int i = 0; i < NQM; i++) {
            m[i] = 0.0;
            if (i < nq) {
                const double q = primes[i].q, qinv = primes[i].qinv;
                double v = rd(i, q);
#pragma unroll
                for (int j = 0; j < i; j++) {
                    // no reduction: every Q prime is below 2.5x the smallest (the launcher routes other
                    // chains to NQM = 0), so |v| <= q_i and 0 <= m_j < 2.5 q_i give |v - m_j| < 3.5 q_i,
                    // a valid mulmod input, and the product is <= 0.97 q_i
                    v = mulmod(v - m[j], invq[(long long)i * T + j], q, qinv);
                }
                m[i] = i == 0 ? v : canon(v, q, qinv);
            }
        }
    } else {
        for (int i = 0; i < nq; i++) {
            const double q = primes[i].q, qinv = primes[i].qinv;
            double v = rd(i, q);
            for (int j = 0; j < i; j++) {
                double a = reduce(v - lds_m[j * kDecBS + tid], q, qinv);
                v = mulmod(a, invq[(long long)i * T + j], q, qinv);
            }
            lds_m[i * kDecBS + tid] = i == 0 ? v : canon(v, q, qinv);
        }
    }
    // Horner over 30-bit words (the digits' base): X = 0; X = X * q_i + m_i for i = nq-1 .. 0.
    // With q = qhi 2^30 + qlo (qhi < 2^21), word k of X * q + m is the low 30 bits of
    //   S_k = x_k qlo + x_{k-1} qhi + C_k,  C_0 = m < 2^51,  C_{k+1} = S_k >> 30,
    // and S_k < 2^60 + 2^51 + 2^51 never overflows: two 32x32+64 multiply-adds, a mask and a shift
    // per word.  After step i, X < q_i ... q_{nq-1} < 2^{51 (nq - i)}, so words at or beyond
    // (51 (nq - i) + 29) / 30 + 1 stay zero and are skipped.
    uint32_t x[W];
#pragma unroll
    for (int k = 0; k < W; k++) x[k] = 0;
    auto step = [&](int k_len, uint64_t q, double mi) {
        const uint32_t qlo = (uint32_t)q & 0x3fffffffu, qhi = (uint32_t)(q >> 30);
        // mi is an integer in [0, 2^51): the low 52 bits of mi + 2^52 are mi itself
        uint64_t C = (uint64_t)__double_as_longlong(mi + 4503599627370496.0) & ((1ull << 52) - 1);
        uint32_t prev = 0;
#pragma unroll
        for (int k = 0; k < W; k++) {
            if (k < k_len) {
                const uint64_t S = (uint64_t)x[k] * qlo + ((uint64_t)prev * qhi + C);
                prev = x[k];
                x[k] = (uint32_t)S & 0x3fffffffu;
                C = S >> 30;
            }
        }
    };
    if constexpr (NQM > 0) {
        // fully unrolled and branch-free: every Garner digit and word index is a constant, the word
        // count of step i is bounded with NQM in place of nq (the extra words compute zeros), and the
        // steps i >= nq multiply by 1 and add 0 (a branch around them would make the compiler copy
        // every word at each join)
#pragma clang loop unroll(full)
        for (int i = NQM - 1; i >= 0; i--) {
            const bool on = i < nq;
            step(min(W, (51 * (NQM - i) + 29) / 30 + 1), on ? primes[i].qu : 1ull, on ? m[i] : 0.0);
        }
    } else {
        for (int i = nq - 1; i >= 0; i--)
            step(min(W, (51 * (nq - i) + 29) / 30 + 1), primes[i].qu, lds_m[i * kDecBS + tid]);
    }
    // centring: X > floor(Q/2)  ->  X - Q, two's complement over the W words (30 W >= bits(Q): the
    // centred value lies in (-Q/2, Q/2))
    // branch-free (selects): the first differing word from the top decides
    bool gt = false, decided = false;
#pragma unroll
    for (int k = W - 1; k >= 0; k--) {
        const uint32_t h = hw[k];
        const bool ne = x[k] != h;
        gt = (!decided & ne) ? x[k] > h : gt;
        decided |= ne;
    }
    {
        uint32_t borrow = 0;
#pragma unroll
        for (int k = 0; k < W; k++) {
            const uint32_t t = x[k] - qw[k] - borrow;  // in (-2^31, 2^30): bit 31 is the borrow
            x[k] = gt ? (t & 0x3fffffffu) : x[k];
            borrow = t >> 31;
        }
    }
    // balanced base-2^30 digits: digit i is word i plus the carry (words past W - 1 are the sign
    // extension); d >= 2^29 becomes d - 2^30 with a carry into the next digit
    const uint32_t sx = (x[W - 1] >> 29) ? 0x3fffffffu : 0u;
    int carry = 0;
    long long mag = 0;  // sum of |d| over the digits (< 2^35)
#pragma unroll
    for (int i = 0; i < W; i++) {
        if (i < levels) {
            int d = (int)x[i] + carry;
            carry = d >= (1 << 29);
            d = carry ? d - (1 << 30) : d;
            digits[(long long)i * n + c] = d;
            mag += d < 0 ? -d : d;
        }
    }
    for (int i = W; i < levels; i++) {
        int d = (int)sx + carry;
        carry = d >= (1 << 29);
        d = carry ? d - (1 << 30) : d;
        digits[(long long)i * n + c] = d;
        mag += d < 0 ? -d : d;
    }
    if (l1 != nullptr) {
        double t = (double)mag;
#pragma unroll
        for (int off = kDecBS / 2; off > 0; off >>= 1) t += __shfl_xor(t, off, kDecBS);
        if (tid == 0) l1[(long long)blockIdx.y * gridDim.x + blockIdx.x] = t;
    }
}

template <int W, int NQM>
static void dec_tq(const Context& c, const uint64_t* poly, int nq, int levels, int32_t* digits, hipStream_t s, int B,
                   long long ps, unsigned long long ginv, double* l1) {
    const uint32_t* qw = c.d_qwords + (size_t)nq * 2 * kWords30;
    ProfScope prof(s, "decompose");
    const size_t lds = NQM == 0 ? (size_t)nq * kDecBS * 8 : 0;
    k_decompose<W, NQM><<<dim3((c.n + kDecBS - 1) / kDecBS, B), kDecBS, lds, s>>>(
        poly, ps, c.n, nq, levels, c.d_primes, c.d_invq, c.L + c.K, qw, qw + kWords30, digits, ginv, l1);
    RCK_CHECK_HIP(hipGetLastError());
}

// Register-resident Garner digits up to 24 limbs (beyond that the unrolled digit array would spill)
// for chains whose Q primes lie within a factor 2.5 of each other (their Garner skips reductions);
// larger limb counts and other chains keep the digits in LDS (NQM = 0).  W words hold at most 30 W
// bits of Q, so NQM limbs of at most 51 bits need W >= 51 NQM / 30: W > 48 only pairs with NQM = 0.
template <int W>
static void dec_t(const Context& c, const uint64_t* poly, int nq, int levels, int32_t* digits, hipStream_t s, int B,
                  long long ps, unsigned long long ginv, double* l1) {
    if constexpr (W <= 48) {
        if (!c.garner_lazy) return dec_tq<W, 0>(c, poly, nq, levels, digits, s, B, ps, ginv, l1);
        if (nq <= 8) return dec_tq<W, 8>(c, poly, nq, levels, digits, s, B, ps, ginv, l1);
        if (nq <= 16) return dec_tq<W, 16>(c, poly, nq, levels, digits, s, B, ps, ginv, l1);
        if (nq <= 24) return dec_tq<W, 24>(c, poly, nq, levels, digits, s, B, ps, ginv, l1);
    }
    dec_tq<W, 0>(c, poly, nq, levels, digits, s, B, ps, ginv, l1);
}

void launch_decompose(const Context& c, const uint64_t* poly, int nq, int levels, int32_t* digits, hipStream_t s, int B,
                      long long ps, unsigned long long ginv, double* l1) {
    require_fp64(c);
    RCK_REQUIRE(l1 == nullptr || c.n % kDecBS == 0, "digit magnitudes need whole blocks");
    RCK_REQUIRE(nq >= 1 && nq <= (int)c.L, "bad limb count for decomposition");
    RCK_REQUIRE(levels >= 1, "decomposition needs at least one digit");
    // 30-bit words holding Q_nq (and the centred value in two's complement)
    const int need = std::max(1, (int)(c.qbits[nq] + 29) / 30);
    if (need <= 8)
        dec_t<8>(c, poly, nq, levels, digits, s, B, ps, ginv, l1);
    else if (need <= 16)
        dec_t<16>(c, poly, nq, levels, digits, s, B, ps, ginv, l1);
    else if (need <= 24)
        dec_t<24>(c, poly, nq, levels, digits, s, B, ps, ginv, l1);
    else if (need <= 32)
        dec_t<32>(c, poly, nq, levels, digits, s, B, ps, ginv, l1);
    else if (need <= 40)
        dec_t<40>(c, poly, nq, levels, digits, s, B, ps, ginv, l1);
    else if (need <= 48)
        dec_t<48>(c, poly, nq, levels, digits, s, B, ps, ginv, l1);
    else if (need <= 64)
        dec_t<64>(c, poly, nq, levels, digits, s, B, ps, ginv, l1);
    else if (need <= kWords30)
        dec_t<kWords30>(c, poly, nq, levels, digits, s, B, ps, ginv, l1);
    else
        RCK_REQUIRE(false, "decomposition needs more than 2160 bits");
}

// ---------------------------------------------------------------------------------------
// Digit column pass: the first log2(N1) stages of the forward NTT of digit z = (b, i) under the
// eval limbs jl = y * jper .. min(g, (y + 1) * jper) - 1 of a group (prime pm(jl)), written to
// out[z][jl][N].  A digit is prime-independent, so a thread loads its 16 int32 values once and
// transforms them under every prime of its range; the stores of one prime drain while the next
// prime's stages run (streaming stores: the 8 bytes per residue are read once, by the MAC, much
// later).  Same thread mapping and LDS exchange as k_fwd_cols.
// ---------------------------------------------------------------------------------------
// Phase-B twiddles of the column pass for one prime (stages 4..LOGR-1 need 16 >> (LOGR - s)
// entries of stage s, starting at (1 << s) + (16 tau >> (LOGR - s))).  They are loaded into
// registers before the previous prime's stores are issued, so waiting for them never waits for
// those stores (phase A's entries are lane-uniform and come through the scalar cache).

template <int LOGR>
struct ColTw {
    static constexpr int count() {
        int c = 0;
        for (int s = 4; s < LOGR; s++) c += 16 >> (LOGR - s);
        return c;
    }
    double w[count() > 0 ? count() : 1];
    __device__ __forceinline__ void load(const double* fw, int tau) {
        int k = 0;
#pragma unroll
        for (int s = 4; s < LOGR; s++) {
#pragma unroll
            for (int i = 0; i < (16 >> (LOGR - s)); i++) w[k++] = tw_load<kTwGlobal>(fw, (1 << s) + ((16 * tau) >> (LOGR - s)) + i);
        }
    }
};

// Phase A runs on digits (|d| < 2^29, far below q) without reducing: its outputs are <= 2.44q
// (stage bounds 0.5, 1.07, 1.71, 2.44; see mac_phase_a for the per-stage rule), so phase B reduces
// only its first stage's 8 lower inputs (then 1.32, 2.0, 2.77, 3.63) before the final reduction.
template <int LOGR>
__device__ __forceinline__ void col_phase_b(double (&v)[16], const ColTw<LOGR>& t, double q, double qinv) {
    int k = 0;
#pragma unroll
    for (int s = 4; s < LOGR; s++) {
        const int half = (1 << LOGR) >> (s + 1);
        if (s == 4) {
#pragma unroll
            for (int e = 0; e < 16; e++)
                if ((e & half) == 0) v[e] = reduce(v[e], q, qinv);
        }
#pragma unroll
        for (int e = 0; e < 16; e++) {
            if ((e & half) == 0) ct_bfly(v[e], v[e + half], t.w[k + (e >> (LOGR - s))], q, qinv);
        }
        k += 16 >> (LOGR - s);
    }
    reduce_all(v, q, qinv);
}


template <int LOGR>
__global__ __launch_bounds__(256) void k_digit_cols(const int32_t* __restrict__ digits, double* __restrict__ out,
                                                    int g, int jper, int logn, PrimeMap pm,
                                                    const DevPrime* __restrict__ primes) {
    constexpr int R = 1 << LOGR, TPS = R / 16, S = 256 / TPS;
    __shared__ double lds[R * S];
    const int n = 1 << logn, N2 = n >> LOGR;
    const int tid = threadIdx.x, col = tid % S, tau = tid / S;
    // S = 16 int32 columns are half a 128-byte line: column blocks 2k and 2k + 1 are given to workgroups
    // 8 apart in dispatch order, which run on the same XCD (workgroups go round-robin over the 8), so
    // each line is fetched into one L2 once
    int cb = blockIdx.x;
    if (S == 16 && (gridDim.x & 15) == 0) cb = (cb & ~15) | ((cb & 7) << 1) | ((cb >> 3) & 1);
    const int r = cb * S + col;
    const long long z = blockIdx.z;
    const int32_t* src = digits + z * n + r;
    int32_t dv[16];
#pragma unroll
    for (int e = 0; e < 16; e++) dv[e] = src[(long long)N2 * (tau + TPS * e)];
    const int jb = blockIdx.y * jper, je = min(g, jb + jper);
    double* dst = out + (z * g + jb) * n;  // wave-uniform base
    const int lane = r + N2 * 16 * tau;      // 32-bit lane offset of row 16 tau
    ColTw<LOGR> tw;
    tw.load(primes[pm(jb)].fw, tau);
#pragma unroll 1
    for (int jl = jb; jl < je; jl++, dst += n) {
        const DevPrime P = primes[pm(jl)];
        const double q = P.q, qinv = P.qinv;
        double v[16];
#pragma unroll
        for (int e = 0; e < 16; e++) {
            // keep the digits as 16 int registers across the prime loop (converted per prime)
            int32_t t = dv[e];
            asm volatile("" : "+v"(t));
            v[e] = (double)t;
        }
        fwd_phase_a<kTwScalar, false>(v, P.fw, 1, q, qinv);
        lds_barrier();  // the previous prime's LDS reads are done
#pragma unroll
        for (int e = 0; e < 16; e++) lds[(tau + TPS * e) * S + col] = v[e];
        lds_barrier();
#pragma unroll
        for (int e = 0; e < 16; e++) v[e] = lds[(16 * tau + e) * S + col];
        col_phase_b<LOGR>(v, tw, q, qinv);
        if (jl + 1 < je) tw.load(primes[pm(jl + 1)].fw, tau);
#pragma unroll
        for (int e = 0; e < 16; e++) __builtin_nontemporal_store(v[e], &(dst + (long long)N2 * e)[lane]);
    }
}

template <int LOGR>
static void digit_cols_t(const Context& c, const int32_t* digits, double* out, int g, int polys, PrimeMap pm,
                         hipStream_t s) {
    constexpr int S = 256 / ((1 << LOGR) / 16);
    const int nx = (1 << (c.logn - LOGR)) / S;
    // primes per workgroup: as many as keeps >= 8 workgroups per CU of the 256 in flight
    const long long base = (long long)nx * polys;
    int jper = (int)std::max(1LL, std::min<long long>(g, base * g / 2048));
    const int ny = (g + jper - 1) / jper;
    jper = (g + ny - 1) / ny;
    ProfScope ps(s, "digit_cols");
    k_digit_cols<LOGR><<<dim3(nx, ny, polys), 256, 0, s>>>(digits, out, g, jper, c.logn, pm, c.d_primes);
    RCK_CHECK_HIP(hipGetLastError());
}

void launch_digit_cols(const Context& c, const int32_t* digits, double* out, int g, int polys, PrimeMap pm,
                       hipStream_t s) {
    int n1, n2;
    split_logn((int)c.logn, &n1, &n2);
    switch (n1) {
        case 6: digit_cols_t<6>(c, digits, out, g, polys, pm, s); break;
        case 7: digit_cols_t<7>(c, digits, out, g, polys, pm, s); break;
        case 8: digit_cols_t<8>(c, digits, out, g, polys, pm, s); break;
        default: RCK_REQUIRE(false, "unsupported ring degree");
    }
}

// ---------------------------------------------------------------------------------------
// Digit MAC: the row pass of the forward NTT of every digit, fused with the multiply-
// accumulate against the NTT-form key.  The column pass ran before into
// cols[b][i][jl][N] for the eval limbs j0 .. j0+G-1 of this group.
//
// Work item = (row block, eval limb, digit split, ciphertext b).  The 1-D grid is laid out
// so that the B work items that read the same key tile get consecutive dispatch slots on
// the same XCD (workgroup w runs on XCD w % 8): the key rows are fetched from HBM once and
// served to the other ciphertexts of the batch from that XCD's L2.
// Output: partial sums part[s][b][comp][j][N] per digit split s (reduced doubles).
// ---------------------------------------------------------------------------------------
// PF (load schedule): 0 = key rows loaded after the NTT stages; 1 = key rows of digit i issued
// before its NTT stages; 2 = the column-pass rows of digit i+1 issued before digit i's stages
// (software pipelining), key rows late; 3 = key rows after the stages in four groups of four
// positions; 4 = key rows issued between the LDS exchange and the second radix-16 phase.
template <int LOGR, int ROWS, int PF>
__global__ __launch_bounds__(ROWS*(1 << LOGR) / 16) void k_digit_mac(const double* __restrict__ cols, int D, int G,
                                                                    int B, int dsplit, int ntiles,
                                                                    const double* __restrict__ key, int E, int j0,
                                                                    double* __restrict__ part, int logn, PrimeMap pm,
                                                                    const DevPrime* __restrict__ primes) {
    constexpr int R = 1 << LOGR, TPS = R / 16;
    __shared__ double lds[ROWS * row_pitch<R>()];
    const int n = 1 << logn, N1 = n >> LOGR, nrb = N1 / ROWS;
    const int g = blockIdx.x, xcd = g & 7, slot = g >> 3;
    const int cb = slot % B;
    const int tile = (slot / B) * 8 + xcd;
    if (tile >= ntiles) return;
    const int rb = tile % nrb, rest = tile / nrb, jl = rest % G, sp = rest / G;
    const int j = j0 + jl;
    const int i0 = (int)((long long)D * sp / dsplit), i1 = (int)((long long)D * (sp + 1) / dsplit);
    const int tid = threadIdx.x, tau = tid % TPS, row = tid / TPS;
    const int b = rb * ROWS + row;
    const DevPrime P = primes[pm(j)];
    const double q = P.q, qinv = P.qinv;
    const long long off = (long long)b * R;
    const long long EN = (long long)E * n;
    const long long col_stride = (long long)G * n;  // digit i -> i + 1 in cols[b][i][jl][N]
    const double* pc = cols + (((long long)cb * D + i0) * G + jl) * n + off + tau;
    // Schedules that keep loads in flight across the NTT stages (1: keys, 2: next digit's rows)
    // read twiddles as global loads and use LDS-only barriers, so neither waits for those loads;
    // the others consume every load at once and keep the generic path (measured faster for them).
    constexpr bool kInFlight = PF == 1 || PF == 2 || PF == 4;
    constexpr int TWM = kInFlight ? kTwGlobal : kTwFlat;
    auto sync = [] {
        if constexpr (kInFlight)
            lds_barrier();
        else
            __syncthreads();
    };
    double a0[16], a1[16];
#pragma unroll
    for (int e = 0; e < 16; e++) a0[e] = a1[e] = 0.0;
    double nx[PF == 2 ? 16 : 1];
    if constexpr (PF == 2) {
#pragma unroll
        for (int e = 0; e < 16; e++) nx[e] = pc[TPS * e];
    }
    for (int i = i0; i < i1; i++, pc += col_stride) {
        const double* k0 = key + (2LL * i) * EN + (long long)j * n + off + tau;
        const double* k1 = k0 + EN;
        double v[16], w0[PF == 1 || PF == 4 ? 16 : 1], w1[PF == 1 || PF == 4 ? 16 : 1];
        if constexpr (PF == 2) {
#pragma unroll
            for (int e = 0; e < 16; e++) v[e] = nx[e];
            if (i + 1 < i1) {
#pragma unroll
                for (int e = 0; e < 16; e++) nx[e] = pc[col_stride + TPS * e];
            }
        } else {
#pragma unroll
            for (int e = 0; e < 16; e++) v[e] = pc[TPS * e];
        }
        if constexpr (PF == 1) {
#pragma unroll
            for (int e = 0; e < 16; e++) {
                w0[e] = k0[e * TPS];
                w1[e] = k1[e * TPS];
            }
        }
        fwd_phase_a<TWM>(v, P.fw, N1 + b, q, qinv);
        sync();
#pragma unroll
        for (int e = 0; e < 16; e++) lds[row_lds<R>(row, tau + TPS * e)] = v[e];
        sync();
#pragma unroll
        for (int e = 0; e < 16; e++) v[e] = lds[row_lds<R>(row, 16 * tau + e)];
        if constexpr (PF == 4) {
#pragma unroll
            for (int e = 0; e < 16; e++) {
                w0[e] = k0[e * TPS];
                w1[e] = k1[e * TPS];
            }
        }
        fwd_phase_b<LOGR, TWM>(v, P.fw, N1 + b, tau, q, qinv);
        if constexpr (PF == 3) {
            // key rows in four groups of four positions: at most 8 key values in flight per thread
            // (fewer live registers, more resident waves)
#pragma unroll
            for (int g4 = 0; g4 < 4; g4++) {
                double x0[4], x1[4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    x0[e] = k0[(4 * g4 + e) * TPS];
                    x1[e] = k1[(4 * g4 + e) * TPS];
                }
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    a0[4 * g4 + e] += mulmod(v[4 * g4 + e], x0[e], q, qinv);
                    a1[4 * g4 + e] += mulmod(v[4 * g4 + e], x1[e], q, qinv);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
#pragma unroll
            for (int e = 0; e < 16; e++) {
                if constexpr (PF == 1 || PF == 4) {
                    a0[e] += mulmod(v[e], w0[e], q, qinv);
                    a1[e] += mulmod(v[e], w1[e], q, qinv);
                } else {
                    a0[e] += mulmod(v[e], k0[e * TPS], q, qinv);
                    a1[e] += mulmod(v[e], k1[e * TPS], q, qinv);
                }
            }
        }
        // reduced rows times reduced keys: products <= 0.57q, so eight of them on a reduced
        // accumulator stay below 5.1q < 2^53 / q
        if (((i - i0) & 7) == 7) {
            reduce_all(a0, q, qinv);
            reduce_all(a1, q, qinv);
        }
    }
    reduce_all(a0, q, qinv);
    reduce_all(a1, q, qinv);
    double* o0 = part + (((long long)sp * B + cb) * 2) * EN + (long long)j * n + off + tau;
    double* o1 = o0 + EN;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        o0[e * TPS] = a0[e];
        o1[e * TPS] = a1[e];
    }
}

// Streaming variant: a workgroup is four waves on one tile (ring rows b0 .. b0 + ROWS - 1 of one
// eval limb, ROWS = 64 / (R / 16) rows per wave): one ciphertext per wave, or (SPLIT) one
// ciphertext and a quarter of its digit range per wave.
//   * the column-pass rows of digit i + 1 are copied HBM -> LDS by global_load_lds while digit i
//     is transformed (no registers held; the copy is waited for when digit i + 1 starts);
//   * the row-pass twiddles of the tile are staged in LDS once and shared by the four waves;
//   * the key rows are read in two groups after the second phase (with one ciphertext per wave the
//     four waves read the same rows, served by the CU's L1 and L2).  Vector loads complete in issue
//     order, so the copy is issued once the exchange has been read and before the key loads: waiting
//     for the keys then includes the copy, which by then has had the whole second phase to land.
// Each wave's LDS buffer holds its rows at the padded pitch of row_lds (R + R/16): the copy
// lands a row linearly, phase A reads it linearly and writes its outputs back in place at the
// padded positions, which phase B reads without bank conflicts.

// s_waitcnt immediates (vmcnt [3:0]+[15:14], expcnt [6:4], lgkmcnt [11:8]; the others at maximum)
constexpr int kWaitVm0 = 0x0F70;
constexpr int kWaitLgkm0 = 0xC07F;

// Row-pass stages with the twiddles staged in LDS in lane-friendly layouts:
//   phase A (stages 0..3, the same 15 entries for every lane of a ring row):
//     twa[k * ROWS + row], k = (1 << s) - 1 + t  (the ROWS rows of one k sit side by side);
//   phase B (stages 4..LOGR-1, 16 >> (LOGR - s) entries of stage s per lane):
//     twb[k * 64 + lane], k = off_s + t  (the 64 lanes of one k sit side by side).
// Entry t of stage s for ring row b is fw[((N1 + b) << s) + t'] with t' = t (phase A) or
// t' = (16 tau >> (LOGR - s)) + t (phase B), as in fwd_phase_a / fwd_phase_b.
template <int LOGR>
struct MacTw {
    static constexpr int R = 1 << LOGR, TPS = R / 16, ROWS = 64 / TPS;
    static constexpr int countb() {
        int c = 0;
        for (int s = 4; s < LOGR; s++) c += 16 >> (LOGR - s);
        return c;
    }
    static constexpr int SIZE = 15 * ROWS + countb() * 64;
};

// Value bounds of the row pass (multiples of q).  With |w| <= q/2 + 1 and |a| <= A q, A <= 3.75,
// mulmod returns |a w - k q| <= (0.5 + 0.133 A) q: the quotient estimate is off by at most
// |a w / q| 2^-52 <= A q 2^-53 < 0.133 A for q < 2^50 + 2^46 (modarith.hpp), and A <= 3.75 keeps
// |a w / q| below 2^51.  A butterfly whose lower input is bounded by L and upper input by U
// outputs at most L + 0.5 + 0.133 U.  Phase A takes the column pass's reduced outputs
// (|x| <= q/2 + 1) through four stages without reducing: 0.5 -> 1.067 -> 1.709 -> 2.436 -> 3.26.
template <int LOGR>
__device__ __forceinline__ void mac_phase_a(double (&v)[16], const double* twa, double q, double qinv) {
    constexpr int ROWS = MacTw<LOGR>::ROWS;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int half = 8 >> s;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            if ((e & half) == 0) ct_bfly(v[e], v[e + half], twa[((1 << s) - 1 + (e >> (4 - s))) * ROWS], q, qinv);
        }
    }
}

// Phase B starts from |v| <= 3.26q.  Every output is a chain of lower butterfly inputs plus one
// twiddle product per stage, so only the chains' lower inputs are reduced: before the first stage
// its 8 lower inputs (every value is then <= 0.5 + (0.5 + 0.133 * 3.26) = 1.433 after it, then
// 2.124 after the second stage), and with four stages, before the third stage the 4 lower inputs
// that feed the last stage's lower inputs (v[e], e % 4 == 0): those pairs give <= 1.282, the other
// pairs <= 2.906, and the last stage <= 1.282 + 0.5 + 0.133 * 2.906 = 2.17.  With three stages the
// last one gives <= 2.906.  Multiplied inputs stay <= 3.26 and the outputs feed the key products
// (|a| <= 3.75q) directly: 12 reductions instead of 16.
template <int LOGR>
__device__ __forceinline__ void mac_phase_b(double (&v)[16], const double* twb, double q, double qinv) {
    constexpr int R = 1 << LOGR;
    int off = 0;
#pragma unroll
    for (int s = 4; s < LOGR; s++) {
        const int half = R >> (s + 1), m = LOGR - s;
        if (s == 4) {
#pragma unroll
            for (int e = 0; e < 16; e++)
                if ((e & half) == 0) v[e] = reduce(v[e], q, qinv);
        }
        if (s == 6 && LOGR == 8) {
#pragma unroll
            for (int e = 0; e < 16; e += 4) v[e] = reduce(v[e], q, qinv);
        }
#pragma unroll
        for (int e = 0; e < 16; e++) {
            if ((e & half) == 0) ct_bfly(v[e], v[e + half], twb[(off + (e >> m)) * 64], q, qinv);
        }
        off += 16 >> m;
    }
}

// INV (single digit range): the epilogue also runs the inverse row stages of both accumulators (after
// adding P * fold[cb][comp] on the Q limbs, when fold is given) and stores them in coefficient row
// order, so the accumulators go straight to the inverse column pass.
// SPLIT: the four waves take one ciphertext and four consecutive parts of its digit range (for small
// batches: four times the waves, the staged twiddles still shared); the four partial accumulators
// are summed through LDS and waves 0 and 1 finish components 0 and 1.
template <int LOGR, bool INV, bool SPLIT>
__global__ __launch_bounds__(256) void k_digit_mac_dma(const double* __restrict__ cols, int D, int G, int B,
                                                       int nbq, int dsplit, int ntiles,
                                                       const double* __restrict__ key, int E, int j0,
                                                       double* __restrict__ part, int logn, PrimeMap pm,
                                                       const DevPrime* __restrict__ primes, MacFold fold) {
    constexpr int R = 1 << LOGR, TPS = R / 16, ROWS = 64 / TPS, PITCH = row_pitch<R>();
    // copy granule per lane: 16 bytes (1 KiB per instruction) when a row is at least 1 KiB, else 4
    constexpr int SZ = R * 8 >= 1024 ? 16 : 4, PER_ROW = R * 8 / (64 * SZ);
    __shared__ double twl[MacTw<LOGR>::SIZE];  // [phase A: 15 x ROWS][phase B: countb x 64]
    __shared__ double bufl[4][ROWS * PITCH];   // per wave: the digit's rows
    const int n = 1 << logn, N1 = n >> LOGR, nrb = N1 / ROWS;
    const int g = blockIdx.x, xcd = g & 7, slot = g >> 3;
    const int cq = slot % nbq;
    const int tile = (slot / nbq) * 8 + xcd;
    if (tile >= ntiles) return;
    const int rb = tile % nrb, rest = tile / nrb, jl = rest % G, sp = rest / G;
    const int j = j0 + jl;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r0 = (int)((long long)D * sp / dsplit), r1 = (int)((long long)D * (sp + 1) / dsplit);
    const int i0 = SPLIT ? r0 + (r1 - r0) * wave / 4 : r0, i1 = SPLIT ? r0 + (r1 - r0) * (wave + 1) / 4 : r1;
    const int tau = lane % TPS, row = lane / TPS;
    const int b0 = rb * ROWS, b = b0 + row;
    const DevPrime P = primes[pm(j)];
    const double q = P.q, qinv = P.qinv;
    const int cb = SPLIT ? cq : cq * 4 + wave;
    const bool active = cb < B && i0 < i1;
    const long long EN = (long long)E * n;
    const long long col_stride = (long long)G * n;
    double* buf = bufl[wave];
    const double* tile0 = cols + (((long long)(active ? cb : 0) * D + i0) * G + jl) * n + (long long)b0 * R;
    auto dma = [&](const double* src) {
#pragma unroll
        for (int r = 0; r < ROWS; r++) {
#pragma unroll
            for (int h = 0; h < PER_ROW; h++) {
                const char* gsrc = reinterpret_cast<const char*>(src + r * R) + (h * 64 + lane) * SZ;
                char* ldst = reinterpret_cast<char*>(buf + r * PITCH) + h * 64 * SZ;
                if constexpr (SZ == 16)
                    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)gsrc,
                                                     (__attribute__((address_space(3))) void*)ldst, 16, 0, 0);
                else
                    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)gsrc,
                                                     (__attribute__((address_space(3))) void*)ldst, 4, 0, 0);
            }
        }
    };
    if (active) dma(tile0);
    // stage the tile's twiddles (layouts at MacTw)
    for (int t = tid; t < MacTw<LOGR>::SIZE; t += 256) {
        int idx;
        if (t < 15 * ROWS) {
            const int k = t / ROWS, r = t % ROWS, st = 31 - __builtin_clz(k + 1);
            idx = ((N1 + b0 + r) << st) + (k + 1 - (1 << st));
        } else {
            const int u = t - 15 * ROWS, k = u / 64, ln = u % 64, r = ln / TPS, ta = ln % TPS;
            int st = 4, off = 0;
            while (k >= off + (16 >> (LOGR - st))) {
                off += 16 >> (LOGR - st);
                st++;
            }
            idx = ((N1 + b0 + r) << st) + ((16 * ta) >> (LOGR - st)) + (k - off);
        }
        twl[t] = tw_load<kTwGlobal>(P.fw, idx);
    }
    __syncthreads();
    if (!SPLIT && !active) return;
    const double* twa = twl + row;
    const double* twb = twl + 15 * ROWS + lane;
    double a0[16], a1[16];
#pragma unroll
    for (int e = 0; e < 16; e++) a0[e] = a1[e] = 0.0;
    const double* src = tile0;
    int pend = 0;  // products added since the accumulators were last reduced
    for (int i = i0; i < i1; i++, src += col_stride) {
        __builtin_amdgcn_s_waitcnt(kWaitVm0);  // this digit's rows are in LDS
        asm volatile("" ::: "memory");
        double v[16];
#pragma unroll
        for (int e = 0; e < 16; e++) v[e] = buf[row * PITCH + tau + TPS * e];
        mac_phase_a<LOGR>(v, twa, q, qinv);
#pragma unroll
        for (int e = 0; e < 16; e++) buf[row_lds<R>(row, tau + TPS * e)] = v[e];
#pragma unroll
        for (int e = 0; e < 16; e++) v[e] = buf[row_lds<R>(row, 16 * tau + e)];
        __builtin_amdgcn_s_waitcnt(kWaitLgkm0);  // every read of the buffer has completed
        asm volatile("" ::: "memory");
        if (i + 1 < i1) dma(src + col_stride);
        __builtin_amdgcn_sched_barrier(0);
        mac_phase_b<LOGR>(v, twb, q, qinv);
        // key rows in two groups of eight positions (issued after the copy: by the time they are
        // waited for, the copy has had the whole second phase to land)
        const double* k0 = key + (2LL * i) * EN + (long long)j * n + (long long)b * R + tau;
        const double* k1 = k0 + EN;
#pragma unroll
        for (int g8 = 0; g8 < 2; g8++) {
            double x0[8], x1[8];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                x0[e] = k0[(8 * g8 + e) * TPS];
                x1[e] = k1[(8 * g8 + e) * TPS];
            }
#pragma unroll
            for (int e = 0; e < 8; e++) {
                a0[8 * g8 + e] += mulmod(v[8 * g8 + e], x0[e], q, qinv);
                a1[8 * g8 + e] += mulmod(v[8 * g8 + e], x1[e], q, qinv);
            }
        }
        // mac_phase_b leaves |v| <= 2.17q (four stages), 2.906q (three) or 2.124q (two), so a product
        // is <= 0.79q, 0.89q or 0.79q; eight of them (seven with the three-stage row pass, LOGR = 7, of
        // N = 2^13 and 2^14) on a reduced accumulator stay below 6.81q (6.71q) < 2^53 / q = 7.52q for
        // q < 2^50 + 2^46: exact integers
        constexpr int kPend = LOGR == 7 ? 7 : 8;
        if (++pend == kPend) {
            pend = 0;
            reduce_all(a0, q, qinv);
            reduce_all(a1, q, qinv);
        }
    }
    reduce_all(a0, q, qinv);
    reduce_all(a1, q, qinv);
    double* o0 = part + (((long long)sp * B + cb) * 2) * EN + (long long)j * n + (long long)b * R + tau;
    // one component's finish: optional fold + inverse row stages (INV), then the store
    auto finish = [&](int comp, double(&acc)[16]) {
        double* o = o0 + comp * EN;
        if constexpr (INV) {
            if (fold.addn != nullptr && j < fold.nq) {
                const double pmq = fold.pmodq[j];
                const double* an = fold.addn + cb * fold.bs + comp * fold.cs + (long long)j * n + (long long)b * R + tau;
#pragma unroll
                for (int e = 0; e < 16; e++) acc[e] = reduce(acc[e] + mulmod(an[e * TPS], pmq, q, qinv), q, qinv);
            }
            inv_phase_b<LOGR>(acc, P.iv, N1 + b, tau, q, qinv);
#pragma unroll
            for (int e = 0; e < 16; e++) buf[row_lds<R>(row, 16 * tau + e)] = acc[e];
#pragma unroll
            for (int e = 0; e < 16; e++) acc[e] = buf[row_lds<R>(row, tau + TPS * e)];
            inv_phase_a(acc, P.iv, N1 + b, q, qinv);
#pragma unroll
            for (int e = 0; e < 16; e++) o[TPS * e] = acc[e];
            __builtin_amdgcn_s_waitcnt(kWaitLgkm0);  // the buffer's reads are done before it is reused
            asm volatile("" ::: "memory");
        } else {
#pragma unroll
            for (int e = 0; e < 16; e++) o[e * TPS] = acc[e];
        }
    };
    if constexpr (SPLIT) {
        // sum the four waves' partial accumulators: component k goes through the buffers, wave k sums it
        // (|sum| <= 4 * 0.51q), then waves 0 and 1 finish their component in their own buffer
#pragma unroll
        for (int comp = 0; comp < 2; comp++) {
            double(&acc)[16] = comp == 0 ? a0 : a1;
#pragma unroll
            for (int e = 0; e < 16; e++) buf[e * 64 + lane] = acc[e];
            __syncthreads();
            if (wave == comp) {
#pragma unroll
                for (int e = 0; e < 16; e++)
                    acc[e] = reduce(bufl[0][e * 64 + lane] + bufl[1][e * 64 + lane] + bufl[2][e * 64 + lane] +
                                        bufl[3][e * 64 + lane],
                                    q, qinv);
            }
            __syncthreads();
        }
        if (wave == 0) finish(0, a0);
        if (wave == 1) finish(1, a1);
    } else {
        finish(0, a0);
        finish(1, a1);
    }
}

// Narrow variant: 8 elements per thread in the row pass (R/8 threads per ring row), so the
// accumulators, the butterfly operands and the key values take half the registers of the 16-wide
// kernel and more waves stay resident.  Stages of a 2^LOGR-point row transform:
//   phase A: x = tau + (R/8) e            stages 0..2   (row bits LOGR-1 .. LOGR-3)
//   phase B: x = hi(tau) + (e << m) + lo  stages 3..5   (bits LOGR-4 .. LOGR-6), m = LOGR - 6
//   phase C: x = 8 tau + e                stages 6..    (bits m-1 .. 0)
// with an LDS exchange between phases.  Twiddle of stage s for the pair starting at x:
// tbl[(base << s) + (x >> (LOGR - s))], as in the 16-wide phases; outputs land in the same
// NTT-domain positions (ntt_phys), so keys and accumulators are shared with the other kernels.
template <int LOGR, int ROWS>
__global__ __launch_bounds__(ROWS*(1 << LOGR) / 8) void k_digit_mac8(const double* __restrict__ cols, int D, int G,
                                                                     int B, int dsplit, int ntiles,
                                                                     const double* __restrict__ key, int E, int j0,
                                                                     double* __restrict__ part, int logn, PrimeMap pm,
                                                                     const DevPrime* __restrict__ primes) {
    constexpr int R = 1 << LOGR, TPS = R / 8, T16 = R / 16, M = LOGR - 6;
    static_assert(LOGR >= 6 && LOGR <= 8, "row length 64 .. 256");
    __shared__ double lds[ROWS * row_pitch<R>()];
    const int n = 1 << logn, N1 = n >> LOGR, nrb = N1 / ROWS;
    const int g = blockIdx.x, xcd = g & 7, slot = g >> 3;
    const int cb = slot % B;
    const int tile = (slot / B) * 8 + xcd;
    if (tile >= ntiles) return;
    const int rb = tile % nrb, rest = tile / nrb, jl = rest % G, sp = rest / G;
    const int j = j0 + jl;
    const int i0 = (int)((long long)D * sp / dsplit), i1 = (int)((long long)D * (sp + 1) / dsplit);
    const int tid = threadIdx.x, tau = tid % TPS, row = tid / TPS;
    const int b = rb * ROWS + row;
    const int base = N1 + b;
    const DevPrime P = primes[pm(j)];
    const double q = P.q, qinv = P.qinv;
    const double* __restrict__ tbl = P.fw;
    const long long off = (long long)b * R;
    const long long EN = (long long)E * n;
    const long long col_stride = (long long)G * n;
    const double* pc = cols + (((long long)cb * D + i0) * G + jl) * n + off + tau;
    // phase B positions: hi(tau) + (e << M) + lo(tau);  phase C / output positions: 8 tau + e
    const int xb0 = (tau >> M) * (8 << M) + (tau & ((1 << M) - 1));
    // output position of element e: x = 8 tau + e stored at (x % 16) * T16 + x / 16
    const int ph0 = (8 * (tau & 1)) * T16 + (tau >> 1);
    double a0[8], a1[8];
#pragma unroll
    for (int e = 0; e < 8; e++) a0[e] = a1[e] = 0.0;
    for (int i = i0; i < i1; i++, pc += col_stride) {
        double v[8];
#pragma unroll
        for (int e = 0; e < 8; e++) v[e] = pc[TPS * e];
        // phase A: stages 0..2, x = tau + TPS e
#pragma unroll
        for (int st = 0; st < 3; st++) {
            const int half = 4 >> st;
#pragma unroll
            for (int e = 0; e < 8; e++)
                if ((e & half) == 0) ct_bfly(v[e], v[e + half], tw_load<kTwFlat>(tbl, (base << st) + (e >> (3 - st))), q, qinv);
        }
#pragma unroll
        for (int e = 0; e < 8; e++) v[e] = reduce(v[e], q, qinv);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 8; e++) lds[row_lds<R>(row, tau + TPS * e)] = v[e];
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 8; e++) v[e] = lds[row_lds<R>(row, xb0 + (e << M))];
        // phase B: stages 3..5
#pragma unroll
        for (int st = 3; st < 6; st++) {
            const int half = 4 >> (st - 3);
#pragma unroll
            for (int e = 0; e < 8; e++)
                if ((e & half) == 0)
                    ct_bfly(v[e], v[e + half], tw_load<kTwFlat>(tbl, (base << st) + ((xb0 + (e << M)) >> (LOGR - st))), q, qinv);
        }
#pragma unroll
        for (int e = 0; e < 8; e++) v[e] = reduce(v[e], q, qinv);
        if constexpr (M > 0) {
            __syncthreads();
#pragma unroll
            for (int e = 0; e < 8; e++) lds[row_lds<R>(row, xb0 + (e << M))] = v[e];
            __syncthreads();
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = lds[row_lds<R>(row, 8 * tau + e)];
            // phase C: stages 6 .. LOGR-1 on x = 8 tau + e
#pragma unroll
            for (int st = 6; st < LOGR; st++) {
                const int half = (1 << LOGR) >> (st + 1);
#pragma unroll
                for (int e = 0; e < 8; e++)
                    if ((e & half) == 0)
                        ct_bfly(v[e], v[e + half], tw_load<kTwFlat>(tbl, (base << st) + ((8 * tau + e) >> (LOGR - st))), q, qinv);
            }
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = reduce(v[e], q, qinv);
        }
        const double* k0 = key + (2LL * i) * EN + (long long)j * n + off + ph0;
        const double* k1 = k0 + EN;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            a0[e] += mulmod(v[e], k0[e * T16], q, qinv);
            a1[e] += mulmod(v[e], k1[e * T16], q, qinv);
        }
        if (((i - i0) & 7) == 7) {  // products <= 0.57q (reduced operands): eight per reduction
#pragma unroll
            for (int e = 0; e < 8; e++) {
                a0[e] = reduce(a0[e], q, qinv);
                a1[e] = reduce(a1[e], q, qinv);
            }
        }
    }
    double* o0 = part + (((long long)sp * B + cb) * 2) * EN + (long long)j * n + off + ph0;
    double* o1 = o0 + EN;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        o0[e * T16] = reduce(a0[e], q, qinv);
        o1[e * T16] = reduce(a1[e], q, qinv);
    }
}

template <int LOGR, int ROWS>
static void mac8_t(const Context& c, const double* cols, int D, int G, int B, int dsplit, const double* key, int E,
                   int j0, double* part, PrimeMap pm, hipStream_t s) {
    constexpr int TPS = (1 << LOGR) / 8;
    const int nrb = (1 << (c.logn - LOGR)) / ROWS;
    const int ntiles = nrb * G * dsplit;
    const int blocks = ((ntiles + 7) / 8) * 8 * B;
    ProfScope ps(s, "digit_mac");
    k_digit_mac8<LOGR, ROWS><<<dim3(blocks), ROWS * TPS, 0, s>>>(cols, D, G, B, dsplit, ntiles, key, E, j0, part,
                                                                 c.logn, pm, c.d_primes);
    RCK_CHECK_HIP(hipGetLastError());
}

template <int LOGR, bool SPLIT>
static void mac_dma_t(const Context& c, const double* cols, int D, int G, int B, int dsplit, const double* key, int E,
                      int j0, double* part, PrimeMap pm, hipStream_t s, const MacFold* inv) {
    constexpr int ROWS = 64 / ((1 << LOGR) / 16);
    const int nrb = (1 << (c.logn - LOGR)) / ROWS;
    const int ntiles = nrb * G * dsplit;
    const int nbq = SPLIT ? B : (B + 3) / 4;
    const int blocks = ((ntiles + 7) / 8) * 8 * nbq;
    ProfScope ps(s, "digit_mac");
    if (inv) {
        RCK_REQUIRE(dsplit == 1, "the fused inverse rows need a single digit range");
        k_digit_mac_dma<LOGR, true, SPLIT><<<dim3(blocks), 256, 0, s>>>(cols, D, G, B, nbq, dsplit, ntiles, key, E, j0,
                                                                        part, c.logn, pm, c.d_primes, *inv);
    } else {
        k_digit_mac_dma<LOGR, false, SPLIT><<<dim3(blocks), 256, 0, s>>>(cols, D, G, B, nbq, dsplit, ntiles, key, E,
                                                                         j0, part, c.logn, pm, c.d_primes, MacFold{});
    }
    RCK_CHECK_HIP(hipGetLastError());
}

template <int LOGR, int ROWS, int PF>
static void mac_t(const Context& c, const double* cols, int D, int G, int B, int dsplit, const double* key, int E,
                  int j0, double* part, PrimeMap pm, hipStream_t s) {
    constexpr int TPS = (1 << LOGR) / 16;
    const int nrb = (1 << (c.logn - LOGR)) / ROWS;
    const int ntiles = nrb * G * dsplit;
    const int blocks = ((ntiles + 7) / 8) * 8 * B;
    ProfScope ps(s, "digit_mac");
    k_digit_mac<LOGR, ROWS, PF><<<dim3(blocks), ROWS * TPS, 0, s>>>(cols, D, G, B, dsplit, ntiles, key, E, j0,
                                                                   part, c.logn, pm, c.d_primes);
    RCK_CHECK_HIP(hipGetLastError());
}

template <int LOGR, int ROWS>
static void mac_r(const Context& c, int pf, const double* cols, int D, int G, int B, int dsplit, const double* key,
                  int E, int j0, double* part, PrimeMap pm, hipStream_t s, const MacFold* inv) {
    RCK_REQUIRE(inv == nullptr || pf == 5 || pf == 6, "the fused inverse rows need the streamed MAC");
    if (pf == 5)
        mac_dma_t<LOGR, false>(c, cols, D, G, B, dsplit, key, E, j0, part, pm, s, inv);
    else if (pf == 6)
        mac_dma_t<LOGR, true>(c, cols, D, G, B, dsplit, key, E, j0, part, pm, s, inv);
    else if (pf == 1)
        mac_t<LOGR, ROWS, 1>(c, cols, D, G, B, dsplit, key, E, j0, part, pm, s);
    else if (pf == 4)
        mac_t<LOGR, ROWS, 4>(c, cols, D, G, B, dsplit, key, E, j0, part, pm, s);
    else if (pf == 3)
        mac_t<LOGR, ROWS, 3>(c, cols, D, G, B, dsplit, key, E, j0, part, pm, s);
    else if (pf == 2)
        mac_t<LOGR, ROWS, 2>(c, cols, D, G, B, dsplit, key, E, j0, part, pm, s);
    else
        mac_t<LOGR, ROWS, 0>(c, cols, D, G, B, dsplit, key, E, j0, part, pm, s);
}

template <int LOGR>
static void mac_v(const Context& c, int rows, int pf, int width, const double* cols, int D, int G, int B, int dsplit,
                  const double* key, int E, int j0, double* part, PrimeMap pm, hipStream_t s, const MacFold* inv) {
    RCK_REQUIRE(inv == nullptr || (rows <= 4 && width == 16), "the fused inverse rows need the streamed MAC");
    if (width == 8) {
        if (rows <= 8)
            mac8_t<LOGR, 8>(c, cols, D, G, B, dsplit, key, E, j0, part, pm, s);
        else
            mac8_t<LOGR, 16>(c, cols, D, G, B, dsplit, key, E, j0, part, pm, s);
        return;
    }
    if (rows <= 4)
        mac_r<LOGR, 4>(c, pf, cols, D, G, B, dsplit, key, E, j0, part, pm, s, inv);
    else
        mac_r<LOGR, 16>(c, pf, cols, D, G, B, dsplit, key, E, j0, part, pm, s, inv);
}

// Small rings: plain forward NTT of every digit for every prime, then pointwise MAC.
// dn: [B][D][E][N] NTT-form digits; part: [B][2][E][N].
__global__ void k_mac_small(const double* __restrict__ dn, int D, const double* __restrict__ key, int E, int n,
                            double* __restrict__ part, PrimeMap pm, const DevPrime* __restrict__ primes) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int y = blockIdx.y, cb = blockIdx.z;
    const DevPrime P = primes[pm(y)];
    const long long EN = (long long)E * n;
    double s0 = 0.0, s1 = 0.0;
    for (int d = 0; d < D; d++) {
        const double v = dn[((long long)cb * D + d) * EN + (long long)y * n + i];
        s0 += mulmod(v, key[(2LL * d) * EN + (long long)y * n + i], P.q, P.qinv);
        s1 += mulmod(v, key[(2LL * d + 1) * EN + (long long)y * n + i], P.q, P.qinv);
        if ((d & 7) == 7) {  // products <= 0.57q (reduced operands): eight per reduction
            s0 = reduce(s0, P.q, P.qinv);
            s1 = reduce(s1, P.q, P.qinv);
        }
    }
    part[(2LL * cb) * EN + (long long)y * n + i] = reduce(s0, P.q, P.qinv);
    part[(2LL * cb + 1) * EN + (long long)y * n + i] = reduce(s1, P.q, P.qinv);
}

void launch_digit_mac(const Context& c, const int32_t* digits, int B, int D, const double* key, int nq, int naux,
                      double* part, int dsplit, int group, int rows, int prefetch, int width, hipStream_t s,
                      const MacFold* inv) {
    require_fp64(c);
    const int E = nq + naux;
    const long long n = c.n, EN = (long long)E * n;
    PrimeMap pm{nq, (int)c.L, 0};
    Context& cc = const_cast<Context&>(c);
    if (c.logn <= 11) {
        RCK_REQUIRE(dsplit == 1, "small rings use a single digit split");
        double* dn = static_cast<double*>(cc.workspace(7, (size_t)B * D * EN * 8));
        // every digit row is NTT'd for every prime: limb stride 0 on the input
        launch_fwd(c, kInI32, digits, 0, n, dn, n, EN, E, B * D, pm, s);
        ProfScope ps(s, "mac_small");
        k_mac_small<<<dim3((c.n + 255) / 256, E, B), 256, 0, s>>>(dn, D, key, E, c.n, part, pm, c.d_primes);
        RCK_CHECK_HIP(hipGetLastError());
        return;
    }
    RCK_REQUIRE(dsplit >= 1 && dsplit <= D, "bad digit split");
    const int G = std::max(1, std::min(group, E));
    const int ngroups = (E + G - 1) / G;
    const size_t slot_elems = (size_t)B * D * G * n;
    double* cols_base = static_cast<double*>(cc.workspace(7, (ngroups > 1 ? 2 : 1) * slot_elems * 8));
    int n1, n2;
    split_logn(c.logn, &n1, &n2);
    // With several groups the column pass of group g+1 runs on the context's second stream while
    // the MAC of group g runs on `s` (double-buffered column scratch, event-ordered both ways).
    hipStream_t s2 = ngroups > 1 ? cc.aux_stream() : s;
    if (ngroups > 1) {
        RCK_CHECK_HIP(hipEventRecord(cc.event(0), s));
        RCK_CHECK_HIP(hipStreamWaitEvent(s2, cc.event(0), 0));
    }
    for (int gi = 0, j0 = 0; j0 < E; gi++, j0 += G) {
        const int g = std::min(G, E - j0);
        const int slot = gi & 1;
        double* cols = cols_base + slot * slot_elems;
        if (ngroups > 1 && gi >= 2) RCK_CHECK_HIP(hipStreamWaitEvent(s2, cc.event(1 + slot), 0));
        // column pass of digit (b, i) for eval limbs j0 .. j0+g-1 -> cols[b][i][jl][N]
        const PrimeMap pmg{nq, (int)c.L, j0};
        switch (n1) {
            case 6: digit_cols_t<6>(c, digits, cols, g, B * D, pmg, s2); break;
            case 7: digit_cols_t<7>(c, digits, cols, g, B * D, pmg, s2); break;
            case 8: digit_cols_t<8>(c, digits, cols, g, B * D, pmg, s2); break;
            default: RCK_REQUIRE(false, "unsupported ring degree");
        }
        if (ngroups > 1) {
            RCK_CHECK_HIP(hipEventRecord(cc.event(3 + slot), s2));
            RCK_CHECK_HIP(hipStreamWaitEvent(s, cc.event(3 + slot), 0));
        }
        switch (n2) {
            case 6: mac_v<6>(c, rows, prefetch, width, cols, D, g, B, dsplit, key, E, j0, part, pm, s, inv); break;
            case 7: mac_v<7>(c, rows, prefetch, width, cols, D, g, B, dsplit, key, E, j0, part, pm, s, inv); break;
            case 8: mac_v<8>(c, rows, prefetch, width, cols, D, g, B, dsplit, key, E, j0, part, pm, s, inv); break;
            default: RCK_REQUIRE(false, "unsupported ring degree");
        }
        if (ngroups > 1) RCK_CHECK_HIP(hipEventRecord(cc.event(1 + slot), s));
    }
}

// ---------------------------------------------------------------------------------------
// mod-down (keyswitching.rs:278-322) with the optional relinearization / rotation add.
//   x_P = Garner(acc_P) in [0, P);  c = centred(x_P) mod q_i;  out = (x_i - c) P^{-1} mod q_i.
// One thread per (coefficient, accumulator): the P-part is converted once and applied to every
// Q-limb (coalesced row reads per limb).
// ---------------------------------------------------------------------------------------
__global__ void k_mod_down(const uint64_t* __restrict__ acc, int n, int nq, int K, int L, int T,
                           const uint64_t* __restrict__ add0, const uint64_t* __restrict__ add1,
                           uint64_t* __restrict__ out, long long out_ps, const DevPrime* __restrict__ primes,
                           const double* __restrict__ invq, const uint32_t* __restrict__ pwords,
                           const double* __restrict__ pmodq, const double* __restrict__ pinvq,
                           const double* __restrict__ pmod_aux) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int z = blockIdx.y;
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
    // X_P as 8 words (Horner, one pass with two carry chains), compared with floor(P/2)
    uint32_t xw[8];
#pragma unroll
    for (int w = 0; w < 8; w++) xw[w] = 0;
    {
        const uint64_t t = (uint64_t)m[K - 1];
        xw[0] = (uint32_t)t;
        xw[1] = (uint32_t)(t >> 32);
    }
    for (int k = K - 2; k >= 0; k--) {
        const uint64_t pk = (uint64_t)primes[L + k].q;
        const uint32_t plo = (uint32_t)pk, phi = (uint32_t)(pk >> 32);
        uint64_t c0 = 0, c1 = (uint64_t)m[k];
        uint32_t prev = 0;
#pragma unroll
        for (int w = 0; w < 8; w++) {
            const uint64_t t = (uint64_t)xw[w] * plo + c0;
            const uint64_t u = (t & 0xffffffffull) + (uint64_t)prev * phi + c1;
            prev = xw[w];
            xw[w] = (uint32_t)u;
            c0 = t >> 32;
            c1 = u >> 32;
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
    const uint64_t* add = z == 0 ? add0 : add1;
    uint64_t* o = out + (long long)z * out_ps;
    for (int i = 0; i < nq; i++) {
        // centred(x_P) mod q_i by Horner (evaluate_mixed_radix_mod, keyswitching.rs:412-423)
        const DevPrime Pi = primes[i];
        const double q = Pi.q, qinv = Pi.qinv;
        double cv = reduce(m[K - 1], q, qinv);
        for (int k = K - 2; k >= 0; k--)
            cv = reduce(mulmod(cv, pmod_aux[(long long)k * L + i], q, qinv) + m[k], q, qinv);
        if (gt) cv -= pmodq[i];
        const double xi = (double)a[(long long)i * n + c];
        double r = mulmod(reduce(xi - cv, q, qinv), pinvq[i], q, qinv);
        if (add) r += (double)add[(long long)i * n + c];
        o[(long long)i * n + c] = f64_to_u64(canon(r, q, qinv));
    }
}

// ---------------------------------------------------------------------------------------
// Inverse column pass of the Q-limbs of the keyswitch accumulators fused with mod-down and the
// relinearization / rotation add: the coefficient x_i leaves the inverse NTT in registers, the
// aux residues of the same coefficient (already in coefficient form, accP) are converted to the
// centred x_P mod q_i, and (x_i - c) * P^{-1} (+ addend) is written straight to the output.
// Same arithmetic as k_mod_down; saves writing and re-reading the Q-part of the accumulators.
// ---------------------------------------------------------------------------------------
struct OutPtrs {
    uint64_t* p[32];
};

// Once per coefficient of the aux part: pg[z][k][c] = Garner digit m_k (k < K), pg[z][K][c] = 1 when
// x_P > floor(P/2) else 0.  accP: [Z][K][N] canonical residues.
__global__ void k_p_garner(const uint64_t* __restrict__ accP, double* __restrict__ pg, int n, int K, int L, int T,
                           const DevPrime* __restrict__ primes, const double* __restrict__ invq,
                           const uint32_t* __restrict__ pwords) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int z = blockIdx.y;
    double r[kMaxP], m[kMaxP];
    for (int k = 0; k < K; k++) r[k] = (double)accP[((long long)z * K + k) * n + c];
    const bool gt = p_garner(r, K, L, T, m, primes, invq, pwords);
    double* o = pg + (long long)z * (K + 1) * n + c;
    for (int k = 0; k < K; k++) o[(long long)k * n] = m[k];
    o[(long long)K * n] = gt ? 1.0 : 0.0;
}

// ginv != 0: add0 is read through the automorphism (the rotated c0 of a rotation, as k_galois).
template <int LOGR, int KT>
__global__ __launch_bounds__(256) void k_inv_cols_md(const double* __restrict__ acc, long long EN, int nq,
                                                     const double* __restrict__ pg, int L,
                                                     const uint64_t* __restrict__ add0, const uint64_t* __restrict__ add1,
                                                     long long add_ps, unsigned long long ginv, OutPtrs outs,
                                                     long long P, int logn,
                                                     const DevPrime* __restrict__ primes,
                                                     const double* __restrict__ pmodq,
                                                     const double* __restrict__ pinvq,
                                                     const double* __restrict__ pmod_aux) {
    constexpr int R = 1 << LOGR, TPS = R / 16, S = 256 / TPS;
    __shared__ double lds[R * S];
    const int n = 1 << logn, N2 = n >> LOGR, ncb = N2 / S;
    // XCD-aware order: the nq limb workgroups of one column block run back to back on one XCD
    // (workgroup w -> XCD w % 8), so that XCD's L2 serves the block's aux-limb values to all of them.
    // With fewer than 8 column blocks (N <= 2^14) each block is spread over 8 / ncb XCDs, each taking
    // every (8 / ncb)-th limb, so that every XCD has work.
    const int w = blockIdx.x;
    int i, cbk;
    if (ncb >= 8) {
        const int k = w >> 3;
        i = k % nq;
        cbk = (k / nq) * 8 + (w & 7);
        if (cbk >= ncb) return;
    } else {
        const int x = w & 7, rep = 8 / ncb;
        cbk = x % ncb;
        i = x / ncb + rep * (w >> 3);
        if (i >= nq) return;
    }
    const int tid = threadIdx.x, col = tid % S, tau = tid / S;
    const int r = cbk * S + col;
    const int z = blockIdx.z, cb = z >> 1, comp = z & 1;
    const DevPrime Pi = primes[i];
    const double q = Pi.q, qinv = Pi.qinv;
    const double* src = acc + (long long)z * EN + (long long)i * n;
    double v[16];
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = src[r + (long long)N2 * (16 * tau + e)];
    inv_phase_b<LOGR>(v, Pi.iv, 1, tau, q, qinv);
#pragma unroll
    for (int e = 0; e < 16; e++) lds[(16 * tau + e) * S + col] = v[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = lds[(tau + TPS * e) * S + col];
    inv_phase_a<kTwScalar>(v, Pi.iv, 1, q, qinv);
    // mod-down epilogue in chunks of four positions, each chunk's loads issued together
    const double* pp = pg + (long long)z * (KT + 1) * n;
    const uint64_t* add = comp == 0 ? add0 : add1;
    const uint64_t* ad = add ? add + (long long)cb * add_ps + (long long)i * n : nullptr;
    uint64_t* o = outs.p[cb] + (long long)comp * P + (long long)i * n;
    const double pinv = pinvq[i], pmq = pmodq[i];
    double paux[KT > 1 ? KT - 1 : 1];
#pragma unroll
    for (int kk = 0; kk + 1 < KT; kk++) paux[kk] = pmod_aux[(long long)kk * L + i];
#pragma unroll
    for (int e0 = 0; e0 < 16; e0 += 4) {
        double mp[4][KT + 1], av[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int idx = r + N2 * (tau + TPS * (e0 + e));
#pragma unroll
            for (int kk = 0; kk <= KT; kk++) mp[e][kk] = pp[(long long)kk * n + idx];
            if (!ad) {
                av[e] = 0.0;
            } else if (comp == 0 && ginv) {
                const unsigned long long k = ((unsigned long long)idx * ginv) & (2ull * n - 1);
                const uint64_t x = ad[k >= (unsigned long long)n ? (long long)k - n : (long long)k];
                av[e] = k >= (unsigned long long)n && x != 0 ? q - (double)x : (double)x;
            } else {
                av[e] = (double)ad[idx];
            }
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int idx = r + N2 * (tau + TPS * (e0 + e));
            const double x = mulmod(v[e0 + e], Pi.ninv, q, qinv);  // x_i, |x| <= 0.97 q
            // centred(x_P) mod q_i from the aux part's Garner digits m_k and its centring flag:
            // Horner over m_{K-1} .. m_0 with p_k mod q_i, minus P mod q_i when x_P > floor(P/2)
            double cv = reduce(mp[e][KT - 1], q, qinv);
#pragma unroll
            for (int kk = KT - 2; kk >= 0; kk--) cv = reduce(mulmod(cv, paux[kk], q, qinv) + mp[e][kk], q, qinv);
            if (mp[e][KT] != 0.0) cv -= pmq;
            const double res = mulmod(reduce(x - cv, q, qinv), pinv, q, qinv) + av[e];
            o[idx] = f64_to_u64(canon(res, q, qinv));
        }
    }
}

void launch_inv_cols_moddown(const Context& c, const double* acc, int nq, int B, const uint64_t* accP,
                             const uint64_t* add0, const uint64_t* add1, long long add_ps,
                             const std::vector<uint64_t*>& outs, hipStream_t s, unsigned long long ginv) {
    require_fp64(c);
    RCK_REQUIRE(c.K > 0 && c.logn >= 12, "fused mod-down needs an aux basis and the two-pass NTT");
    const long long n = c.n, EN = (long long)(nq + c.K) * n, P = (long long)nq * n;
    int n1, n2;
    split_logn(c.logn, &n1, &n2);
    // the aux part's Garner digits and centring flag, once per coefficient
    Context& cc = const_cast<Context&>(c);
    double* pg = static_cast<double*>(cc.workspace(9, (size_t)B * 2 * (c.K + 1) * n * 8));
    {
        ProfScope ps(s, "p_garner");
        k_p_garner<<<dim3((c.n + 255) / 256, 2 * B), 256, 0, s>>>(accP, pg, c.n, c.K, c.L, c.L + c.K, c.d_primes,
                                                                 c.d_invq, c.d_pwords);
        RCK_CHECK_HIP(hipGetLastError());
    }
    for (int b0 = 0; b0 < B; b0 += 32) {
        const int bc = std::min(32, B - b0);
        OutPtrs op{};
        for (int b = 0; b < bc; b++) op.p[b] = outs[b0 + b];
        const double* accb = acc + (long long)b0 * 2 * EN;
        const double* accPb = pg + (long long)b0 * 2 * (c.K + 1) * n;
        const uint64_t* a0 = add0 ? add0 + b0 * add_ps : nullptr;
        const uint64_t* a1 = add1 ? add1 + b0 * add_ps : nullptr;
        ProfScope ps(s, "inv_cols_moddown");
        const int S = 256 / ((1 << n1) / 16);
        const int ncb = (1 << n2) / S;
        // see k_inv_cols_md for the order (ncb is a power of two)
        const int nw = ncb >= 8 ? ((ncb + 7) / 8) * 8 * nq : 8 * ((nq + 8 / ncb - 1) / (8 / ncb));
        dim3 grid(nw, 1, 2 * bc);
#define RCK_MD(LR, KK)                                                                                      \
    k_inv_cols_md<LR, KK><<<grid, 256, 0, s>>>(accb, EN, nq, accPb, c.L, a0, a1, add_ps, ginv, op, P, c.logn, c.d_primes, \
                                               c.d_pmodq, c.d_pinvq, c.d_pmod_aux)
#define RCK_MDK(LR)                                     \
    switch (c.K) {                                      \
        case 1: RCK_MD(LR, 1); break;                   \
        case 2: RCK_MD(LR, 2); break;                   \
        case 3: RCK_MD(LR, 3); break;                   \
        default: RCK_MD(LR, 4); break;                  \
    }
        switch (n1) {
            case 6: RCK_MDK(6); break;
            case 7: RCK_MDK(7); break;
            case 8: RCK_MDK(8); break;
            default: RCK_REQUIRE(false, "unsupported ring degree");
        }
#undef RCK_MDK
#undef RCK_MD
        RCK_CHECK_HIP(hipGetLastError());
    }
}

void launch_mod_down(const Context& c, const uint64_t* acc, int nq, const uint64_t* add0, const uint64_t* add1,
                     uint64_t* out, long long out_ps, hipStream_t s) {
    require_fp64(c);
    RCK_REQUIRE(c.K > 0, "mod-down requires an auxiliary basis");
    ProfScope ps(s, "mod_down");
    k_mod_down<<<dim3((c.n + 255) / 256, 2), 256, 0, s>>>(acc, c.n, nq, c.K, c.L, c.L + c.K, add0, add1, out, out_ps,
                                                          c.d_primes, c.d_invq, c.d_pwords, c.d_pmodq, c.d_pinvq,
                                                          c.d_pmod_aux);
    RCK_CHECK_HIP(hipGetLastError());
}

}  // namespace rck
