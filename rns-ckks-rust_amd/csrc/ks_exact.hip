// Keyswitch by exact integer convolution through two auxiliary NTT primes.
//
// keyswitch_rns_hybrid (src/rns/keyswitching.rs:159-194) needs, for every eval prime p_j of Q_l || P,
//   acc_j = sum_i d_i * k_ij  mod (X^N + 1, p_j),
// where the digits d_i (|d| <= 2^29) are the same integer polynomials for every p_j and k_ij is key
// digit i on limb j.  Take k_ij as its centred integer polynomial (|k| <= (p_j - 1)/2): then every
// coefficient of the INTEGER negacyclic sum V_j = sum_{i in a group} d_i * k_ij is bounded by
//   |V_j| <= G * N * 2^29 * max_j (p_j - 1)/2,
// and with G (digits per group) small enough that bound stays below (s1 s2 - 1)/2 for two auxiliary
// NTT primes s1, s2 < 2^50 + 2^46.  V_j is then recovered exactly from its residues mod s1 and s2
// (centred CRT), and V_j mod p_j summed over the groups is acc_j itself -- the same residues the
// crate's per-prime products give, bit for bit.
//
// So each digit is transformed under s1 and s2 only (2 limb NTTs instead of one per eval prime), the
// key is stored once as the s1/s2 NTT forms of its centred limbs, the products are, at every NTT slot,
// a (ciphertexts x digits) by (digits x 2 (l + K) key columns) matrix product mod s computed exactly on
// the int8 matrix cores (every key value serves all ciphertexts of a batch), and the accumulators are
// transformed back under s1, s2, recombined, reduced mod p_j and taken through mod-down by P.
#include <algorithm>

#include "moddown.hpp"
#include "pipeline.hpp"

namespace rck {

// Constants of the recombination: the auxiliary primes, s1^{-1} mod s2, and the centring thresholds.
struct SpecConst {
    double s1, s1inv, s2, s2inv, inv12;  // inv12 = s1^{-1} mod s2 (centred)
    double half_t, half_a;               // (s2 - 1)/2, (s1 - 1)/2
};

static SpecConst spec_const(const Context& c) {
    SpecConst k;
    k.s1 = (double)c.sp[0];
    k.s1inv = 1.0 / k.s1;
    k.s2 = (double)c.sp[1];
    k.s2inv = 1.0 / k.s2;
    k.inv12 = c.sp_inv12;
    k.half_t = (double)((c.sp[1] - 1) / 2);
    k.half_a = (double)((c.sp[0] - 1) / 2);
    return k;
}

// V mod q for the integer V in (-(s1 s2)/2, (s1 s2)/2] with residues a1 = V mod s1, a2 = V mod s2
// (canonical).  Garner: V = a1 + s1 t, t = (a2 - a1) s1^{-1} mod s2 in [0, s2), minus s1 s2 when
// a1 + s1 t > (s1 s2 - 1)/2, i.e. t > (s2 - 1)/2, or t == (s2 - 1)/2 and a1 > (s1 - 1)/2.
// p1m = s1 mod q, p12m = s1 s2 mod q (centred).  Returns a representative with |r| <= q/2 + 1.
__device__ __forceinline__ double crt_mod_q(double a1, double a2, const SpecConst& K, double q, double qinv,
                                            double p1m, double p12m) {
    const double d = reduce(a2 - a1, K.s2, K.s2inv);                          // |a2 - a1| < 2^51
    const double t = canon(mulmod(d, K.inv12, K.s2, K.s2inv), K.s2, K.s2inv);  // [0, s2)
    // both comparisons evaluated (selects, no branch)
    const bool neg = (t > K.half_t) | ((t == K.half_t) & (a1 > K.half_a));
    double r = reduce(a1, q, qinv) + mulmod(reduce(t, q, qinv), p1m, q, qinv);  // <= q/2 + 1 + 0.97q
    r -= neg ? p12m : 0.0;
    return reduce(r, q, qinv);
}

// Centred key limbs reduced mod s1, s2: in [2][E][N] canonical mod p_j -> out [2][E][2][N] (reduced
// doubles, the input of the forward NTT under s1, s2).
__global__ void k_key_to_special(const uint64_t* __restrict__ in, double* __restrict__ out, int n, int E, int nq,
                                 int L, int sbase, const DevPrime* __restrict__ primes) {
    const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int j = blockIdx.y, comp = blockIdx.z;
    const DevPrime Pj = primes[prime_of(j, nq, L)];
    const double v = load_centred(in[((long long)comp * E + j) * n + x], Pj.q);
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const DevPrime Ps = primes[sbase + s];
        out[(((long long)comp * E + j) * 2 + s) * n + x] = reduce(v, Ps.q, Ps.qinv);
    }
}

// Inverse of k_key_to_special after the inverse NTT: in [2][E][2][N] canonical mod s -> out [2][E][N]
// canonical mod p_j (the key's coefficient form).
__global__ void k_key_from_special(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, int n, int E, int nq,
                                   int L, SpecConst K, const double* __restrict__ spmod,
                                   const DevPrime* __restrict__ primes) {
    const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int j = blockIdx.y, comp = blockIdx.z;
    const int pj = prime_of(j, nq, L);
    const DevPrime Pj = primes[pj];
    const long long base = (((long long)comp * E + j) * 2) * n + x;
    const double v = crt_mod_q((double)in[base], (double)in[base + n], K, Pj.q, Pj.qinv, spmod[2 * pj],
                               spmod[2 * pj + 1]);
    out[((long long)comp * E + j) * n + x] = f64_to_u64(canon(v, Pj.q, Pj.qinv));
}

// ---------------------------------------------------------------------------------------
// Multiply-accumulate on the int8 matrix cores.
//
// At NTT slot x under auxiliary prime s the accumulators of a batch are one matrix product mod s:
//   acc[b][jc] = sum_i dn[b][i] * key[i][jc]   (B ciphertexts x D digits, by D digits x 2E key columns),
// every operand a residue mod s < 2^50.1, centred.  Each operand is written in balanced signed bytes,
// x = sum_{u<7} x_u 256^u with x_u in [-128, 128).  With the byte planes dn_u[b][i] and key_v[i][jc],
//   acc = sum_p 256^p C_p,   C_p = sum_{u+v=p} key_v^T dn_u      (p = 0 .. 12, |C_p| <= 7 D 2^14 < 2^23),
// 49 products of int8 matrices.  One v_mfma_i32_16x16x64_i8 multiplies a 16-column key tile by 16
// ciphertexts over K = 64 digits (D <= 64): lane l holds key column / ciphertext l % 16 and the digits
// 16 (l / 16) + t, t < 16, of one byte plane (16 bytes); output key column 4 (l / 16) + r, ciphertext
// l % 16 in register r.  The 13 sums are recombined in FP64 from three exact chunks (positions 0-3, 4-7
// and 8-12, each below 2^47.1; see mfma_tile) with two exact modular products by 2^32 mod s, then
// reduced: the value the crate's per-prime products sum to, mod s.
// ---------------------------------------------------------------------------------------
typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned long long v2u __attribute__((ext_vector_type(2)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

// word 3 of a raw buffer resource on gfx9-family GPUs (untyped dword access, range-checked against the
// record count: an out-of-range store is dropped)
constexpr int kBufferRsrcWord3 = 0x00020000;

constexpr uint64_t kByteBias = 0x0080808080808080ull;  // 128 in bytes 0 .. 6

// integer-valued double |v| < 2^51 -> int64: v + 1.5 2^52 has the bit pattern 0x4338000000000000 + v
__device__ __forceinline__ int64_t f64_to_i64(double v) {
    return (int64_t)__double_as_longlong(v + kShift) - 0x4338000000000000ll;
}
// balanced signed-byte word of |x| < 2^55 (byte u = x_u as two's complement, byte 7 zero) and back
__device__ __forceinline__ uint64_t sbyte_word(int64_t x) { return ((uint64_t)x + kByteBias) ^ kByteBias; }
__device__ __forceinline__ int64_t sbyte_value(uint64_t w) { return (int64_t)((w ^ kByteBias) - kByteBias); }

// 4 x 4 byte transpose (v_perm_b32 picks bytes of {src0 : src1}, src1 the low word): y_u byte i = x_i byte u
__device__ __forceinline__ void byte_transpose4(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3, uint32_t (&y)[4]) {
    const uint32_t t0 = __builtin_amdgcn_perm(x1, x0, 0x05010400u);  // x0.b0 x1.b0 x0.b1 x1.b1
    const uint32_t t1 = __builtin_amdgcn_perm(x1, x0, 0x07030602u);  // x0.b2 x1.b2 x0.b3 x1.b3
    const uint32_t t2 = __builtin_amdgcn_perm(x3, x2, 0x05010400u);
    const uint32_t t3 = __builtin_amdgcn_perm(x3, x2, 0x07030602u);
    y[0] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
    y[1] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
    y[2] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
    y[3] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
}

// A slot's accumulator record (B ciphertexts x cols key columns) is N-tile-major: the columns of N-tile nt
// (16, the last one w = cols - 16 (NT - 1)) of ciphertext b at 16 nt B + b w + (col - 16 nt), so the MAC
// writes each tile's outputs as one contiguous block.  acc_record_z is the inverse, to z = b cols + col.
__host__ __device__ __forceinline__ int acc_record_pos(int b, int col, int B, int cols) {
    const int nt = col >> 4, NT = (cols + 15) >> 4;
    const int w = nt == NT - 1 ? cols - 16 * (NT - 1) : 16;
    return 16 * nt * B + b * w + (col - 16 * nt);
}
__host__ __device__ __forceinline__ int acc_record_z(int p, int B, int cols) {
    const int NT = (cols + 15) >> 4, full = 16 * B * (NT - 1);
    int nt, b, jc;
    if (p < full) {
        nt = p / (16 * B);
        b = (p >> 4) % B;
        jc = p & 15;
    } else {
        const int w = cols - 16 * (NT - 1);
        nt = NT - 1;
        b = (p - full) / w;
        jc = (p - full) % w;
    }
    return b * cols + 16 * nt + jc;
}

struct MfmaArgs {
    const uint64_t* dsb;   // [2][N][B][Dp] digit words (slot-major)
    const uint8_t* key;    // [2][N][7 cols Dk] key byte planes (Dk = digits rounded up to 8), then 64 bytes of padding
    double* acc;           // [G][2][N] slot records of B cols accumulators (reduced), in acc_record_pos order
    int n, B, Dp, Dk, cols, D, gsize;
    double q[2], qinv[2], c32[2];  // the auxiliary primes, 1/s, 2^32 mod s (centred)
};

// The digit planes of the lane's 16 digits from their 16 words w (two per v2u): plane u, word k holds
// byte u of digits 4k .. 4k + 3.  Digits outside [i0, i1) (index d0 + t) are zeroed first: those past the
// ciphertext's digits (the key's bytes there are other data) and, with several groups, the other groups'.
__device__ __forceinline__ void digit_planes(const v2u (&w)[8], v4i (&dp)[7], int d0, int i0, int i1) {
    uint32_t lo[16], hi[16];
#pragma unroll
    for (int h = 0; h < 8; h++) {
        uint64_t a = w[h].x, b = w[h].y;
        const int e = d0 + 2 * h;
        a = e >= i0 && e < i1 ? a : 0ull;
        b = e + 1 >= i0 && e + 1 < i1 ? b : 0ull;
        lo[2 * h] = (uint32_t)a;
        hi[2 * h] = (uint32_t)(a >> 32);
        lo[2 * h + 1] = (uint32_t)b;
        hi[2 * h + 1] = (uint32_t)(b >> 32);
    }
    uint32_t P[8][4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint32_t y[4];
        byte_transpose4(lo[4 * k], lo[4 * k + 1], lo[4 * k + 2], lo[4 * k + 3], y);
#pragma unroll
        for (int u = 0; u < 4; u++) P[u][k] = y[u];
        byte_transpose4(hi[4 * k], hi[4 * k + 1], hi[4 * k + 2], hi[4 * k + 3], y);
#pragma unroll
        for (int u = 0; u < 4; u++) P[4 + u][k] = y[u];  // plane 7 is zero and unused
    }
#pragma unroll
    for (int u = 0; u < 7; u++) dp[u] = v4i{(int)P[u][0], (int)P[u][1], (int)P[u][2], (int)P[u][3]};
}

// One step: the 16 x 16 output tile (N-tile of key columns, M-tile of ciphertexts) of one slot: the 49
// plane products accumulated into 13 positions, key plane by key plane.  refill(v) runs right after key
// plane v's last product: it loads the next step's plane into the registers this one no longer needs,
// so every load has about one step to arrive and no second buffer is held.
// Recombination: the chunks H_low (positions 0-3), H_mid (4-7) and H_top (8-12) are exact doubles.  Byte
// 6 of a residue below 2^50.1 (centred) is at most 3 in size, so |C_12| <= 9 D and every chunk stays
// below D 2^40.9 <= 2^46.9 (D <= 64); V = (H_top c + H_mid) c + H_low mod s, c = 2^32 mod s.
template <class Refill>
__device__ __forceinline__ void mfma_tile(const v4i (&dp)[7], v4i (&kp)[7], double (&V)[4], double q, double qinv,
                                          double c32, Refill refill) {
    v4i c[13];
#pragma unroll
    for (int p = 0; p < 13; p++) c[p] = v4i{0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 7; v++) {
#pragma unroll
        for (int u = 0; u < 7; u++) c[u + v] = __builtin_amdgcn_mfma_i32_16x16x64_i8(kp[v], dp[u], c[u + v], 0, 0, 0);
        refill(v);
        // keeps the refill load here (the compiler would otherwise sink it to its use); the barrier reads
        // no register, so it adds no wait
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        double top = (double)c[12][r];
#pragma unroll
        for (int p = 11; p >= 8; p--) top = __builtin_fma(top, 256.0, (double)c[p][r]);
        double mid = (double)c[7][r];
#pragma unroll
        for (int p = 6; p >= 4; p--) mid = __builtin_fma(mid, 256.0, (double)c[p][r]);
        double low = (double)c[3][r];
#pragma unroll
        for (int p = 2; p >= 0; p--) low = __builtin_fma(low, 256.0, (double)c[p][r]);
        const double t = mulmod(top, c32, q, qinv) + mid;
        V[r] = reduce(mulmod(t, c32, q, qinv) + low, q, qinv);
        __builtin_amdgcn_sched_barrier(0);  // one output register at a time: 13 sums read, not 52
    }
}

// A wave's slots xbeg .. xbeg+nslots-1 (xbeg wave-uniform): for each (slot, M-tile) its N-tiles in order.
// The products are D = key^T x digits, so lane l holds ciphertext 16 mt + l % 16 and the four key columns
// 16 nt + 4 (l / 16) + r: 32 contiguous bytes of the tile's block in the slot's accumulator record
// (acc_record_pos), two 16-byte stores.
// No memory operation of the loops is conditional, so the in-order memory counter lets every wait cover
// exactly the loads it needs:
//  * each step's refills load the next N-tile's key planes; the last N-tile's also load the next
//    (slot, M-tile)'s digit words (a slot's digit words are loaded once); the last step reloads its own
//    operands (valid addresses, never used);
//  * key columns past the last tile's width read its last column (those products are not stored);
//    digit words past D read the row's last words, and the digits outside the group (past D included)
//    are zeroed when the planes are formed, so the key bytes they meet (other data or padding) add nothing;
//  * a step's stores are issued during the next step, after its first refill, as buffer stores over the
//    slot's accumulator record; lanes with no output (padding columns, rows past the batch) give an offset
//    past the record, and the buffer range check drops their stores.
// Each address is a wave-uniform base plus a lane offset fixed for the launch.
__device__ __forceinline__ void mfma_wave(const MfmaArgs& A, int s, long long xbeg, int nslots, int g, int i0, int i1,
                                          int lane) {
    const int j = lane & 15, gq = lane >> 4;
    const int NT = (A.cols + 15) >> 4, MT = (A.B + 15) >> 4;
    const int wl = A.cols - 16 * (NT - 1);           // width of the last N-tile
    const long long kslot = 7LL * A.cols * A.Dk;    // key bytes per slot
    // lane byte offsets (32-bit) within a plane of an N-tile: column j (the last tile's columns past its
    // width read its last column; those products are not stored), digits 16 gq .. 16 gq + 15 (past the key's
    // digits the bytes are other data or padding, multiplied by zeroed digits)
    const uint32_t lk_full = (uint32_t)(j * A.Dk + 16 * gq), lk_last = (uint32_t)(min(j, wl - 1) * A.Dk + 16 * gq);
    const double q = A.q[s], qinv = A.qinv[s], c32 = A.c32[s];
    const long long zc = (long long)A.B * A.cols;
    const int d0 = 16 * gq;  // the lane's first digit
    typedef const __attribute__((address_space(1))) char gchar;
    auto uniform_ptr = [](const void* p) -> gchar* {
        const uint64_t v = (uint64_t)p;
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
        return (gchar*)(((uint64_t)hi << 32) | lo);
    };
    auto load_k = [&](long long sl, int nt, int v) -> v4i {
        const bool last = nt == NT - 1;
        const int w = last ? wl : 16;
        const long long base = sl * kslot + (7LL * 16 * nt + (long long)v * w) * A.Dk;
        gchar* kp = uniform_ptr(A.key + base);
        return *(const __attribute__((address_space(1))) v4i*)(kp + (uint64_t)(last ? lk_last : lk_full));
    };
    // digit words 2h, 2h + 1 of the lane's 16 (ciphertext 16 mt + j; rows past the batch repeat the last one,
    // they are not stored)
    auto load_w = [&](long long sl, int mt, int h) -> v2u {
        gchar* dp = uniform_ptr(A.dsb + (sl * A.B + 16LL * mt) * A.Dp);
        const int e = min(d0 + 2 * h, A.Dp - 2);
        const uint32_t la = (uint32_t)(min(j, A.B - 1 - 16 * mt) * A.Dp + e) * 8u;
        return *(const __attribute__((address_space(1))) v2u*)(dp + (uint64_t)la);
    };
    // output of a step: ciphertext 16 mt + j, columns 16 nt + 4 gq + 2 pr + {0, 1} (cols is even, so a pair
    // is valid or not as a whole), as a byte offset in the slot's record of zc doubles (acc_record_pos: a
    // full N-tile's 16 ciphertexts are 2 KiB of whole lines); kDrop for lanes without an output
    constexpr uint32_t kDrop = 0x80000000u;
    double* const accg = A.acc + (long long)g * 2 * A.n * zc;
    auto out_off = [&](int mt, int nt, int pr) -> uint32_t {
        const int bb = 16 * mt + j, col = 16 * nt + 4 * gq + 2 * pr;
        const bool ok = (bb < A.B) & (col < A.cols);  // both evaluated: a select, no branch
        const uint32_t off = (uint32_t)acc_record_pos(bb, col, A.B, A.cols) * 8u;
        return ok ? off : kDrop;
    };
    typedef double v2d __attribute__((ext_vector_type(2)));
    const int groups = nslots * MT;
    long long sl = (long long)s * A.n + xbeg;
    int mt = 0;
    v2u wb[8];  // digit words of the next (slot, M-tile)
    v4i kp[7], dp[7];
    // first step's operands, in the order the refills issue them
#pragma unroll
    for (int v = 0; v < 7; v++) {
        kp[v] = load_k(sl, 0, v);
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int h = 0; h < 8; h++) wb[h] = load_w(sl, 0, h);
    // the previous step's outputs, stored during this step (the first step's stores are all dropped)
    v2d Vp[2];
    uint32_t oo[2];
    long long osl = sl;  // the slot of those outputs (wave-uniform)
#pragma unroll
    for (int pr = 0; pr < 2; pr++) {
        Vp[pr] = v2d{0.0, 0.0};
        oo[pr] = kDrop;
    }
    auto store_prev = [&]() {
        const __amdgpu_buffer_rsrc_t rec = __builtin_amdgcn_make_buffer_rsrc(
            (void*)uniform_ptr(accg + osl * zc), (short)0, (int)(zc * 8), kBufferRsrcWord3);
#pragma unroll
        for (int pr = 0; pr < 2; pr++)
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, Vp[pr]), rec, (int)oo[pr], 0, 0);
    };
    auto keep = [&](const double (&V)[4], long long sl0, int mt0, int nt0) {
        osl = sl0;
#pragma unroll
        for (int pr = 0; pr < 2; pr++) {
            Vp[pr] = v2d{V[2 * pr], V[2 * pr + 1]};
            oo[pr] = out_off(mt0, nt0, pr);
        }
    };
    for (int u = 0; u < groups; u++) {
        int mt1 = mt + 1;
        long long sl1 = sl;
        if (mt1 == MT) {
            mt1 = 0;
            sl1++;
        }
        if (u + 1 == groups) {  // the last step reloads its own operands (a valid address; never used)
            mt1 = mt;
            sl1 = sl;
        }
        digit_planes(wb, dp, d0, i0, i1);
        double V[4];
        for (int nt = 0; nt < NT - 1; nt++) {
            mfma_tile(dp, kp, V, q, qinv, c32, [&](int v) {
                kp[v] = load_k(sl, nt + 1, v);
                if (v == 0) store_prev();
            });
            keep(V, sl, mt, nt);
        }
        mfma_tile(dp, kp, V, q, qinv, c32, [&](int v) {
            kp[v] = load_k(sl1, 0, v);
            if (v == 0) {
                store_prev();
#pragma unroll
                for (int h = 0; h < 8; h++) wb[h] = load_w(sl1, mt1, h);
            }
        });
        keep(V, sl, mt, NT - 1);
        sl = sl1;
        mt = mt1;
    }
    store_prev();
}

// Work item (digit group g, prime s, block of 4 spw slots); each of the four waves takes spw consecutive
// slots.  gate (nullable, device): the launch does nothing unless *gate != 0.
__global__ __launch_bounds__(256) void k_ks_mac_mfma(MfmaArgs A, int spw, int nxb, const int* __restrict__ gate) {
    if (gate != nullptr && *gate == 0) return;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    int r = blockIdx.x;
    const int xb = r % nxb;
    r /= nxb;
    const int s = r & 1, g = r >> 1;
    const int i0 = g * A.gsize, i1 = min(A.D, i0 + A.gsize);
    mfma_wave(A, s, ((long long)xb * 4 + wave) * spw, spw, g, i0, i1, lane);
}

static MfmaArgs mfma_args(const Context& c, const uint64_t* dsb, const uint8_t* key, double* acc,
                          int B, int D, int Dp, int Dk, int cols, int gsize) {
    MfmaArgs A{};
    A.dsb = dsb;
    A.key = key;
    A.acc = acc;
    A.n = (int)c.n;
    A.B = B;
    A.Dp = Dp;
    A.Dk = Dk;
    A.cols = cols;
    A.D = D;
    A.gsize = gsize;
    for (int s = 0; s < 2; s++) {
        const uint64_t sp = c.sp[s];
        A.q[s] = (double)sp;
        A.qinv[s] = 1.0 / (double)sp;
        const uint64_t r = (uint64_t)(((unsigned __int128)1 << 32) % sp);
        A.c32[s] = r > sp / 2 ? (double)r - (double)sp : (double)r;
    }
    return A;
}

// acc: G groups of [2][N][B cols]; kdig: the key's digits (its K-groups); spw slots per wave (1, 2 or 4);
// digits in groups of gsize (G groups); gate as in k_ks_mac_mfma
static void launch_mac_mfma(const Context& c, const uint64_t* dsb, const uint8_t* key, int kdig, double* acc,
                            int B, int D, int Dp, int cols, int gsize, int G, int spw, hipStream_t s,
                            const int* gate = nullptr) {
    RCK_REQUIRE(spw == 1 || spw == 2 || spw == 4, "bad exact keyswitch shape");
    RCK_REQUIRE(c.n % (4 * spw) == 0 && B >= 1 && D <= kdig && kdig <= 64 && Dp >= 2 && Dp % 2 == 0 && cols % 2 == 0,
                "bad exact keyswitch size");
    const MfmaArgs A = mfma_args(c, dsb, key, acc, B, D, Dp, 8 * ((kdig + 7) / 8), cols, gsize);
    const int nxb = (int)(c.n / (4 * spw));
    const long long blocks = (long long)G * 2 * nxb;
    RCK_REQUIRE(blocks < (1LL << 31), "exact keyswitch launch too large");
    // a slot's accumulator record is one buffer range; its byte offsets stay below the dropped offset 2^31
    RCK_REQUIRE((long long)B * cols * 8 < (1LL << 31), "exact keyswitch batch too large");
    ProfScope ps(s, gate ? "ks_mac_gated" : "ks_mac");
    k_ks_mac_mfma<<<dim3((unsigned)blocks), 256, 0, s>>>(A, spw, nxb, gate);
    RCK_CHECK_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// Row passes between the standard limb rows and the slot-major operands of the matrix-core MAC.
// A workgroup takes ring row rb of S consecutive rows z under one prime; the tile goes through LDS
// once more so that each store (or load) instruction covers S consecutive z of a few slots.  That
// exchange is stored position-major, tile_lds(z, x) = x (S + 1) + z: the S rows of one position are
// adjacent (the slot-major side reads or writes S consecutive z of one position) and a position's
// block is S + 1 doubles, so the row side's 16 consecutive positions of one z fall on distinct banks.
// ---------------------------------------------------------------------------------------
template <int S>
__device__ __forceinline__ int tile_lds(int z, int x) {
    return x * (S + 1) + z;  // R (S + 1) <= S row_pitch<R>() doubles
}

// Forward row stages of the digit transforms: in rows (z, s) at in + (z 2 + s) N after the column pass,
// z = b D + i -> out[(s N + x) B Dp + b Dp + i] as signed-byte words (x the NTT-domain position), over
// the B Dp rows zp = b Dp + i: the padding digits D <= i < Dp are written as zeros, so the
// multiply-accumulate reads whole K-blocks without masks.
template <int LOGR>
__global__ __launch_bounds__(256) void k_fwd_rows_sm(const double* __restrict__ in, uint64_t* __restrict__ out,
                                                     int logn, int D, int Dp, int B, int sbase,
                                                     const DevPrime* __restrict__ primes) {
    constexpr int R = 1 << LOGR, TPS = R / 16, S = 256 / TPS;
    static_assert(R * (S + 1) <= S * row_pitch<R>(), "the exchange tile must fit the row buffer");
    __shared__ double lds[S * row_pitch<R>()];
    const int n = 1 << logn, N1 = n >> LOGR, Zp = B * Dp;
    const int tid = threadIdx.x, tau = tid % TPS, row = tid / TPS;
    const int z0 = blockIdx.x * S, rb = blockIdx.y, s = blockIdx.z;
    const DevPrime P = primes[sbase + s];
    const double q = P.q, qinv = P.qinv;
    // rows past Zp or on a padding digit transform digit row 0 (a valid row); their results are replaced
    const int zp = min(z0 + row, Zp - 1), bi = zp / Dp, di = zp % Dp;
    const bool real = z0 + row < Zp && di < D;
    const int z = real ? bi * D + di : 0;
    const double* p = in + ((long long)z * 2 + s) * n + (long long)rb * R;
    double v[16];
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = p[tau + TPS * e];
    fwd_phase_a(v, P.fw, N1 + rb, q, qinv);
#pragma unroll
    for (int e = 0; e < 16; e++) lds[row_lds<R>(row, tau + TPS * e)] = v[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = lds[row_lds<R>(row, 16 * tau + e)];
    fwd_phase_b<LOGR>(v, P.fw, N1 + rb, tau, q, qinv);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; e++) lds[tile_lds<S>(row, e * TPS + tau)] = real ? v[e] : 0.0;  // position e TPS + tau
    __syncthreads();
    const int zz = tid % S, pos0 = tid / S, zo = z0 + zz;
    if (zo >= Zp) return;
    uint64_t* o = out + ((long long)s * n + (long long)rb * R) * Zp + zo;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int pos = pos0 + TPS * e;
        o[(long long)pos * Zp] = sbyte_word(f64_to_i64(lds[tile_lds<S>(zz, pos)]));
    }
}

// Inverse row stages of slot-major accumulators: in [G][2][N] records of Zc = B cols (acc_record_pos
// order) -> rows ((g Zc + z) 2 + s) N, z = b cols + col, the layout of the fused inverse column pass and
// recombination.  A workgroup reads S consecutive record positions.
template <int LOGR>
__global__ __launch_bounds__(256) void k_inv_rows_sm(const double* __restrict__ in, double* __restrict__ out,
                                                     int logn, int Zc, int B, int sbase,
                                                     const DevPrime* __restrict__ primes,
                                                     const int* __restrict__ gate) {
    constexpr int R = 1 << LOGR, TPS = R / 16, S = 256 / TPS;
    static_assert(R * (S + 1) <= S * row_pitch<R>(), "the exchange tile must fit the row buffer");
    __shared__ double lds[S * row_pitch<R>()];
    if (gate != nullptr && *gate == 0) return;
    const int n = 1 << logn, N1 = n >> LOGR;
    const int tid = threadIdx.x, tau = tid % TPS, row = tid / TPS;
    const int z0 = blockIdx.x * S, rb = blockIdx.y, s = blockIdx.z & 1, g = blockIdx.z >> 1;
    const DevPrime P = primes[sbase + s];
    const double q = P.q, qinv = P.qinv;
    const int zz = tid % S, pos0 = tid / S;
    const double* src =
        in + ((long long)(g * 2 + s) * n + (long long)rb * R) * Zc + min(z0 + zz, Zc - 1);
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int pos = pos0 + TPS * e;
        lds[tile_lds<S>(zz, pos)] = src[(long long)pos * Zc];
    }
    __syncthreads();
    double v[16];
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = lds[tile_lds<S>(row, e * TPS + tau)];  // position e TPS + tau: x = 16 tau + e
    inv_phase_b<LOGR>(v, P.iv, N1 + rb, tau, q, qinv);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; e++) lds[row_lds<R>(row, 16 * tau + e)] = v[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = lds[row_lds<R>(row, tau + TPS * e)];
    inv_phase_a(v, P.iv, N1 + rb, q, qinv);
    const int z = z0 + row;
    if (z >= Zc) return;
    const int zo = acc_record_z(z, B, Zc / B);
    double* o = out + (((long long)g * Zc + zo) * 2 + s) * n + (long long)rb * R;
#pragma unroll
    for (int e = 0; e < 16; e++) o[tau + TPS * e] = v[e];
}

template <int LOGR>
static void rows_sm_t(const Context& c, bool fwd, const double* din, uint64_t* dout, const double* ain, double* aout,
                      int Z, int D, int Dp, int B, int G, hipStream_t s, const int* gate) {
    constexpr int S = 256 / ((1 << LOGR) / 16);
    const int n1 = 1 << (c.logn - LOGR);
    const int sbase = (int)(c.L + c.K);
    if (fwd) {
        ProfScope ps(s, "ntt_fwd_rows");
        const int Zp = B * Dp;
        k_fwd_rows_sm<LOGR><<<dim3((Zp + S - 1) / S, n1, 2), 256, 0, s>>>(din, dout, (int)c.logn, D, Dp, B, sbase,
                                                                          c.d_primes);
    } else {
        ProfScope ps(s, gate ? "ntt_inv_rows_gated" : "ntt_inv_rows");
        RCK_REQUIRE(B >= 1 && Z % B == 0, "bad exact keyswitch accumulator rows");
        k_inv_rows_sm<LOGR><<<dim3((Z + S - 1) / S, n1, 2 * G), 256, 0, s>>>(ain, aout, (int)c.logn, Z, B, sbase,
                                                                             c.d_primes, gate);
    }
    RCK_CHECK_HIP(hipGetLastError());
}

static void launch_rows_sm(const Context& c, bool fwd, const double* din, uint64_t* dout, const double* ain,
                           double* aout, int Z, int D, int Dp, int B, int G, hipStream_t s,
                           const int* gate = nullptr) {
    RCK_REQUIRE(Z >= 1 && 2 * G <= 65535, "bad exact keyswitch rows");
    int n1, n2;
    split_logn((int)c.logn, &n1, &n2);
    switch (n2) {
        case 6: rows_sm_t<6>(c, fwd, din, dout, ain, aout, Z, D, Dp, B, G, s, gate); break;
        case 7: rows_sm_t<7>(c, fwd, din, dout, ain, aout, Z, D, Dp, B, G, s, gate); break;
        case 8: rows_sm_t<8>(c, fwd, din, dout, ain, aout, Z, D, Dp, B, G, s, gate); break;
        default: RCK_REQUIRE(false, "unsupported ring degree");
    }
}

// ---------------------------------------------------------------------------------------
// The key's matrix-core form: byte planes.  Per slot (s, x), records of Dk bytes (the key's digits
// rounded up to 8; a lane's 16-byte operand load is 8-byte aligned, which the hardware serves) in the order N-tile nt (16 key columns; the last one w = cols - 16 nt wide), byte
// plane v (7), column j: record = byte v of every key digit of column 16 nt + j (balanced signed bytes;
// the rounding digits are zero).  A lane's operand (16 digits of one plane) is 16 consecutive bytes of a
// record, so a wave's load covers the w records of one plane.  64 bytes of padding follow the last slot
// (a lane's 16 bytes may run past the last record).
// ---------------------------------------------------------------------------------------
size_t key_mfma_bytes(const Context& c, int digits, int cols) {
    return (size_t)2 * c.n * 7 * cols * (8 * ((digits + 7) / 8)) + 64;
}

__device__ __forceinline__ long long key_plane_byte(int n, int cols, int Dk, int s, long long x, int col, int d, int v) {
    const int nt = col >> 4, j = col & 15, w = min(16, cols - 16 * nt);
    return ((long long)s * n + x) * 7LL * cols * Dk + (7LL * 16 * nt + (long long)v * w + j) * Dk + d;
}

// digits d0 .. d0 + nd - 1 of the key (nd <= 8, d0 a multiple of 8): stage [nd][cols][2][N] (the NTT forms
// under s1, s2, reduced) -> their bytes in the planes, one 8-byte word per (slot, column, plane); the bytes
// of digits past nd are written as zeros
__global__ void k_key_chunk_to_mfma(const double* __restrict__ stage, uint8_t* __restrict__ key, int n, int cols,
                                    int Dk, int d0, int nd) {
    const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int col = blockIdx.y, s = blockIdx.z;
    const long long dstride = (long long)cols * 2 * n;
    uint64_t w[8];
#pragma unroll
    for (int t = 0; t < 8; t++)
        w[t] = t < nd ? sbyte_word(f64_to_i64(stage[t * dstride + ((long long)col * 2 + s) * n + x])) : 0ull;
#pragma unroll
    for (int v = 0; v < 7; v++) {
        uint64_t o = 0;
#pragma unroll
        for (int t = 0; t < 8; t++) o |= ((w[t] >> (8 * v)) & 0xffull) << (8 * t);
        *reinterpret_cast<uint64_t*>(key + key_plane_byte(n, cols, Dk, s, x, col, d0, v)) = o;
    }
}

// the inverse: digit d's bytes -> stage [cols][2][N] (centred doubles)
__global__ void k_key_digit_from_mfma(const uint8_t* __restrict__ key, double* __restrict__ stage, int n, int cols,
                                      int Dk, int d) {
    const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int col = blockIdx.y, s = blockIdx.z;
    uint64_t w = 0;
#pragma unroll
    for (int v = 0; v < 7; v++) w |= (uint64_t)key[key_plane_byte(n, cols, Dk, s, x, col, d, v)] << (8 * v);
    stage[((long long)col * 2 + s) * n + x] = (double)sbyte_value(w);
}

// Inverse column pass of the accumulators fused with the recombination: a workgroup transforms the
// same S columns of one accumulator (digit group g, ciphertext b, key column jc) under s1 and then s2
// (each exactly as k_inv_cols, with the N^{-1} scaling) and writes V mod p_j, V the group's exact sum
// (crt_mod_q, |V mod p_j| <= p_j/2 + 1 as a double).  acc: [Z][2][N] after the inverse row pass,
// Z = G B 2E; out: [Z][N].  Half the bytes of writing both residues for the recombination to read back.
// addr (nullable): for group 0 and the Q limbs j < nq, the addend d_comp limb j of ciphertext b (after the
// inverse row stages: addr + b addr_ps + (comp nq + j) N) gets the same inverse column pass under p_j and is
// added times P (pmodq[j] = P mod p_j): mod-down by P then returns keyswitch + d_comp.
template <int LOGR>
__global__ __launch_bounds__(256) void k_inv_cols_crt(const double* __restrict__ acc, double* __restrict__ out,
                                                      int logn, int E, int nq, int L, int sbase, SpecConst SK,
                                                      const double* __restrict__ spmod,
                                                      const DevPrime* __restrict__ primes,
                                                      const int* __restrict__ gate, const double* __restrict__ addr,
                                                      long long addr_ps, int B, const double* __restrict__ pmodq) {
    constexpr int R = 1 << LOGR, TPS = R / 16, S = 256 / TPS;
    __shared__ double lds[R * S];
    if (gate != nullptr && *gate == 0) return;
    const int n = 1 << logn, N2 = n >> LOGR;
    const int tid = threadIdx.x, col = tid % S, tau = tid / S;
    const int r = blockIdx.x * S + col;
    const long long z = blockIdx.y;
    const int jc = (int)(z % (2 * E)), comp = jc / E, j = jc % E;
    // element e of the thread's column: a wave-uniform row base (scalar registers) plus the 32-bit lane
    // offset lane0 + N2 e, so each load holds one address register, not two
    const uint32_t lane0 = (uint32_t)(r + N2 * 16 * tau);
    auto ld_lane = [](const double* base, uint32_t idx) -> double {
        typedef const __attribute__((address_space(1))) char gchar;
        const uint64_t v = (uint64_t)base;
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
        gchar* p = (gchar*)(((uint64_t)hi << 32) | lo);
        return *(const __attribute__((address_space(1))) double*)(p + (uint64_t)(idx * 8u));  // idx < N <= 2^17
    };
    const int pj = prime_of(j, nq, L);
    // both primes' columns are loaded up front (the second in flight during the first transform); the
    // addend's after the recombination, when the second prime's registers are free (with the scalar row
    // bases and the branch-free recombination, 128 registers: 4 waves per SIMD instead of 2)
    const int g = (int)(z / (2LL * E * B));
    const bool fold = addr != nullptr && g == 0 && j < nq;
    double a1[16], a2[16], a3[16];
#pragma unroll
    for (int e = 0; e < 16; e++) a1[e] = ld_lane(acc + (z * 2) * n, lane0 + (uint32_t)(N2 * e));
#pragma unroll
    for (int e = 0; e < 16; e++) a2[e] = ld_lane(acc + (z * 2 + 1) * n, lane0 + (uint32_t)(N2 * e));
    auto column = [&](const DevPrime& Ps, double (&v)[16]) {
        // this transform's twiddle loads stay here (hoisted above the previous transform, each prime's
        // would hold 30 more registers)
        asm volatile("" ::: "memory");
        inv_phase_b<LOGR>(v, Ps.iv, 1, tau, Ps.q, Ps.qinv);
        // LDS-only barriers: the next column's loads stay in flight across this exchange
        lds_barrier();  // the previous transform's reads of lds are done
#pragma unroll
        for (int e = 0; e < 16; e++) lds[(16 * tau + e) * S + col] = v[e];
        lds_barrier();
#pragma unroll
        for (int e = 0; e < 16; e++) v[e] = lds[(tau + TPS * e) * S + col];
        inv_phase_a<kTwScalar>(v, Ps.iv, 1, Ps.q, Ps.qinv);
    };
    const DevPrime P1 = primes[sbase], P2 = primes[sbase + 1];
    column(P1, a1);
#pragma unroll
    for (int e = 0; e < 16; e++) a1[e] = canon(mulmod(a1[e], P1.ninv, P1.q, P1.qinv), P1.q, P1.qinv);
    column(P2, a2);
#pragma unroll
    for (int e = 0; e < 16; e++) a2[e] = canon(mulmod(a2[e], P2.ninv, P2.q, P2.qinv), P2.q, P2.qinv);
    const DevPrime Pj = primes[pj];
#pragma unroll
    for (int e = 0; e < 16; e++) a1[e] = crt_mod_q(a1[e], a2[e], SK, Pj.q, Pj.qinv, spmod[2 * pj], spmod[2 * pj + 1]);
    if (fold) {
        const int b = (int)((z / (2 * E)) % B);
        const double* dsrc = addr + (long long)b * addr_ps + ((long long)comp * nq + j) * n;
        asm volatile("" ::: "memory");  // keeps these loads here (hoisted, they would hold 32 more registers)
#pragma unroll
        for (int e = 0; e < 16; e++) a3[e] = ld_lane(dsrc, lane0 + (uint32_t)(N2 * e));
        column(Pj, a3);
        const double pm = reduce(mulmod(Pj.ninv, pmodq[j], Pj.q, Pj.qinv), Pj.q, Pj.qinv);  // N^{-1} P mod p_j
#pragma unroll
        for (int e = 0; e < 16; e++) a1[e] = reduce(a1[e] + mulmod(a3[e], pm, Pj.q, Pj.qinv), Pj.q, Pj.qinv);
    }
    double* o = out + z * n;
#pragma unroll
    for (int e = 0; e < 16; e++) o[r + (long long)N2 * (tau + TPS * e)] = a1[e];
}

static void launch_inv_cols_crt(const Context& c, const double* acc, double* out, int nq, int B, int G, hipStream_t s,
                                const int* gate = nullptr, const double* addr = nullptr, long long addr_ps = 0) {
    const int E = nq + (int)c.K;
    const SpecConst SK = spec_const(c);
    int n1, n2;
    split_logn((int)c.logn, &n1, &n2);
    const int S = 256 / ((1 << n1) / 16);
    const long long Z = (long long)G * B * 2 * E;
    RCK_REQUIRE(Z <= 65535, "exact keyswitch batch too large for one recombination launch");
    const dim3 grid((unsigned)((1 << n2) / S), (unsigned)Z);
    ProfScope prof(s, gate ? "inv_cols_crt_gated" : "inv_cols_crt");
    switch (n1) {
#define RCK_IC(LR)                                                                                           \
    case LR:                                                                                                 \
        k_inv_cols_crt<LR><<<grid, 256, 0, s>>>(acc, out, (int)c.logn, E, nq, (int)c.L, (int)(c.L + c.K), SK, \
                                                c.d_spmod, c.d_primes, gate, addr, addr_ps, B, c.d_pmodq);   \
        break;
        RCK_IC(6)
        RCK_IC(7)
        RCK_IC(8)
#undef RCK_IC
        default: RCK_REQUIRE(false, "unsupported ring degree");
    }
    RCK_CHECK_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// Group sum and mod-down: one thread per (coefficient x, component, ciphertext).
//   x_j = sum_g accq[g][b][comp][j] (each the recombined group sum mod p_j) for every eval limb j;
//   the aux part through Garner (p_garner), then for each Q limb i
//   out_i = (x_i - centred(x_P) mod q_i) P^{-1} mod q_i (+ addend)          (keyswitching.rs:278-322)
// addends: add0 for component 0 (read through X -> X^g when ginv != 0, the rotated c0 of a rotation),
// add1 for component 1, each at element offsets relative to ciphertext b (stride add_ps).
// ---------------------------------------------------------------------------------------
struct ExactOut {
    uint64_t* p[32];
};

template <int KT>
__global__ __launch_bounds__(256) void k_crt_moddown(const double* __restrict__ accq, int n, int nq, int E, int B,
                                                     int G, int L, int T,
                                                     const uint64_t* __restrict__ add0,
                                                     const uint64_t* __restrict__ add1, long long add_ps,
                                                     unsigned long long ginv, ExactOut outs, int b0,
                                                     const DevPrime* __restrict__ primes,
                                                     const double* __restrict__ invq,
                                                     const uint32_t* __restrict__ pwords,
                                                     const double* __restrict__ pmodq,
                                                     const double* __restrict__ pinvq,
                                                     const double* __restrict__ pmod_aux,
                                                     const int* __restrict__ gate, int want) {
    if (gate != nullptr && (*gate != 0) != (want != 0)) return;
    const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int z = blockIdx.y, bl = z >> 1, comp = z & 1, b = b0 + bl;
    const int cols = 2 * E;
    // x_j for eval limb j (prime index pj), summed over the digit groups
    auto xj = [&](int j) -> double {
        double v = 0.0;
        for (int g = 0; g < G; g++) v += accq[(((long long)g * B + b) * cols + comp * E + j) * n + x];
        return v;  // |v| <= G (q/2 + 1)
    };
    double r[KT], m[KT];
#pragma unroll
    for (int k = 0; k < KT; k++) {
        const DevPrime Pk = primes[L + k];
        r[k] = canon(xj(nq + k), Pk.q, Pk.qinv);
    }
    const bool gt = p_garner(r, KT, L, T, m, primes, invq, pwords);
    const uint64_t* add = comp == 0 ? add0 : add1;
    uint64_t* o = outs.p[bl] + (long long)comp * nq * n;
    for (int i = 0; i < nq; i++) {
        const DevPrime Pi = primes[i];
        const double q = Pi.q, qinv = Pi.qinv;
        const double xi = reduce(xj(i), q, qinv);
        double cv = reduce(m[KT - 1], q, qinv);
#pragma unroll
        for (int kk = KT - 2; kk >= 0; kk--)
            cv = reduce(mulmod(cv, pmod_aux[(long long)kk * L + i], q, qinv) + m[kk], q, qinv);
        if (gt) cv -= pmodq[i];
        double res = mulmod(reduce(xi - cv, q, qinv), pinvq[i], q, qinv);
        if (add) {
            const uint64_t* ad = add + (long long)b * add_ps + (long long)i * n;
            if (comp == 0 && ginv) {
                const unsigned long long k = ((unsigned long long)x * ginv) & (2ull * n - 1);
                const uint64_t v = ad[k >= (unsigned long long)n ? (long long)k - n : (long long)k];
                res += k >= (unsigned long long)n && v != 0 ? q - (double)v : (double)v;
            } else {
                res += (double)ad[x];
            }
        }
        o[(long long)i * n + x] = f64_to_u64(canon(res, q, qinv));
    }
}

// gate (nullable, device): the launch does nothing unless (*gate != 0) == want
static void launch_crt_moddown(const Context& c, const double* accq, int nq, int B, int G, const uint64_t* add0,
                               const uint64_t* add1, long long add_ps, unsigned long long ginv,
                               const std::vector<uint64_t*>& outs, hipStream_t s, const int* gate = nullptr,
                               int want = 0) {
    const int E = nq + (int)c.K;
    for (int b0 = 0; b0 < B; b0 += 32) {
        const int bc = std::min(32, B - b0);
        ExactOut op{};
        for (int b = 0; b < bc; b++) op.p[b] = outs[b0 + b];
        dim3 grid((c.n + 255) / 256, 2 * bc);
        ProfScope ps(s, gate && want ? "crt_moddown_gated" : "crt_moddown");
#define RCK_CM(KK)                                                                                             \
    k_crt_moddown<KK><<<grid, 256, 0, s>>>(accq, (int)c.n, nq, E, B, G, (int)c.L, (int)(c.L + c.K), add0, add1,    \
                                           add_ps, ginv, op, b0, c.d_primes, c.d_invq, c.d_pwords, c.d_pmodq,   \
                                           c.d_pinvq, c.d_pmod_aux, gate, want)
        switch (c.K) {
            case 1: RCK_CM(1); break;
            case 2: RCK_CM(2); break;
            case 3: RCK_CM(3); break;
            default: RCK_CM(4); break;
        }
#undef RCK_CM
        RCK_CHECK_HIP(hipGetLastError());
    }
}

// Digits per group: the largest G with G N 2^29 max_j (p_j - 1)/2 <= (s1 s2 - 1)/2 over the eval
// primes of `nq` Q limbs and the aux basis.
int exact_group_size(const Context& c, int nq) {
    using u128 = unsigned __int128;
    uint64_t maxhalf = 0;
    for (int j = 0; j < nq; j++) maxhalf = std::max(maxhalf, (c.hp[j].q - 1) / 2);
    for (uint32_t k = 0; k < c.K; k++) maxhalf = std::max(maxhalf, (c.hp[c.L + k].q - 1) / 2);
    const u128 lim = ((u128)c.sp[0] * c.sp[1] - 1) / 2;
    const u128 per = (u128)c.n * ((u128)1 << 29) * maxhalf;
    const u128 g = lim / per;
    RCK_REQUIRE(g >= 1, "auxiliary primes too small for one digit");
    return (int)std::min<u128>(g, 1 << 20);
}

// One group for a batch whose digits allow it.  A ciphertext's one-group sums satisfy |V_j| <= S maxhalf,
// S = sum_i ||d_i||_1 (the L1 norm of all its digits), so they are exact when S <= floor(lim / maxhalf),
// lim = (s1 s2 - 1)/2.  The digits of a uniform residue average 2^28 in magnitude (S near levels N 2^28),
// so when levels N 2^28 (5/4) is below that threshold the keyswitch runs one group, the decomposition
// records every ciphertext's S, and a batch with some S above the threshold is recomputed with the
// worst-case groups (the gate is set on the device; the launches of the other path do nothing).
struct ExactPlan {
    int G, gsize;    // grouping of the main path
    int Gw, gsw;     // worst-case grouping (the recomputation when gated)
    bool gated;      // main path one group, checked per batch
    double thr;      // the threshold on S
};

static ExactPlan exact_plan(const Context& c, int nq, int levels) {
    using u128 = unsigned __int128;
    ExactPlan pl{};
    const int gw = exact_group_size(c, nq);
    pl.Gw = (levels + gw - 1) / gw;
    pl.gsw = (levels + pl.Gw - 1) / pl.Gw;  // balanced groups
    pl.G = pl.Gw;
    pl.gsize = pl.gsw;
    if (pl.Gw == 1 || c.n % 64 != 0) return pl;
    uint64_t maxhalf = 0;
    for (int j = 0; j < nq; j++) maxhalf = std::max(maxhalf, (c.hp[j].q - 1) / 2);
    for (uint32_t k = 0; k < c.K; k++) maxhalf = std::max(maxhalf, (c.hp[c.L + k].q - 1) / 2);
    const u128 lim = ((u128)c.sp[0] * c.sp[1] - 1) / 2;
    const u128 thr = lim / maxhalf;
    const u128 typical = (u128)levels * c.n * ((u128)1 << 28) * 5 / 4;
    if (typical > thr) return pl;
    pl.gated = true;
    pl.G = 1;
    pl.gsize = levels;
    pl.thr = thr >= ((u128)1 << 53) ? 9007199254740992.0 : (double)(uint64_t)thr;  // exact below 2^53
    return pl;
}

void exact_grouping(const Context& c, int nq, int levels, int* gsize, int* G) {
    const ExactPlan pl = exact_plan(c, nq, levels);
    *gsize = pl.gsize;
    *G = pl.G;
}

// gate[0] = 1 when some ciphertext's S = sum of l1[b][0 .. nblk) exceeds thr, else 0 (every partial
// is an integer and every sum stays below 2^53, so the sums are exact in any order).  Wave w of the
// 16 sums ciphertexts w, w + 16, ...
__global__ __launch_bounds__(1024) void k_exact_gate(const double* __restrict__ l1, int B, int nblk, double thr,
                                                     int* __restrict__ gate) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int flag = 0;
    for (int b = wave; b < B; b += 16) {
        double t = 0.0;
        for (int i = lane; i < nblk; i += 64) t += l1[(long long)b * nblk + i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        flag |= t > thr;
    }
    flag = __syncthreads_or(flag);
    if (threadIdx.x == 0) gate[0] = flag;
}

bool exact_keyswitch_available(const Context& c, uint32_t aux) { return c.S == 2 && aux > 0 && c.logn >= 12; }

// Converts a key from its NTT form mod the eval primes (data) to the exact-convolution form (mdata, the
// matrix-core byte planes of its centred limbs' NTT forms under s1 and s2); data is released unless keep_data.
void key_make_exact(Context& c, KeyDev* k, bool keep_data) {
    require_fp64(c);
    RCK_REQUIRE(k && k->data, "null key");
    RCK_REQUIRE(exact_keyswitch_available(c, k->aux), "exact-convolution keys need hybrid keys at N >= 2^12");
    RCK_REQUIRE(k->digits <= 64, "exact-convolution keys hold at most 64 digits");
    const long long n = c.n;
    const int E = (int)k->limbs, nq = (int)k->level + 1, cols = 2 * E;
    const int Dk = 8 * (((int)k->digits + 7) / 8);
    const long long P = (long long)E * n;
    const size_t bytes = key_mfma_bytes(c, (int)k->digits, cols);
    uint8_t* md = nullptr;
    uint64_t* coeff = nullptr;
    double* stage = nullptr;   // 2 P doubles (inverse NTT input), then 4 P (the centred limbs mod s1, s2)
    double* stage8 = nullptr;  // the NTT forms of 8 digits, [8][cols][2][N]
    try {
        RCK_CHECK_HIP(hipMalloc(&md, bytes));
        RCK_CHECK_HIP(hipMemset(md, 0, bytes));  // the padding stays zero
        RCK_CHECK_HIP(hipMalloc(&coeff, (size_t)2 * P * 8));
        RCK_CHECK_HIP(hipMalloc(&stage, (size_t)4 * P * 8));
        RCK_CHECK_HIP(hipMalloc(&stage8, (size_t)8 * 4 * P * 8));
        const PrimeMap pm{nq, (int)c.L, 0};
        const PrimeMap ps{0, (int)(c.L + c.K), 0};
        for (uint32_t d0 = 0; d0 < k->digits; d0 += 8) {
            const int nd = std::min(8, (int)k->digits - (int)d0);
            for (int t = 0; t < nd; t++) {
                RCK_CHECK_HIP(hipMemcpy(stage, k->data + (size_t)(d0 + t) * 2 * P, 2 * P * 8,
                                        hipMemcpyDeviceToDevice));
                launch_inv(c, stage, n, P, kOutU64, coeff, n, P, E, 2, pm, nullptr);
                k_key_to_special<<<dim3((c.n + 255) / 256, E, 2), 256>>>(coeff, stage, (int)c.n, E, nq, (int)c.L,
                                                                        (int)(c.L + c.K), c.d_primes);
                RCK_CHECK_HIP(hipGetLastError());
                launch_fwd(c, kInF64, stage, n, 2 * n, stage8 + (size_t)t * 4 * P, n, 2 * n, 2, cols, ps, nullptr);
            }
            k_key_chunk_to_mfma<<<dim3((c.n + 255) / 256, cols, 2), 256>>>(stage8, md, (int)c.n, cols, Dk, (int)d0,
                                                                         nd);
            RCK_CHECK_HIP(hipGetLastError());
        }
        RCK_CHECK_HIP(hipDeviceSynchronize());
    } catch (...) {
        if (md) (void)hipFree(md);
        if (coeff) (void)hipFree(coeff);
        if (stage) (void)hipFree(stage);
        if (stage8) (void)hipFree(stage8);
        throw;
    }
    (void)hipFree(coeff);
    (void)hipFree(stage);
    (void)hipFree(stage8);
    if (!keep_data) {
        RCK_CHECK_HIP(hipFree(k->data));
        k->data = nullptr;
    }
    k->mdata = md;
}

// The crate's coefficient form of digit d of an exact-convolution key: out [2][E][N] (device).
void key_export_exact_digit(Context& c, const KeyDev* k, uint32_t d, uint64_t* out) {
    const long long n = c.n;
    const int E = (int)k->limbs, nq = (int)k->level + 1, cols = 2 * E;
    const int Dk = 8 * (((int)k->digits + 7) / 8);
    const long long P = (long long)E * n;
    double* stage = nullptr;
    uint64_t* coeff = nullptr;
    try {
        RCK_CHECK_HIP(hipMalloc(&stage, (size_t)2 * P * 2 * 8));
        RCK_CHECK_HIP(hipMalloc(&coeff, (size_t)2 * P * 2 * 8));
        k_key_digit_from_mfma<<<dim3((c.n + 255) / 256, cols, 2), 256>>>(k->mdata, stage, (int)c.n, cols, Dk,
                                                                         (int)d);
        RCK_CHECK_HIP(hipGetLastError());
        launch_inv(c, stage, n, 2 * n, kOutU64, coeff, n, 2 * n, 2, cols, PrimeMap{0, (int)(c.L + c.K), 0}, nullptr);
        k_key_from_special<<<dim3((c.n + 255) / 256, E, 2), 256>>>(coeff, out, (int)c.n, E, nq, (int)c.L,
                                                                  spec_const(c), c.d_spmod, c.d_primes);
        RCK_CHECK_HIP(hipGetLastError());
        RCK_CHECK_HIP(hipDeviceSynchronize());
    } catch (...) {
        if (stage) (void)hipFree(stage);
        if (coeff) (void)hipFree(coeff);
        throw;
    }
    (void)hipFree(stage);
    (void)hipFree(coeff);
}

// Keyswitch of B polynomials (stride in_ps) at `limbs` Q limbs with an exact-convolution key, `levels`
// digits; adds add0/add1 (or the row-form addends addr) and writes outs[b] = [c0][c1] (keyswitch_batch's
// contract).  spw: slots per wave of the multiply-accumulate.
void keyswitch_exact(Context& c, const uint64_t* poly, long long in_ps, int B, int limbs, const KeyDev* key, int levels,
                     const uint64_t* add0, const uint64_t* add1, long long add_ps, const std::vector<uint64_t*>& outs,
                     hipStream_t s, unsigned long long ginv, int spw, const double* addr, long long addr_ps) {
    require_fp64(c);
    const long long n = c.n;
    const int E = limbs + (int)key->aux, cols = 2 * E;
    const int Dp = 8 * ((levels + 7) / 8);
    RCK_REQUIRE(levels <= (int)key->digits && cols == 2 * (int)key->limbs, "exact keyswitch key does not fit");
    const ExactPlan pl = exact_plan(c, limbs, levels);
    const int G = pl.G, Gw = pl.Gw;
    const int Zc = B * cols;
    int32_t* digits = static_cast<int32_t*>(c.workspace(5, (size_t)B * levels * n * 4));
    // workspace slots 30-36 (slots 20-26 hold the BSGS buffers of a transform whose rotations come here);
    // the accumulators are sized for the worst-case groups, which the recomputation reuses
    double* dn = static_cast<double*>(c.workspace(30, (size_t)B * levels * 2 * n * 8));
    uint64_t* dsb = static_cast<uint64_t*>(c.workspace(35, (size_t)2 * n * B * Dp * 8));
    double* accm = static_cast<double*>(c.workspace(36, (size_t)Gw * 2 * n * Zc * 8));
    double* acc = static_cast<double*>(c.workspace(31, (size_t)Gw * Zc * 2 * n * 8));
    double* accq = static_cast<double*>(c.workspace(32, (size_t)Gw * Zc * n * 8));
    double* l1 = nullptr;
    int* gate = nullptr;
    if (pl.gated) {
        l1 = static_cast<double*>(c.workspace(33, (size_t)B * (n / 64) * 8));
        gate = static_cast<int*>(c.workspace(34, 256));
    }
    launch_decompose(c, poly, limbs, levels, digits, s, B, in_ps, ginv, l1);
    if (pl.gated) {
        ProfScope prof(s, "exact_gate");
        k_exact_gate<<<1, 1024, 0, s>>>(l1, B, (int)(n / 64), pl.thr, gate);
        RCK_CHECK_HIP(hipGetLastError());
    }
    const PrimeMap ps{0, (int)(c.L + c.K), 0};
    // every digit under s1 and s2: column pass (each digit row read once for both primes), then the row
    // stages written slot-major as signed-byte words
    launch_digit_cols(c, digits, dn, 2, B * levels, ps, s);
    launch_rows_sm(c, true, dn, dsb, nullptr, nullptr, B * levels, levels, Dp, B, 1, s);
    launch_mac_mfma(c, dsb, key->mdata, (int)key->digits, accm, B, levels, Dp, cols, pl.gsize, G, spw, s);
    // inverse rows under s1, s2 (back to limb rows); inverse columns fused with the recombination (one
    // value per accumulator and coefficient: the group's exact sum mod p_j); then the group sum and
    // mod-down (skipped when the gate is set)
    launch_rows_sm(c, false, nullptr, nullptr, accm, acc, Zc, 0, 0, B, G, s);
    launch_inv_cols_crt(c, acc, accq, limbs, B, G, s, nullptr, addr, addr_ps);
    launch_crt_moddown(c, accq, limbs, B, G, add0, add1, add_ps, ginv, outs, s, gate, 0);
    if (pl.gated) {
        // some ciphertext's digits exceed the one-group bound: the batch again with the worst-case groups
        launch_mac_mfma(c, dsb, key->mdata, (int)key->digits, accm, B, levels, Dp, cols, pl.gsw, Gw, spw, s,
                        gate);
        launch_rows_sm(c, false, nullptr, nullptr, accm, acc, Zc, 0, 0, B, Gw, s, gate);
        launch_inv_cols_crt(c, acc, accq, limbs, B, Gw, s, gate, addr, addr_ps);
        launch_crt_moddown(c, accq, limbs, B, Gw, add0, add1, add_ps, ginv, outs, s, gate, 1);
    }
}

}  // namespace rck
