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
// key is stored as the s1/s2 NTT forms of its centred limbs ([digit][comp][limb][s][N]), the products
// are a per-coefficient multiply-accumulate over the digits of a group (every digit value reused for
// all 2 (l + K) key columns, every key value for all ciphertexts of a batch), and the accumulators are
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
    const bool neg = t > K.half_t || (t == K.half_t && a1 > K.half_a);
    double r = reduce(a1, q, qinv) + mulmod(reduce(t, q, qinv), p1m, q, qinv);  // <= q/2 + 1 + 0.97q
    if (neg) r -= p12m;
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
// Multiply-accumulate: for coefficient x, auxiliary prime s and digit group g,
//   acc[g][b][jc][s][x] = sum_{i in g} dn[b][i][s][x] * key[i][jc][s][x]  mod s,  jc = comp * E + j.
// A workgroup is four waves over 64 consecutive coefficients; wave w takes BT ciphertexts of its
// b-tile and JT key columns of its j-tile (WB waves along the ciphertexts, 4 / WB along the
// columns), so the waves of a workgroup read the same key values (served by the CU's caches) or the
// same digit values.  Work items are numbered key-column group fastest, then coefficient block,
// ciphertext group, prime and digit group, and each XCD takes a contiguous range of them (workgroup
// w runs on XCD w % 8), so consecutive items on an XCD share their digit tiles in its L2.
// ---------------------------------------------------------------------------------------
// One wave's tile: ciphertexts b0 .. b0+BT-1 by key columns c0 .. c0+JT-1 over digits i0 .. i1-1 of
// auxiliary prime s, for the 64 coefficients from x0.
template <int BT, int JT>
__device__ __forceinline__ void mac_tile(const double* __restrict__ dn, const double* __restrict__ key,
                                         double* __restrict__ acc, int n, int B, int D, int cols, int b0, int c0,
                                         int s, int g, int i0, int i1, long long x0, int lane, double q, double qinv) {
    const long long n2 = 2LL * n;
    // rows past the batch or past the key columns read the last valid one (their products are dropped);
    // every load is one of these wave-uniform offsets from a kernel argument plus the lane
    long long dbase[BT], kbase[JT];
#pragma unroll
    for (int bt = 0; bt < BT; bt++) dbase[bt] = ((long long)min(b0 + bt, B - 1) * D * 2 + s) * n + x0;
#pragma unroll
    for (int jt = 0; jt < JT; jt++) kbase[jt] = (long long)min(c0 + jt, cols - 1) * n2 + s * n + x0;
    const long long kstep = (long long)cols * n2;  // key digit i -> i + 1
    double a[BT][JT];
#pragma unroll
    for (int bt = 0; bt < BT; bt++)
#pragma unroll
        for (int jt = 0; jt < JT; jt++) a[bt][jt] = 0.0;
    auto load = [&](int t, double (&d)[BT], double (&k)[JT]) {
#pragma unroll
        for (int bt = 0; bt < BT; bt++) d[bt] = (dn + dbase[bt] + t * n2)[lane];
#pragma unroll
        for (int jt = 0; jt < JT; jt++) k[jt] = (key + kbase[jt] + t * kstep)[lane];
        // keeps the loads ahead of the multiplies that follow (the compiler would otherwise sink them to
        // their first use); neither barrier reads a register, so neither adds a wait
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
    };
    auto mac = [&](const double (&d)[BT], const double (&k)[JT]) {
#pragma unroll
        for (int bt = 0; bt < BT; bt++)
#pragma unroll
            for (int jt = 0; jt < JT; jt++) a[bt][jt] += mulmod(d[bt], k[jt], q, qinv);
    };
    // reduced operands: each product is <= 0.57 s, so eight of them on a reduced accumulator stay
    // below 5.1 s < 2^53 / s
    auto fold = [&]() {
#pragma unroll
        for (int bt = 0; bt < BT; bt++)
#pragma unroll
            for (int jt = 0; jt < JT; jt++) a[bt][jt] = reduce(a[bt][jt], q, qinv);
    };
    // two register buffers, unrolled by two: the digit after the one being multiplied is always in
    // flight (its index clamped to the group, so every pass issues the same loads and the only wait
    // is on loads issued one digit earlier)
    double dA[BT], kA[JT], dB[BT], kB[JT];
    load(i0, dA, kA);
    for (int t = i0; t < i1; t += 2) {
        load(min(t + 1, i1 - 1), dB, kB);
        mac(dA, kA);
        load(min(t + 2, i1 - 1), dA, kA);
        // an odd group's last pass has no second digit: its key operands become 0, and every product
        // with 0 is exactly 0
        const bool none = t + 1 == i1;
#pragma unroll
        for (int jt = 0; jt < JT; jt++) kB[jt] = none ? 0.0 : kB[jt];
        mac(dB, kB);
        if (((t - i0) & 7) == 6) fold();
    }
    fold();
#pragma unroll
    for (int bt = 0; bt < BT; bt++) {
        if (b0 + bt >= B) continue;
#pragma unroll
        for (int jt = 0; jt < JT; jt++) {
            if (c0 + jt >= cols) continue;
            acc[((((long long)g * B + b0 + bt) * cols + c0 + jt) * 2 + s) * n + x0 + lane] = a[bt][jt];
        }
    }
}

// The last column group of a row may hold fewer columns than JT: it runs the narrowest tile that still
// covers them, so no products are computed for columns past the key's.
template <int BT, int JT>
__device__ __forceinline__ void mac_tile_fit(const double* __restrict__ dn, const double* __restrict__ key,
                                             double* __restrict__ acc, int n, int B, int D, int cols, int b0, int c0,
                                             int s, int g, int i0, int i1, long long x0, int lane, double q,
                                             double qinv) {
    if constexpr (JT > 4) {
        if (cols - c0 <= JT / 2)
            return mac_tile_fit<BT, JT / 2>(dn, key, acc, n, B, D, cols, b0, c0, s, g, i0, i1, x0, lane, q, qinv);
    }
    mac_tile<BT, JT>(dn, key, acc, n, B, D, cols, b0, c0, s, g, i0, i1, x0, lane, q, qinv);
}

// The same tile when the four waves of a workgroup take four b-tiles of the same key columns: the
// workgroup's key tile (JT columns x 64 coefficients of one digit) is copied into LDS once, by
// global_load_lds (each lane 16 bytes, a wave two columns), and every wave reads it from there.
// The waves' own vector loads are then only their digits, BT per digit instead of BT + JT: the
// texture address and data units, which each vector load instruction occupies, bound the kernel.
// Per digit t: wait for this wave's loads of t, barrier (every wave's copy of tile t has landed and
// every read of the other buffer is done), start the copy of tile t + 1 and the digit loads of t + 1,
// then read tile t and multiply.
constexpr int kWaitVm0Exact = 0x0F70;  // s_waitcnt vmcnt(0), the other counters at their maximum

template <int BT, int JT>
__device__ __forceinline__ void mac_tile_lds(const double* __restrict__ dn, const double* __restrict__ key,
                                             double* __restrict__ acc, double* kt, int n, int B, int D, int cols,
                                             int b0, int c0, int s, int g, int i0, int i1, long long x0, int wave,
                                             int lane, double q, double qinv) {
    const long long n2 = 2LL * n;
    long long dbase[BT];
#pragma unroll
    for (int bt = 0; bt < BT; bt++) dbase[bt] = ((long long)min(b0 + bt, B - 1) * D * 2 + s) * n + x0;
    const long long kstep = (long long)cols * n2;  // key digit i -> i + 1
    // column pair m of the tile (m = wave, wave + 4, ..): lanes 0-31 copy column c0 + 2m, lanes 32-63
    // column c0 + 2m + 1 (columns past the key's read its last one; their products are dropped)
    const int half = lane >> 5, sub = 2 * (lane & 31);
    auto copy = [&](int t, int buf) {
#pragma unroll
        for (int m = wave; m < JT / 2; m += 4) {
            const int col = min(c0 + 2 * m + half, cols - 1);
            const double* src = key + (long long)t * kstep + (long long)col * n2 + (long long)s * n + x0 + sub;
            double* dst = kt + (buf * JT + 2 * m) * 64;
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                             (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
        }
    };
    auto load_d = [&](int t, double (&d)[BT]) {
#pragma unroll
        for (int bt = 0; bt < BT; bt++) d[bt] = (dn + dbase[bt] + t * n2)[lane];
    };
    double a[BT][JT];
#pragma unroll
    for (int bt = 0; bt < BT; bt++)
#pragma unroll
        for (int jt = 0; jt < JT; jt++) a[bt][jt] = 0.0;
    // reduced operands: each product is <= 0.57 s, so eight of them on a reduced accumulator stay
    // below 5.1 s < 2^53 / s
    auto fold = [&]() {
#pragma unroll
        for (int bt = 0; bt < BT; bt++)
#pragma unroll
            for (int jt = 0; jt < JT; jt++) a[bt][jt] = reduce(a[bt][jt], q, qinv);
    };
    auto step = [&](int t, const double (&cur)[BT], double (&nxt)[BT], int buf) {
        __builtin_amdgcn_s_waitcnt(kWaitVm0Exact);
        __syncthreads();
        if (t + 1 < i1) {
            copy(t + 1, buf ^ 1);
            load_d(t + 1, nxt);
        }
        // keeps the loads ahead of the multiplies (the compiler would otherwise sink them to their
        // first use); neither barrier reads a register, so neither adds a wait
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
        double k[JT];
#pragma unroll
        for (int jt = 0; jt < JT; jt++) k[jt] = kt[(buf * JT + jt) * 64 + lane];
#pragma unroll
        for (int bt = 0; bt < BT; bt++)
#pragma unroll
            for (int jt = 0; jt < JT; jt++) a[bt][jt] += mulmod(cur[bt], k[jt], q, qinv);
        if (((t - i0) & 7) == 7) fold();
    };
    double dA[BT], dB[BT];
    copy(i0, 0);
    load_d(i0, dA);
    for (int t = i0; t < i1; t += 2) {
        step(t, dA, dB, 0);
        if (t + 1 < i1) step(t + 1, dB, dA, 1);
    }
    fold();
#pragma unroll
    for (int bt = 0; bt < BT; bt++) {
        if (b0 + bt >= B) continue;
#pragma unroll
        for (int jt = 0; jt < JT; jt++) {
            if (c0 + jt >= cols) continue;
            acc[((((long long)g * B + b0 + bt) * cols + c0 + jt) * 2 + s) * n + x0 + lane] = a[bt][jt];
        }
    }
}

// the narrowest tile that still covers the last column group (uniform over the workgroup)
template <int BT, int JT>
__device__ __forceinline__ void mac_tile_lds_fit(const double* __restrict__ dn, const double* __restrict__ key,
                                                 double* __restrict__ acc, double* kt, int n, int B, int D, int cols,
                                                 int b0, int c0, int s, int g, int i0, int i1, long long x0, int wave,
                                                 int lane, double q, double qinv) {
    if constexpr (JT > 4) {
        if (cols - c0 <= JT / 2)
            return mac_tile_lds_fit<BT, JT / 2>(dn, key, acc, kt, n, B, D, cols, b0, c0, s, g, i0, i1, x0, wave, lane,
                                                q, qinv);
    }
    mac_tile_lds<BT, JT>(dn, key, acc, kt, n, B, D, cols, b0, c0, s, g, i0, i1, x0, wave, lane, q, qinv);
}

// Four waves along the ciphertexts (WB = 4): work items as in k_ks_mac_exact.  A wave whose b-tile
// lies past the batch still takes part in every barrier and copy; it stores nothing.
template <int BT, int JT>
__global__ __launch_bounds__(256) void k_ks_mac_exact_lds(const double* __restrict__ dn,
                                                          const double* __restrict__ key, double* __restrict__ acc,
                                                          int n, int B, int D, int gsize, int cols, int njg, int nxb,
                                                          int nbg, int items, int sbase,
                                                          const DevPrime* __restrict__ primes,
                                                          const int* __restrict__ gate) {
    __shared__ double kt[2 * JT * 64];
    if (gate != nullptr && *gate == 0) return;
    const int w = blockIdx.x;
    const int per = gridDim.x >> 3;
    const int item = (w & 7) * per + (w >> 3);
    if (item >= items) return;
    int r = item;
    const int jg = r % njg;
    r /= njg;
    const int xb = r % nxb;
    r /= nxb;
    const int bg = r % nbg;
    r /= nbg;
    const int s = r & 1, g = r >> 1;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int b0 = (bg * 4 + wave) * BT;
    const int c0 = jg * JT;
    const DevPrime P = primes[sbase + s];
    const int i0 = g * gsize, i1 = min(D, i0 + gsize);
    mac_tile_lds_fit<BT, JT>(dn, key, acc, kt, n, B, D, cols, b0, c0, s, g, i0, i1, (long long)xb * 64, wave, lane,
                             P.q, P.qinv);
}

template <int BT, int JT>
__global__ __launch_bounds__(256) void k_ks_mac_exact(const double* __restrict__ dn, const double* __restrict__ key,
                                                      double* __restrict__ acc, int n, int B, int D, int gsize,
                                                      int cols, int WB, int njg, int nxb, int nbg, int items,
                                                      int sbase, const DevPrime* __restrict__ primes,
                                                      const int* __restrict__ gate) {
    if (gate != nullptr && *gate == 0) return;
    const int w = blockIdx.x;
    const int per = gridDim.x >> 3;
    const int item = (w & 7) * per + (w >> 3);
    if (item >= items) return;
    int r = item;
    const int jg = r % njg;
    r /= njg;
    const int xb = r % nxb;
    r /= nxb;
    const int bg = r % nbg;
    r /= nbg;
    const int s = r & 1, g = r >> 1;
    // wave-uniform quantities live in scalar registers: every load below is a scalar base plus the
    // lane's offset, with no per-lane address arithmetic
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int WJ = 4 / WB;
    const int wb = wave % WB, wj = wave / WB;
    const int b0 = (bg * WB + wb) * BT;
    const int c0 = (jg * WJ + wj) * JT;
    if (b0 >= B || c0 >= cols) return;
    const DevPrime P = primes[sbase + s];
    const int i0 = g * gsize, i1 = min(D, i0 + gsize);
    mac_tile_fit<BT, JT>(dn, key, acc, n, B, D, cols, b0, c0, s, g, i0, i1, (long long)xb * 64, lane, P.q, P.qinv);
}

template <int BT, int JT>
static void mac_exact_t(const Context& c, const double* dn, const double* key, double* acc, int B, int D, int gsize,
                        int G, int E, hipStream_t s, const int* gate) {
    const int n = (int)c.n;
    const int nbt = (B + BT - 1) / BT;  // b-tiles; WB of them per workgroup (1, 2 or 4 waves along them)
    const int WB = nbt >= 3 ? 4 : nbt;
    const int WJ = 4 / WB;
    const int cols = 2 * E;
    const int njt = (cols + JT - 1) / JT;
    const int njg = (njt + WJ - 1) / WJ;
    const int nxb = n / 64;
    const int nbg = (B + WB * BT - 1) / (WB * BT);
    const long long items = (long long)njg * nxb * nbg * 2 * G;
    RCK_REQUIRE(items < (1LL << 30), "exact keyswitch launch too large");
    const int per = (int)((items + 7) / 8);
    ProfScope ps(s, gate ? "ks_mac_gated" : "ks_mac");
    if (WB == 4)
        k_ks_mac_exact_lds<BT, JT><<<dim3(per * 8), 256, 0, s>>>(dn, key, acc, n, B, D, gsize, cols, njg, nxb, nbg,
                                                                 (int)items, (int)(c.L + c.K), c.d_primes, gate);
    else
        k_ks_mac_exact<BT, JT><<<dim3(per * 8), 256, 0, s>>>(dn, key, acc, n, B, D, gsize, cols, WB, njg, nxb, nbg,
                                                             (int)items, (int)(c.L + c.K), c.d_primes, gate);
    RCK_CHECK_HIP(hipGetLastError());
}

template <int BT>
static void mac_exact_j(const Context& c, int jt, const double* dn, const double* key, double* acc, int B, int D,
                        int gsize, int G, int E, hipStream_t s, const int* gate) {
    if (jt == 4)
        mac_exact_t<BT, 4>(c, dn, key, acc, B, D, gsize, G, E, s, gate);
    else if (jt == 16)
        mac_exact_t<BT, 16>(c, dn, key, acc, B, D, gsize, G, E, s, gate);
    else
        mac_exact_t<BT, 8>(c, dn, key, acc, B, D, gsize, G, E, s, gate);
}

// bt ciphertexts by jt key columns per thread (bt in {1, 2, 4}, jt in {4, 8, 16}); gate (nullable,
// device): the launch does nothing unless *gate != 0
static void launch_mac_exact(const Context& c, int bt, int jt, const double* dn, const double* key, double* acc, int B,
                             int D, int gsize, int G, int E, hipStream_t s, const int* gate = nullptr) {
    RCK_REQUIRE((bt == 1 || bt == 2 || bt == 4) && (jt == 4 || jt == 8 || jt == 16), "bad exact keyswitch tile");
    if (bt == 4)
        mac_exact_j<4>(c, jt, dn, key, acc, B, D, gsize, G, E, s, gate);
    else if (bt == 2)
        mac_exact_j<2>(c, jt, dn, key, acc, B, D, gsize, G, E, s, gate);
    else
        mac_exact_j<1>(c, jt, dn, key, acc, B, D, gsize, G, E, s, gate);
}

// Inverse column pass of the accumulators fused with the recombination: a workgroup transforms the
// same S columns of one accumulator (digit group g, ciphertext b, key column jc) under s1 and then s2
// (each exactly as k_inv_cols, with the N^{-1} scaling) and writes V mod p_j, V the group's exact sum
// (crt_mod_q, |V mod p_j| <= p_j/2 + 1 as a double).  acc: [Z][2][N] after the inverse row pass,
// Z = G B 2E; out: [Z][N].  Half the bytes of writing both residues for the recombination to read back.
template <int LOGR>
__global__ __launch_bounds__(256) void k_inv_cols_crt(const double* __restrict__ acc, double* __restrict__ out,
                                                      int logn, int E, int nq, int L, int sbase, SpecConst SK,
                                                      const double* __restrict__ spmod,
                                                      const DevPrime* __restrict__ primes,
                                                      const int* __restrict__ gate) {
    constexpr int R = 1 << LOGR, TPS = R / 16, S = 256 / TPS;
    __shared__ double lds[R * S];
    if (gate != nullptr && *gate == 0) return;
    const int n = 1 << logn, N2 = n >> LOGR;
    const int tid = threadIdx.x, col = tid % S, tau = tid / S;
    const int r = blockIdx.x * S + col;
    const long long z = blockIdx.y;
    const int j = (int)(z % (2 * E)) % E;
    const int pj = prime_of(j, nq, L);
    // both primes' columns are loaded up front (the second set is in flight during the first transform)
    double a1[16], a2[16];
#pragma unroll
    for (int e = 0; e < 16; e++) a1[e] = acc[(z * 2) * n + r + (long long)N2 * (16 * tau + e)];
#pragma unroll
    for (int e = 0; e < 16; e++) a2[e] = acc[(z * 2 + 1) * n + r + (long long)N2 * (16 * tau + e)];
    auto column = [&](int s, double (&v)[16]) {
        const DevPrime Ps = primes[sbase + s];
        inv_phase_b<LOGR>(v, Ps.iv, 1, tau, Ps.q, Ps.qinv);
        // LDS-only barriers: the second prime's loads stay in flight across the first exchange
        if (s) lds_barrier();  // the first transform's reads of lds are done
#pragma unroll
        for (int e = 0; e < 16; e++) lds[(16 * tau + e) * S + col] = v[e];
        lds_barrier();
#pragma unroll
        for (int e = 0; e < 16; e++) v[e] = lds[(tau + TPS * e) * S + col];
        inv_phase_a<kTwScalar>(v, Ps.iv, 1, Ps.q, Ps.qinv);
#pragma unroll
        for (int e = 0; e < 16; e++) v[e] = canon(mulmod(v[e], Ps.ninv, Ps.q, Ps.qinv), Ps.q, Ps.qinv);
    };
    column(0, a1);
    column(1, a2);
    const DevPrime Pj = primes[pj];
    double* o = out + z * n;
#pragma unroll
    for (int e = 0; e < 16; e++)
        o[r + (long long)N2 * (tau + TPS * e)] =
            crt_mod_q(a1[e], a2[e], SK, Pj.q, Pj.qinv, spmod[2 * pj], spmod[2 * pj + 1]);
}

static void launch_inv_cols_crt(const Context& c, const double* acc, double* out, int nq, int Z, hipStream_t s,
                                const int* gate = nullptr) {
    const int E = nq + (int)c.K;
    const SpecConst SK = spec_const(c);
    int n1, n2;
    split_logn((int)c.logn, &n1, &n2);
    const int S = 256 / ((1 << n1) / 16);
    const dim3 grid((unsigned)((1 << n2) / S), (unsigned)Z);
    ProfScope prof(s, gate ? "inv_cols_crt_gated" : "inv_cols_crt");
    switch (n1) {
#define RCK_IC(LR)                                                                                           \
    case LR:                                                                                                 \
        k_inv_cols_crt<LR><<<grid, 256, 0, s>>>(acc, out, (int)c.logn, E, nq, (int)c.L, (int)(c.L + c.K), SK, \
                                                c.d_spmod, c.d_primes, gate);                                \
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

// Converts a key from its NTT form mod the eval primes (data) to the exact-convolution form (sdata);
// data is released unless keep_data.
void key_make_exact(Context& c, KeyDev* k, bool keep_data) {
    require_fp64(c);
    RCK_REQUIRE(k && k->data, "null key");
    RCK_REQUIRE(exact_keyswitch_available(c, k->aux), "exact-convolution keys need hybrid keys at N >= 2^12");
    const long long n = c.n;
    const int E = (int)k->limbs, nq = (int)k->level + 1;
    const long long P = (long long)E * n;
    double* sd = nullptr;
    uint64_t* coeff = nullptr;
    double* stage = nullptr;  // 2 P doubles (inverse NTT input), then 4 P (the centred limbs mod s1, s2)
    try {
        RCK_CHECK_HIP(hipMalloc(&sd, (size_t)k->digits * 2 * P * 2 * 8));
        RCK_CHECK_HIP(hipMalloc(&coeff, (size_t)2 * P * 8));
        RCK_CHECK_HIP(hipMalloc(&stage, (size_t)4 * P * 8));
        const PrimeMap pm{nq, (int)c.L, 0};
        const PrimeMap ps{0, (int)(c.L + c.K), 0};
        for (uint32_t d = 0; d < k->digits; d++) {
            RCK_CHECK_HIP(hipMemcpy(stage, k->data + (size_t)d * 2 * P, 2 * P * 8, hipMemcpyDeviceToDevice));
            launch_inv(c, stage, n, P, kOutU64, coeff, n, P, E, 2, pm, nullptr);
            k_key_to_special<<<dim3((c.n + 255) / 256, E, 2), 256>>>(coeff, stage, (int)c.n, E, nq, (int)c.L,
                                                                    (int)(c.L + c.K), c.d_primes);
            RCK_CHECK_HIP(hipGetLastError());
            launch_fwd(c, kInF64, stage, n, 2 * n, sd + (size_t)d * 2 * P * 2, n, 2 * n, 2, 2 * E, ps, nullptr);
        }
        RCK_CHECK_HIP(hipDeviceSynchronize());
    } catch (...) {
        if (sd) (void)hipFree(sd);
        if (coeff) (void)hipFree(coeff);
        if (stage) (void)hipFree(stage);
        throw;
    }
    (void)hipFree(coeff);
    (void)hipFree(stage);
    if (!keep_data) {
        RCK_CHECK_HIP(hipFree(k->data));
        k->data = nullptr;
    }
    k->sdata = sd;
}

// The crate's coefficient form of digit d of an exact-convolution key: out [2][E][N] (device).
void key_export_exact_digit(Context& c, const KeyDev* k, uint32_t d, uint64_t* out) {
    const long long n = c.n;
    const int E = (int)k->limbs, nq = (int)k->level + 1;
    const long long P = (long long)E * n;
    double* stage = nullptr;
    uint64_t* coeff = nullptr;
    try {
        RCK_CHECK_HIP(hipMalloc(&stage, (size_t)2 * P * 2 * 8));
        RCK_CHECK_HIP(hipMalloc(&coeff, (size_t)2 * P * 2 * 8));
        RCK_CHECK_HIP(hipMemcpy(stage, k->sdata + (size_t)d * 2 * P * 2, 2 * P * 2 * 8, hipMemcpyDeviceToDevice));
        launch_inv(c, stage, n, 2 * n, kOutU64, coeff, n, 2 * n, 2, 2 * E, PrimeMap{0, (int)(c.L + c.K), 0}, nullptr);
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
// digits; adds add0/add1 and writes outs[b] = [c0][c1] (keyswitch_batch's contract).
void keyswitch_exact(Context& c, const uint64_t* poly, long long in_ps, int B, int limbs, const KeyDev* key, int levels,
                     const uint64_t* add0, const uint64_t* add1, long long add_ps, const std::vector<uint64_t*>& outs,
                     hipStream_t s, unsigned long long ginv, int bt, int jt) {
    require_fp64(c);
    const long long n = c.n;
    const int E = limbs + (int)key->aux;
    const ExactPlan pl = exact_plan(c, limbs, levels);
    const int G = pl.G, Gw = pl.Gw;
    int32_t* digits = static_cast<int32_t*>(c.workspace(5, (size_t)B * levels * n * 4));
    // workspace slots 30-34 (slots 20-26 hold the BSGS buffers of a transform whose rotations come here);
    // the accumulators are sized for the worst-case groups, which the recomputation reuses
    double* dn = static_cast<double*>(c.workspace(30, (size_t)B * levels * 2 * n * 8));
    double* acc = static_cast<double*>(c.workspace(31, (size_t)Gw * B * 2 * E * 2 * n * 8));
    double* accq = static_cast<double*>(c.workspace(32, (size_t)Gw * B * 2 * E * n * 8));
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
    // every digit under s1 and s2: limb stride 0 on the input (the same digit row for both)
    launch_fwd(c, kInI32, digits, 0, n, dn, n, 2 * n, 2, B * levels, ps, s);
    launch_mac_exact(c, bt, jt, dn, key->sdata, acc, B, levels, pl.gsize, G, E, s);
    // inverse rows under s1, s2; inverse columns fused with the recombination (one value per
    // accumulator and coefficient: the group's exact sum mod p_j); then the group sum and mod-down
    // (skipped when the gate is set)
    launch_inv_rows(c, acc, n, 2 * n, 2, G * B * 2 * E, ps, s);
    launch_inv_cols_crt(c, acc, accq, limbs, G * B * 2 * E, s);
    launch_crt_moddown(c, accq, limbs, B, G, add0, add1, add_ps, ginv, outs, s, gate, 0);
    if (pl.gated) {
        // some ciphertext's digits exceed the one-group bound: the batch again with the worst-case groups
        launch_mac_exact(c, bt, jt, dn, key->sdata, acc, B, levels, pl.gsw, Gw, E, s, gate);
        launch_inv_rows(c, acc, n, 2 * n, 2, Gw * B * 2 * E, ps, s, 1, 0, nullptr, 0, 0, 0, gate);
        launch_inv_cols_crt(c, acc, accq, limbs, Gw * B * 2 * E, s, gate);
        launch_crt_moddown(c, accq, limbs, B, Gw, add0, add1, add_ps, ginv, outs, s, gate, 1);
    }
}

}  // namespace rck
