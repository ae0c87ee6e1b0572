// Mod-down helper shared by the keyswitch kernels: the aux part of a coefficient in mixed radix.
#pragma once
#include "ops.hpp"

namespace rck {

// Garner digits m_0 .. m_{K-1} of x_P from its aux residues r[0..K) and the centring flag
// x_P > floor(P/2) (mixed_radix_digits / mod_q_biguint_to_centered, keyswitching.rs:397-410, 227-235)
__device__ __forceinline__ bool p_garner(const double* r, int K, int L, int T, double* m,
                                         const DevPrime* __restrict__ primes, const double* __restrict__ invq,
                                         const uint32_t* __restrict__ pwords) {
    for (int k = 0; k < K; k++) {
        const DevPrime Pk = primes[L + k];
        double v = r[k];
        for (int j = 0; j < k; j++) {
            double d = reduce(v - m[j], Pk.q, Pk.qinv);
            v = mulmod(d, invq[(long long)(L + k) * T + (L + j)], Pk.q, Pk.qinv);
        }
        m[k] = k == 0 ? v : canon(v, Pk.q, Pk.qinv);
    }
    bool gt;
    if (K == 2) {  // X_P = m0 + m1 p0 < 2^102 as (hi, lo) 64-bit words
        const uint64_t p0 = (uint64_t)primes[L].q, m1 = (uint64_t)m[1];
        uint64_t lo = m1 * p0, hi = __umul64hi(m1, p0);
        const uint64_t l2 = lo + (uint64_t)m[0];
        hi += l2 < lo;
        lo = l2;
        const uint64_t hlo = (uint64_t)pwords[8] | ((uint64_t)pwords[9] << 32);
        const uint64_t hhi = (uint64_t)pwords[10] | ((uint64_t)pwords[11] << 32);
        gt = hi > hhi || (hi == hhi && lo > hlo);
    } else {
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
        gt = false;
        bool decided = false;
#pragma unroll
        for (int w = 7; w >= 0; w--) {
            const uint32_t h = pwords[8 + w];
            if (!decided && xw[w] != h) {
                gt = xw[w] > h;
                decided = true;
            }
        }
    }
    return gt;
}

}  // namespace rck
