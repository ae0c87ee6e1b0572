// Exact modular arithmetic for primes q < 2^50 + 2^46 on the FP64 pipes of CDNA4.
//
// MI355X issues FP64 FMA at full rate (64 lanes/clk/CU) while 32-bit integer multiplies run
// at quarter rate, so residues live in registers as integer-valued doubles and every modular
// product is computed exactly with FMA:
//
//   h = fl(a*b), l = a*b - h (exact, FMA two-product), k = round(h/q) via the 1.5*2^52 shift,
//   r = (h - k*q) + l  (both steps exact because the results are integers below 2^53).
//
// Representation: a residue x mod q is any integer-valued double; "reduced" means |x| <= q/2+1.
// Exactness conditions (q < 2^50 + 2^46 < 2^50.1, enforced by the context; operands integer-valued):
//   mulmod(a, w): |a| <= 3.5q and |w| <= q/2 + 1 ->  |result| <= 0.97 q
//                 (|a*w/q| <= 1.75q < 2^51 keeps the shift rounding exact; the computed quotient
//                  is off by < 1.75q * 2^-52 < 0.47, so |a*w - k*q| <= 0.97q; h - k*q and its sum
//                  with l are integers below 2^53 and therefore exact).
//   reduce(x):    |x| < 2^53                     ->  |result| <= q/2 + 1.
// Outputs crossing the C ABI are canonical u64 in [0, q).  Nothing here is approximate: each
// function returns the exact representative described above.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rck {

// 1.5 * 2^52: x + kShift rounds any |x| < 2^51 to the nearest integer (ties to even).
constexpr double kShift = 6755399441055744.0;

__device__ __forceinline__ double round_int(double x) { return (x + kShift) - kShift; }

// a*w mod q as a signed representative with |r| <= 0.97q (see header for input bounds).
__device__ __forceinline__ double mulmod(double a, double w, double q, double qinv) {
    double h = a * w;
    double l = __builtin_fma(a, w, -h);
    double k = __builtin_fma(h, qinv, kShift) - kShift;
    double r = __builtin_fma(-k, q, h);
    return r + l;
}

// x mod q with |result| <= q/2 + 1, for |x| < 2^51.
__device__ __forceinline__ double reduce(double x, double q, double qinv) {
    double k = __builtin_fma(x, qinv, kShift) - kShift;
    return __builtin_fma(-k, q, x);
}

// canonical representative in [0, q) of any |x| < 2^51.
__device__ __forceinline__ double canon(double x, double q, double qinv) {
    double r = reduce(x, q, qinv);
    return r < 0.0 ? r + q : r;
}

__device__ __forceinline__ double u64_to_f64(uint64_t x) { return (double)x; }  // exact: x < 2^53

__device__ __forceinline__ uint64_t f64_to_u64(double x) { return (uint64_t)x; }  // exact: integer in [0, 2^53)

// canonical u64 in [0,q) -> centred double in (-q/2, q/2]
__device__ __forceinline__ double load_centred(uint64_t x, double q) {
    double d = (double)x;
    return d > 0.5 * q ? d - q : d;
}

}  // namespace rck
