// CKKS encode / decode on the host (src/rns/encoding.rs:9-87, src/math/canonical_embedding.rs:21-95,
// src/rns/polynomial.rs:57-85,162-235): the crate's radix-2 complex FFT with its twiddle
// recurrence, the xi^{-j} twist, scaling and rounding to i128, then residues per prime (encode);
// CRT reconstruction, centring against floor(Q/2), conversion to f64 and the forward embedding
// (decode).  Complex arithmetic follows num-complex 0.4 (products (ac - bd, ad + bc), from_polar
// = (r cos t, r sin t), powi by the inverse and square-and-multiply); no FMA contraction.
// Encoded polynomials are written to, and decoded ones read from, device memory.
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/rnsckks.h"
#include "algorithms.hpp"
#include "capi_types.hpp"

namespace rck {
namespace {

struct Cx {
    double re, im;
};
inline Cx cmul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline Cx cadd(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
inline Cx csub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
inline Cx polar(double r, double t) { return {r * std::cos(t), r * std::sin(t)}; }
inline Cx cinv(Cx a) {
    const double ns = a.re * a.re + a.im * a.im;
    return {a.re / ns, 0.0 - a.im / ns};
}

// base^e, e >= 0, by squaring (num-traits pow)
Cx cpow_u(Cx base, uint32_t e) {
    if (e == 0) return {1.0, 0.0};
    while ((e & 1) == 0) {
        base = cmul(base, base);
        e >>= 1;
    }
    if (e == 1) return base;
    Cx acc = base;
    while (e > 1) {
        e >>= 1;
        base = cmul(base, base);
        if (e & 1) acc = cmul(acc, base);
    }
    return acc;
}

Cx cpowi(Cx base, int32_t e) { return e < 0 ? cpow_u(cinv(base), (uint32_t)(-(int64_t)e)) : cpow_u(base, (uint32_t)e); }

void bit_reverse_permute(std::vector<Cx>& v) {
    const size_t n = v.size();
    size_t j = 0;
    for (size_t i = 1; i < n; i++) {
        size_t bit = n >> 1;
        while (j & bit) {
            j ^= bit;
            bit >>= 1;
        }
        j ^= bit;
        if (i < j) std::swap(v[i], v[j]);
    }
}

// fft_in_place (canonical_embedding.rs:44-78)
void fft(std::vector<Cx>& v, bool inverse) {
    const size_t n = v.size();
    bit_reverse_permute(v);
    for (size_t len = 2; len <= n; len *= 2) {
        const double angle = inverse ? -2.0 * M_PI / (double)len : 2.0 * M_PI / (double)len;
        const Cx wl = polar(1.0, angle);
        for (size_t c = 0; c < n; c += len) {
            Cx w{1.0, 0.0};
            for (size_t i = 0; i < len / 2; i++) {
                const Cx u = v[c + i];
                const Cx t = cmul(v[c + i + len / 2], w);
                v[c + i] = cadd(u, t);
                v[c + i + len / 2] = csub(u, t);
                w = cmul(w, wl);
            }
        }
    }
    if (inverse) {
        const double s = 1.0 / (double)n;
        for (auto& x : v) {
            x.re *= s;
            x.im *= s;
        }
    }
}

std::vector<size_t> slot_indices(size_t n) {
    std::vector<size_t> idx(n / 2);
    size_t cur = 1;
    for (size_t i = 0; i < n / 2; i++) {
        idx[i] = (cur - 1) / 2;
        cur = cur * 5 % (2 * n);
    }
    return idx;
}

__int128 round_to_i128(double x) {
    const double r = std::round(x);
    if (std::isnan(r)) return 0;
    const double lim = 170141183460469231731687303715884105728.0;  // 2^127
    if (r >= lim) return (__int128)(((unsigned __int128)1 << 127) - 1);
    if (r <= -lim) return -(__int128)(((unsigned __int128)1 << 127) - 1) - 1;
    return (__int128)r;
}

template <typename F>
void parallel_for(size_t n, F&& f) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t chunk = (n + hw - 1) / hw;
    std::vector<std::thread> ts;
    for (unsigned t = 0; t < hw; t++) {
        const size_t a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        ts.emplace_back([&f, a, b] {
            for (size_t i = a; i < b; i++) f(i);
        });
    }
    for (auto& t : ts) t.join();
}

// little-endian 32-bit word integers
using Big = std::vector<uint32_t>;

void big_mul_add(Big& x, uint64_t m, uint64_t a) {
    unsigned __int128 carry = a;
    for (auto& w : x) {
        const unsigned __int128 t = (unsigned __int128)w * m + carry;
        w = (uint32_t)t;
        carry = t >> 32;
    }
    while (carry) {
        x.push_back((uint32_t)carry);
        carry >>= 32;
    }
}

int big_cmp(const Big& a, const Big& b) {
    size_t na = a.size(), nb = b.size();
    while (na && a[na - 1] == 0) na--;
    while (nb && b[nb - 1] == 0) nb--;
    if (na != nb) return na < nb ? -1 : 1;
    for (size_t i = na; i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

Big big_sub(const Big& a, const Big& b) {  // a >= b
    Big r(a.size());
    int64_t borrow = 0;
    for (size_t i = 0; i < a.size(); i++) {
        int64_t t = (int64_t)a[i] - (i < b.size() ? (int64_t)b[i] : 0) - borrow;
        borrow = t < 0;
        r[i] = (uint32_t)(t + (borrow << 32));
    }
    return r;
}

// nearest double to a nonnegative big integer (round half to even)
double big_to_f64(const Big& x) {
    size_t n = x.size();
    while (n && x[n - 1] == 0) n--;
    if (n == 0) return 0.0;
    const int top = 31 - __builtin_clz(x[n - 1]);
    const long long bits = (long long)(n - 1) * 32 + top + 1;
    if (bits <= 64) {
        uint64_t v = 0;
        for (size_t i = n; i-- > 0;) v = (v << 32) | x[i];
        return (double)v;  // correctly rounded conversion
    }
    // top 64 bits plus a sticky bit for everything below them
    const long long shift = bits - 64;
    uint64_t v = 0;
    for (int k = 63; k >= 0; k--) {
        const long long b = shift + k;
        v |= (uint64_t)((x[b / 32] >> (b % 32)) & 1) << k;
    }
    bool sticky = false;
    for (long long b = 0; b < shift && !sticky; b++) sticky = (x[b / 32] >> (b % 32)) & 1;
    // keep 53 bits + guard; fold the rest into the sticky bit, then let the hardware round
    const uint64_t low = v & 0x3FF;  // the 10 bits below the 54 kept
    v = (v >> 10) << 10;
    if (low || sticky) v |= 1;  // odd marker below the guard bit decides ties
    return std::ldexp((double)v, (int)shift);
}

}  // namespace

void encode_host(const Context& c, const double* re, const double* im, uint32_t limbs, uint32_t scale_bits,
                 uint64_t* out, bool threads) {
    const size_t n = c.n;
    std::vector<Cx> ev(n, Cx{0.0, 0.0});
    size_t cur = 1;
    for (size_t i = 0; i < n / 2; i++) {
        const Cx v{re[i], im ? im[i] : 0.0};
        ev[(cur - 1) / 2] = v;
        ev[((2 * n - cur) % (2 * n) - 1) / 2] = Cx{v.re, -v.im};
        cur = cur * 5 % (2 * n);
    }
    fft(ev, true);
    const Cx xi = polar(1.0, M_PI / (double)n);
    const double scale = std::ldexp(1.0, (int)scale_bits);
    auto one = [&](size_t j) {
        const Cx tw = cmul(ev[j], cpowi(xi, -(int32_t)j));
        const __int128 v = round_to_i128(std::fma(tw.re, scale, 0.0));
        for (uint32_t l = 0; l < limbs; l++) {
            const __int128 q = (__int128)c.hp[l].q;
            __int128 r = v % q;
            if (r < 0) r += q;
            out[(size_t)l * n + j] = (uint64_t)r;
        }
    };
    if (threads) {
        parallel_for(n, one);
    } else {
        for (size_t j = 0; j < n; j++) one(j);
    }
}

}  // namespace rck

using namespace rck;

extern "C" {

int rck_encode(rck_context* ctx, const double* re, const double* im, uint32_t slots, uint32_t limbs,
               uint32_t scale_bits, uint64_t* out) {
    try {
        RCK_REQUIRE(ctx && ctx->c, "null context");
        Context& c = *ctx->c;
        RCK_REQUIRE(re != nullptr && out != nullptr, "null argument");
        RCK_REQUIRE(slots == c.n / 2, "slot count must be N/2");
        RCK_REQUIRE(limbs >= 1 && limbs <= c.L, "requested level context must exist for encoding");
        std::vector<uint64_t> host((size_t)limbs * c.n);
        encode_host(c, re, im, limbs, scale_bits, host.data(), true);
        RCK_CHECK_HIP(hipMemcpy(out, host.data(), host.size() * 8, hipMemcpyHostToDevice));
        return RCK_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return RCK_ERR_INTERNAL;
    }
}

int rck_decode(rck_context* ctx, const uint64_t* poly, uint32_t limbs, uint32_t scale_bits, double* re, double* im) {
    try {
        RCK_REQUIRE(ctx && ctx->c, "null context");
        Context& c = *ctx->c;
        RCK_REQUIRE(poly != nullptr && re != nullptr, "null argument");
        RCK_REQUIRE(limbs >= 1 && limbs <= c.L, "polynomial level context must exist for decoding");
        const size_t n = c.n;
        std::vector<uint64_t> host((size_t)limbs * n);
        RCK_CHECK_HIP(hipMemcpy(host.data(), poly, host.size() * 8, hipMemcpyDeviceToHost));
        // Garner constants: inv[i][j] = q_j^{-1} mod q_i (j < i)
        std::vector<uint64_t> inv((size_t)limbs * limbs, 0);
        for (uint32_t i = 0; i < limbs; i++)
            for (uint32_t j = 0; j < i; j++) {
                uint64_t v = 0;
                RCK_REQUIRE(h_invmod(c.hp[j].q % c.hp[i].q, c.hp[i].q, &v), "RNS moduli must be pairwise coprime");
                inv[(size_t)i * limbs + j] = v;
            }
        Big Q{1};
        for (uint32_t i = 0; i < limbs; i++) big_mul_add(Q, c.hp[i].q, 0);
        Big half(Q.size());
        for (size_t w = 0; w < Q.size(); w++) half[w] = (Q[w] >> 1) | (w + 1 < Q.size() ? Q[w + 1] << 31 : 0);
        const double inv_delta = 1.0 / std::ldexp(1.0, (int)scale_bits);
        const Cx xi = polar(1.0, M_PI / (double)n);
        std::vector<Cx> b(n);
        parallel_for(n, [&](size_t j) {
            uint64_t v[kMaxQ];
            for (uint32_t i = 0; i < limbs; i++) {
                const uint64_t qi = c.hp[i].q;
                uint64_t t = host[(size_t)i * n + j];
                for (uint32_t k = 0; k < i; k++) {
                    const uint64_t vk = v[k] % qi;
                    t = t >= vk ? t - vk : t + qi - vk;
                    t = h_mulmod(t, inv[(size_t)i * limbs + k], qi);
                }
                v[i] = t;
            }
            // mixed radix -> integer: x = v_{L-1}; x = x * q_i + v_i for i = L-2 .. 0
            Big x{0};
            big_mul_add(x, 1, v[limbs - 1]);
            for (uint32_t i = limbs - 1; i-- > 0;) big_mul_add(x, c.hp[i].q, v[i]);
            double val;
            if (big_cmp(x, half) <= 0) {
                val = big_to_f64(x);
            } else {
                val = -big_to_f64(big_sub(Q, x));
            }
            const Cx cj{val * inv_delta, 0.0};
            b[j] = cmul(cj, cpowi(xi, (int32_t)j));
        });
        fft(b, false);
        const std::vector<size_t> idx = slot_indices(n);
        for (size_t i = 0; i < n / 2; i++) {
            re[i] = b[idx[i]].re;
            if (im) im[i] = b[idx[i]].im;
        }
        return RCK_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return RCK_ERR_INTERNAL;
    }
}

}  // extern "C"
