// Seeded sampling, key generation, encryption and decryption.
//
// The crate samples from rand::thread_rng (unseeded; src/rns/keys.rs:83-151, ops.rs:17-20,
// keyswitching.rs:113-115).  Here every draw comes from one seeded stream, xoshiro256** seeded
// through splitmix64, consumed in the crate's order:
//   uniform poly   — for each limb, for each coefficient: next() % q          (keys.rs:88-100)
//   error poly     — N draws of round(sigma * z), z ~ N(0,1) by Box-Muller (the second normal of
//                    each pair is kept for the next draw), lifted to every limb (keys.rs:102-124)
//   ternary poly   — N draws of next() % 3 -> {0, 1, -1}                      (keys.rs:126-151)
//   sparse ternary — repeat: idx = next() % N; if free: sign = next() & 1 ? 1 : -1 (keys.rs:18-52)
// Sampling runs on the host (the stream is sequential); all polynomial arithmetic — the products
// a*s and pk*u, the key plaintexts, the NTT-form key upload — runs on the device.  Keys come out
// in the same NTT-resident form as rck_ksk_create.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/rnsckks.h"
#include "capi_types.hpp"
#include "wide.hpp"

struct rck_rng {
    uint64_t s[4];
    bool have_spare;
    double spare;
};

namespace rck {
namespace {

uint64_t splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t next(rck_rng* r) {
    uint64_t* s = r->s;
    const uint64_t out = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return out;
}

// uniform double in [0, 1) with 53 random bits
double unit(rck_rng* r) { return (double)(next(r) >> 11) * (1.0 / 9007199254740992.0); }

int64_t gaussian(rck_rng* r, double sigma) {
    double z;
    if (r->have_spare) {
        r->have_spare = false;
        z = r->spare;
    } else {
        double u1 = unit(r);
        const double u2 = unit(r);
        if (u1 < 1e-300) u1 = 1e-300;
        const double rad = std::sqrt(-2.0 * std::log(u1));
        z = rad * std::cos(2.0 * M_PI * u2);
        r->spare = rad * std::sin(2.0 * M_PI * u2);
        r->have_spare = true;
    }
    return (int64_t)std::llround(z * sigma);
}

uint64_t prime_at(const Context& c, int l, int nq) { return c.hp[l < nq ? l : (int)c.L + (l - nq)].q; }

std::vector<uint64_t> host_uniform(const Context& c, rck_rng* r, int nq, int np) {
    std::vector<uint64_t> v((size_t)(nq + np) * c.n);
    for (int l = 0; l < nq + np; l++) {
        const uint64_t q = prime_at(c, l, nq);
        uint64_t* row = v.data() + (size_t)l * c.n;
        for (uint32_t i = 0; i < c.n; i++) row[i] = next(r) % q;
    }
    return v;
}

std::vector<int64_t> host_error(const Context& c, rck_rng* r, double sigma) {
    std::vector<int64_t> v(c.n);
    for (auto& x : v) x = gaussian(r, sigma);
    return v;
}

std::vector<int64_t> host_ternary(const Context& c, rck_rng* r) {
    std::vector<int64_t> v(c.n);
    for (auto& x : v) {
        const uint64_t k = next(r) % 3;
        x = k == 0 ? 0 : (k == 1 ? 1 : -1);
    }
    return v;
}

std::vector<int64_t> host_sparse_ternary(const Context& c, rck_rng* r, uint32_t hamming) {
    std::vector<int64_t> v(c.n, 0);
    uint32_t placed = 0;
    while (placed < hamming) {
        const size_t idx = (size_t)(next(r) % c.n);
        if (v[idx] == 0) {
            v[idx] = (next(r) & 1) ? 1 : -1;
            placed++;
        }
    }
    return v;
}

// small signed coefficients -> residues of every limb (|v| < q)
__global__ void k_lift_small(const int64_t* __restrict__ v, uint64_t* __restrict__ out, int n, PrimeMap pm,
                             const DevPrime* __restrict__ primes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    if (i >= n) return;
    const uint64_t q = primes[pm(l)].qu;
    const int64_t x = v[i];
    uint64_t r = x >= 0 ? (uint64_t)x % q : q - ((uint64_t)(-x) % q);
    out[(size_t)l * n + i] = r == q ? 0 : r;
}

// Residues of every limb of a small polynomial given as residues mod the first Q prime (centred).
__global__ void k_small_from_limb0(const uint64_t* __restrict__ s0, uint64_t* __restrict__ out, int n, PrimeMap pm,
                                   const DevPrime* __restrict__ primes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    if (i >= n) return;
    const uint64_t q0 = primes[0].qu;
    const uint64_t q = primes[pm(l)].qu;
    const uint64_t x = s0[i];
    uint64_t r;
    if (x <= q0 / 2) {
        r = x % q;
    } else {
        const uint64_t m = (q0 - x) % q;
        r = m == 0 ? 0 : q - m;
    }
    out[(size_t)l * n + i] = r;
}

// out = a * f[l] (+ b) (- c) mod q, canonical (b, c optional)
__global__ void k_scale_add_sub(const uint64_t* __restrict__ a, const uint64_t* __restrict__ f,
                                const uint64_t* __restrict__ b, const uint64_t* __restrict__ cc,
                                uint64_t* __restrict__ out, int n, PrimeMap pm, const DevPrime* __restrict__ primes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    if (i >= n) return;
    const DevPrime P = primes[pm(l)];
    const size_t k = (size_t)l * n + i;
    const uint64_t q = (uint64_t)P.q;
    double t = canon(mulmod(load_centred(a[k], P.q), load_centred(f[l], P.q), P.q, P.qinv), P.q, P.qinv);
    uint64_t r = (uint64_t)t;
    if (b) {
        r += b[k];
        if (r >= q) r -= q;
    }
    if (cc) r = r >= cc[k] ? r - cc[k] : r + q - cc[k];
    out[k] = r;
}

// k_scale_add_sub for 64-bit integer contexts
__global__ void k_w_scale_add_sub(const uint64_t* __restrict__ a, const uint64_t* __restrict__ f,
                                  const uint64_t* __restrict__ b, const uint64_t* __restrict__ cc,
                                  uint64_t* __restrict__ out, int n, PrimeMap pm, const WidePrime* __restrict__ wp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    if (i >= n) return;
    const WidePrime P = wp[pm(l)];
    const size_t k = (size_t)l * n + i;
    uint64_t r = w_mul(w_mul(a[k], f[l], P.q, P.qn), P.r2, P.q, P.qn);
    if (b) r = w_add(r, b[k], P.q);
    if (cc) r = w_sub(r, cc[k], P.q);
    out[k] = r;
}

struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) { RCK_CHECK_HIP(hipMalloc(&p, bytes ? bytes : 8)); }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

void upload(void* dst, const void* src, size_t bytes) {
    RCK_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

dim3 grid_of(const Context& c, int limbs) { return dim3((c.n + 255) / 256, limbs); }

void lift_small(const Context& c, const std::vector<int64_t>& v, uint64_t* out, int nq, int np) {
    DevBuf d(v.size() * 8);
    upload(d.p, v.data(), v.size() * 8);
    hipLaunchKernelGGL(k_lift_small, grid_of(c, nq + np), dim3(256), 0, nullptr, d.as<int64_t>(), out, (int)c.n,
                       PrimeMap{nq, (int)c.L, 0}, c.d_primes);
    RCK_CHECK_HIP(hipGetLastError());
    RCK_CHECK_HIP(hipDeviceSynchronize());
}

// the secret's residues on every limb of the basis Q[0..nq) || P[0..np)
void secret_to_basis(const Context& c, const uint64_t* secret, uint64_t* out, int nq, int np) {
    hipLaunchKernelGGL(k_small_from_limb0, grid_of(c, nq + np), dim3(256), 0, nullptr, secret, out, (int)c.n,
                       PrimeMap{nq, (int)c.L, 0}, c.d_primes);
    RCK_CHECK_HIP(hipGetLastError());
}

void scale_add_sub(const Context& c, const uint64_t* a, const std::vector<uint64_t>& f, const uint64_t* b,
                   const uint64_t* cc, uint64_t* out, int nq, int np) {
    DevBuf df(f.size() * 8);
    upload(df.p, f.data(), f.size() * 8);
    if (c.wide)
        hipLaunchKernelGGL(k_w_scale_add_sub, grid_of(c, nq + np), dim3(256), 0, nullptr, a, df.as<uint64_t>(), b, cc,
                           out, (int)c.n, PrimeMap{nq, (int)c.L, 0}, c.d_wp);
    else
        hipLaunchKernelGGL(k_scale_add_sub, grid_of(c, nq + np), dim3(256), 0, nullptr, a, df.as<uint64_t>(), b, cc,
                           out, (int)c.n, PrimeMap{nq, (int)c.L, 0}, c.d_primes);
    RCK_CHECK_HIP(hipGetLastError());
    RCK_CHECK_HIP(hipDeviceSynchronize());
}

// (c0, c1) = (pk0 * u + e0 + m, pk1 * u + e1) over `limbs` limbs (ops.rs:8-46); pk has pk_limbs
// limbs per component (truncated to `limbs`).  Draws u, e0, e1 in that order.
void encrypt_into(Context& c, rck_rng* r, double sigma, const uint64_t* m, const uint64_t* pk, int pk_limbs,
                  int limbs, uint64_t* out) {
    const long long n = c.n, P = (long long)limbs * n;
    const PrimeMap pm{limbs, (int)c.L, 0};
    DevBuf u(P * 8), e0(P * 8), e1(P * 8);
    lift_small(c, host_ternary(c, r), u.as<uint64_t>(), limbs, 0);
    lift_small(c, host_error(c, r, sigma), e0.as<uint64_t>(), limbs, 0);
    lift_small(c, host_error(c, r, sigma), e1.as<uint64_t>(), limbs, 0);
    poly_mul(c, pk, (long long)pk_limbs * n, u.as<uint64_t>(), 0, out, P, limbs, 2, pm, nullptr);
    launch_addsub(c, {out, n, 0}, {e0.as<uint64_t>(), n, 0}, {out, n, 0}, limbs, 1, 0, pm, nullptr);
    if (m) launch_addsub(c, {out, n, 0}, {m, n, 0}, {out, n, 0}, limbs, 1, 0, pm, nullptr);
    launch_addsub(c, {out + P, n, 0}, {e1.as<uint64_t>(), n, 0}, {out + P, n, 0}, limbs, 1, 0, pm, nullptr);
    RCK_CHECK_HIP(hipDeviceSynchronize());
}

uint32_t decomp_digits(const Context& c, int limbs) { return (c.qbits[limbs] + 29) / 30; }

// generate_keyswitch_key_rns_hybrid (keyswitching.rs:81-128) / generate_keyswitch_key_rns (:44-79)
// for a small source polynomial given on the eval basis (hybrid) or on Q (Q-only).
rck_ksk* keygen_switch(rck_context* ctx, rck_rng* r, double sigma, const uint64_t* source, const uint64_t* secret,
                       const uint64_t* pk, int pk_limbs, int level, uint64_t galois) {
    Context& c = *ctx->c;
    const int nq = level + 1, K = (int)c.K, E = nq + K;
    const long long n = c.n, P = (long long)E * n;
    const uint32_t D = decomp_digits(c, nq);
    const PrimeMap pm{nq, (int)c.L, 0};
    rck_ksk* k = new rck_ksk();
    k->owner = ctx;
    k->digits = D;
    k->level = (uint32_t)level;
    k->aux = (uint32_t)K;
    k->limbs = (uint32_t)E;
    k->galois = galois;
    // 64-bit integer contexts collect the key in coefficient form and convert it at the end
    uint64_t* coef = nullptr;
    auto store_digit = [&](uint32_t i, const uint64_t* a0) {
        if (coef)
            RCK_CHECK_HIP(hipMemcpy(coef + (size_t)i * 2 * P, a0, (size_t)2 * P * 8, hipMemcpyDeviceToDevice));
        else
            launch_fwd(c, kInU64, a0, n, P, k->data + (size_t)i * 2 * P, n, P, E, 2, pm, nullptr);
    };
    try {
        if (c.wide)
            RCK_CHECK_HIP(hipMalloc(&coef, (size_t)D * 2 * P * 8));
        else
            RCK_CHECK_HIP(hipMalloc(&k->data, (size_t)D * 2 * P * 8));
        DevBuf stage(2 * P * 8);
        uint64_t* a0 = stage.as<uint64_t>();
        uint64_t* a1 = a0 + P;
        if (K > 0) {
            DevBuf tsec(P * 8), err(P * 8), a1s(P * 8);
            secret_to_basis(c, secret, tsec.as<uint64_t>(), nq, K);
            // P mod p_j for every eval prime
            std::vector<uint64_t> pmod(E);
            for (int l = 0; l < E; l++) {
                const uint64_t pj = prime_at(c, l, nq);
                uint64_t m = 1 % pj;
                for (int a = 0; a < K; a++) m = h_mulmod(m, c.hp[c.L + a].q % pj, pj);
                pmod[l] = m;
            }
            for (uint32_t i = 0; i < D; i++) {
                std::vector<uint64_t> f(E);
                for (int l = 0; l < E; l++) {
                    const uint64_t pj = prime_at(c, l, nq);
                    f[l] = h_mulmod(h_powmod(2 % pj, 30ull * i, pj), pmod[l], pj);
                }
                const std::vector<uint64_t> ua = host_uniform(c, r, nq, K);
                upload(a1, ua.data(), ua.size() * 8);
                lift_small(c, host_error(c, r, sigma), err.as<uint64_t>(), nq, K);
                poly_mul(c, a1, P, tsec.as<uint64_t>(), 0, a1s.as<uint64_t>(), P, E, 1, pm, nullptr);
                scale_add_sub(c, source, f, err.as<uint64_t>(), a1s.as<uint64_t>(), a0, nq, K);
                store_digit(i, a0);
                RCK_CHECK_HIP(hipDeviceSynchronize());
            }
        } else {
            DevBuf pt(P * 8);
            for (uint32_t i = 0; i < D; i++) {
                std::vector<uint64_t> f(E);
                for (int l = 0; l < E; l++) f[l] = h_powmod(2 % c.hp[l].q, 30ull * i, c.hp[l].q);
                scale_add_sub(c, source, f, nullptr, nullptr, pt.as<uint64_t>(), nq, 0);
                encrypt_into(c, r, sigma, pt.as<uint64_t>(), pk, pk_limbs, nq, a0);
                store_digit(i, a0);
                RCK_CHECK_HIP(hipDeviceSynchronize());
            }
        }
        if (coef) {
            wide_key_create(c, k, coef);
            RCK_CHECK_HIP(hipFree(coef));
            coef = nullptr;
        } else if (c.ks_method != RCK_KS_DIGIT_NTT && exact_keyswitch_available(c, (uint32_t)K) &&
                   (c.ks_method != RCK_KS_AUTO || k->digits <= 64)) {
            key_make_exact(c, k, c.ks_method == RCK_KS_AUTO);
        }
    } catch (...) {
        if (coef) (void)hipFree(coef);
        if (k->data) (void)hipFree(k->data);
        if (k->mdata) (void)hipFree(k->mdata);
        if (k->wdata) (void)hipFree(k->wdata);
        delete k;
        throw;
    }
    return k;
}

template <typename F>
int kguard(F&& f) {
    try {
        f();
        return RCK_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return RCK_ERR_INTERNAL;
    } catch (const std::exception& e) {
        set_error(e.what());
        return RCK_ERR_INTERNAL;
    }
}

Context& ctx_of(rck_context* ctx) {
    RCK_REQUIRE(ctx && ctx->c, "null context");
    return *ctx->c;
}

void check_basis(const Context& c, uint32_t nq, uint32_t np) {
    RCK_REQUIRE(nq >= 1 && nq <= c.L, "limb count must be in 1..=" + std::to_string(c.L));
    RCK_REQUIRE(np == 0 || np == c.K, "aux limb count must be 0 or the aux basis size");
}

void check_sigma(double sigma) { RCK_REQUIRE(sigma > 0.0 && sigma < 1e6, "sigma must be positive"); }

void check_rng(rck_rng* r) { RCK_REQUIRE(r != nullptr, "null rng"); }

}  // namespace
}  // namespace rck

using namespace rck;

extern "C" {

int rck_rng_create(uint64_t seed, rck_rng** out) {
    return kguard([&] {
        RCK_REQUIRE(out != nullptr, "null output");
        rck_rng* r = new rck_rng();
        uint64_t x = seed;
        for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
        r->have_spare = false;
        r->spare = 0;
        *out = r;
    });
}

int rck_rng_destroy(rck_rng* rng) {
    delete rng;
    return RCK_OK;
}

uint64_t rck_rng_next(rck_rng* rng) { return rng ? next(rng) : 0; }

int rck_sample_uniform(rck_context* ctx, rck_rng* rng, uint64_t* out, uint32_t nq, uint32_t np) {
    return kguard([&] {
        Context& c = ctx_of(ctx);
        check_basis(c, nq, np);
        check_rng(rng);
        check_ptr(out, "out");
        const std::vector<uint64_t> v = host_uniform(c, rng, (int)nq, (int)np);
        upload(out, v.data(), v.size() * 8);
    });
}

int rck_sample_error(rck_context* ctx, rck_rng* rng, double sigma, uint64_t* out, uint32_t nq, uint32_t np) {
    return kguard([&] {
        Context& c = ctx_of(ctx);
        check_basis(c, nq, np);
        check_rng(rng);
        check_sigma(sigma);
        check_ptr(out, "out");
        lift_small(c, host_error(c, rng, sigma), out, (int)nq, (int)np);
    });
}

int rck_sample_ternary(rck_context* ctx, rck_rng* rng, uint64_t* out, uint32_t nq, uint32_t np) {
    return kguard([&] {
        Context& c = ctx_of(ctx);
        check_basis(c, nq, np);
        check_rng(rng);
        check_ptr(out, "out");
        lift_small(c, host_ternary(c, rng), out, (int)nq, (int)np);
    });
}

int rck_sample_sparse_ternary(rck_context* ctx, rck_rng* rng, uint32_t hamming_weight, uint64_t* out, uint32_t nq,
                              uint32_t np) {
    return kguard([&] {
        Context& c = ctx_of(ctx);
        check_basis(c, nq, np);
        check_rng(rng);
        check_ptr(out, "out");
        RCK_REQUIRE(hamming_weight <= c.n, "hamming_weight must be <= poly_degree");
        lift_small(c, host_sparse_ternary(c, rng, hamming_weight), out, (int)nq, (int)np);
    });
}

int rck_keygen_public(rck_context* ctx, rck_rng* rng, double sigma, const uint64_t* secret, uint32_t limbs,
                      uint64_t* pk) {
    return kguard([&] {
        Context& c = ctx_of(ctx);
        check_basis(c, limbs, 0);
        check_rng(rng);
        check_sigma(sigma);
        check_ptr(secret, "secret");
        check_ptr(pk, "public key");
        const long long n = c.n, P = (long long)limbs * n;
        const PrimeMap pm{(int)limbs, (int)c.L, 0};
        const std::vector<uint64_t> a = host_uniform(c, rng, (int)limbs, 0);
        upload(pk + P, a.data(), a.size() * 8);
        DevBuf err(P * 8), as(P * 8);
        lift_small(c, host_error(c, rng, sigma), err.as<uint64_t>(), (int)limbs, 0);
        poly_mul(c, pk + P, P, secret, P, as.as<uint64_t>(), P, (int)limbs, 1, pm, nullptr);
        launch_addsub(c, {err.as<uint64_t>(), n, 0}, {as.as<uint64_t>(), n, 0}, {pk, n, 0}, (int)limbs, 1, 1, pm,
                      nullptr);
        RCK_CHECK_HIP(hipDeviceSynchronize());
    });
}

int rck_keygen_relin(rck_context* ctx, rck_rng* rng, double sigma, const uint64_t* secret, const uint64_t* pk,
                     uint32_t pk_limbs, uint32_t level, rck_ksk** out) {
    return kguard([&] {
        Context& c = ctx_of(ctx);
        check_rng(rng);
        check_sigma(sigma);
        check_ptr(secret, "secret");
        check_ptr(out, "output");
        *out = nullptr;
        RCK_REQUIRE(level < c.L, "relinearization level context must exist");
        const int nq = (int)level + 1, E = nq + (int)c.K;
        const long long n = c.n;
        if (c.K == 0) {
            check_ptr(pk, "public key");
            RCK_REQUIRE(pk_limbs >= (uint32_t)nq && pk_limbs <= c.L, "public key levels must cover the key level");
        }
        // source = s^2: exact on every limb because its integer coefficients are below q/2
        DevBuf sb((size_t)E * n * 8), src((size_t)E * n * 8);
        secret_to_basis(c, secret, sb.as<uint64_t>(), nq, (int)c.K);
        poly_mul(c, sb.as<uint64_t>(), 0, sb.as<uint64_t>(), 0, src.as<uint64_t>(), 0, E, 1,
                 PrimeMap{nq, (int)c.L, 0}, nullptr);
        RCK_CHECK_HIP(hipDeviceSynchronize());
        *out = keygen_switch(ctx, rng, sigma, src.as<uint64_t>(), secret, pk, (int)pk_limbs, (int)level, 0);
    });
}

int rck_keygen_rotation(rck_context* ctx, rck_rng* rng, double sigma, const uint64_t* secret, const uint64_t* pk,
                        uint32_t pk_limbs, int64_t steps, uint32_t level, rck_ksk** out) {
    return kguard([&] {
        Context& c = ctx_of(ctx);
        check_rng(rng);
        check_sigma(sigma);
        check_ptr(secret, "secret");
        check_ptr(out, "output");
        *out = nullptr;
        RCK_REQUIRE(level < c.L, "keyswitch level context must exist");
        const int nq = (int)level + 1, E = nq + (int)c.K;
        const long long n = c.n;
        if (c.K == 0) {
            check_ptr(pk, "public key");
            RCK_REQUIRE(pk_limbs >= (uint32_t)nq && pk_limbs <= c.L, "public key levels must cover the key level");
        }
        const uint64_t g = rck_galois_element_for_rotation(c.n, steps);
        // source = s(X^g) (the rotated secret, context/rotation.rs:30-31), on the key's basis
        DevBuf sb((size_t)E * n * 8), src((size_t)E * n * 8);
        secret_to_basis(c, secret, sb.as<uint64_t>(), nq, (int)c.K);
        launch_galois(c, {sb.as<uint64_t>(), n, 0}, {src.as<uint64_t>(), n, 0}, E, 1, g, PrimeMap{nq, (int)c.L, 0},
                      nullptr);
        RCK_CHECK_HIP(hipDeviceSynchronize());
        *out = keygen_switch(ctx, rng, sigma, src.as<uint64_t>(), secret, pk, (int)pk_limbs, (int)level, g);
    });
}

int rck_ksk_export(rck_context* ctx, const rck_ksk* key, uint64_t* out) {
    return kguard([&] {
        Context& c = ctx_of(ctx);
        RCK_REQUIRE(key != nullptr && (key->data != nullptr || key->mdata != nullptr || key->wdata != nullptr),
                    "null key");
        check_ptr(out, "out");
        if (key->wdata) return wide_key_export(c, key, out);
        require_fp64(c);
        const long long n = c.n, P = (long long)key->limbs * n;
        const PrimeMap pm{(int)key->level + 1, (int)c.L, 0};
        DevBuf t(2 * P * 8), u(2 * P * 8);
        if (!key->data) {
            for (uint32_t d = 0; d < key->digits; d++) {
                key_export_exact_digit(c, key, d, u.as<uint64_t>());
                RCK_CHECK_HIP(hipMemcpy(out + (size_t)d * 2 * P, u.p, 2 * P * 8, hipMemcpyDeviceToHost));
            }
            return;
        }
        for (uint32_t d = 0; d < key->digits; d++) {
            RCK_CHECK_HIP(hipMemcpy(t.p, key->data + (size_t)d * 2 * P, 2 * P * 8, hipMemcpyDeviceToDevice));
            launch_inv(c, t.as<double>(), n, P, kOutU64, u.p, n, P, (int)key->limbs, 2, pm, nullptr);
            RCK_CHECK_HIP(hipMemcpy(out + (size_t)d * 2 * P, u.p, 2 * P * 8, hipMemcpyDeviceToHost));
        }
    });
}

int rck_ksk_info(const rck_ksk* key, uint32_t* digits, uint32_t* level, uint32_t* aux, uint64_t* galois) {
    return kguard([&] {
        RCK_REQUIRE(key != nullptr, "null key");
        if (digits) *digits = key->digits;
        if (level) *level = key->level;
        if (aux) *aux = key->aux;
        if (galois) *galois = key->galois;
    });
}

int rck_encrypt(rck_context* ctx, rck_rng* rng, double sigma, const uint64_t* plain, const uint64_t* pk,
                uint32_t pk_limbs, uint32_t limbs, uint64_t* out) {
    return kguard([&] {
        Context& c = ctx_of(ctx);
        check_basis(c, limbs, 0);
        check_rng(rng);
        check_sigma(sigma);
        check_ptr(plain, "plaintext");
        check_ptr(pk, "public key");
        check_ptr(out, "out");
        RCK_REQUIRE(pk_limbs >= limbs && pk_limbs <= c.L, "public key levels must match the active RNS context");
        encrypt_into(c, rng, sigma, plain, pk, (int)pk_limbs, (int)limbs, out);
    });
}

int rck_decrypt(rck_context* ctx, const uint64_t* ct, uint32_t limbs, const uint64_t* secret, uint64_t* out,
                void* stream) {
    return kguard([&] {
        Context& c = ctx_of(ctx);
        check_basis(c, limbs, 0);
        check_ptr(ct, "ciphertext");
        check_ptr(secret, "secret");
        check_ptr(out, "out");
        const long long n = c.n, P = (long long)limbs * n;
        const PrimeMap pm{(int)limbs, (int)c.L, 0};
        hipStream_t s = static_cast<hipStream_t>(stream);
        poly_mul(c, ct + P, P, secret, P, out, P, (int)limbs, 1, pm, s);
        launch_addsub(c, {ct, n, 0}, {out, n, 0}, {out, n, 0}, (int)limbs, 1, 0, pm, s);
    });
}

}  // extern "C"
