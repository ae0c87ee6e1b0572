// C ABI (include/rnsckks.h): argument checks with the crate's assertion messages, scale/level
// metadata rules, and the per-op composition of the kernels.
#include <cstring>
#include <exception>
#include <new>

#include "../../include/rnsckks.h"
#include "capi_types.hpp"
#include "wide.hpp"

namespace rck {
Context* context_create(uint32_t n, const uint64_t* q, const uint64_t* g, uint32_t L, const uint64_t* p,
                        const uint64_t* gp, uint32_t K, uint32_t scale_bits);
void noalloc_setup_current_device();
}

using namespace rck;


template <typename F>
static int guard(F&& f) {
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

static hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

static Context& C(rck_context* ctx) {
    RCK_REQUIRE(ctx && ctx->c, "null context");
    return *ctx->c;
}

static void check_limbs(const Context& c, uint32_t nq, uint32_t np) {
    RCK_REQUIRE(nq >= 1 && nq <= c.L, "limb count must be in 1..=" + std::to_string(c.L) + ", got " +
                                          std::to_string(nq));
    RCK_REQUIRE(np == 0 || np == c.K, "aux limb count must be 0 or the aux basis size");
}


// floor(log2 q) of the dropped prime (ops.rs:74)
static uint32_t floor_log2(uint64_t q) { return 63 - __builtin_clzll(q); }

extern "C" {

const char* rck_last_error(void) { return last_error(); }
int rck_version(void) { return 100; }

int rck_malloc(void** ptr, size_t bytes) {
    return guard([&] {
        check_ptr(ptr, "pointer");
        RCK_CHECK_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    });
}
int rck_free(void* ptr) {
    return guard([&] { RCK_CHECK_HIP(hipFree(ptr)); });
}
int rck_memcpy_htod(void* dst, const void* src, size_t bytes) {
    return guard([&] { RCK_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int rck_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    return guard([&] { RCK_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int rck_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
    return guard([&] { RCK_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream))); });
}
int rck_device_synchronize(void) {
    return guard([&] { RCK_CHECK_HIP(hipDeviceSynchronize()); });
}
int rck_stream_create(void** stream) {
    return guard([&] {
        check_ptr(stream, "stream");
        hipStream_t s;
        RCK_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        *stream = s;
    });
}
int rck_stream_destroy(void* stream) {
    return guard([&] { RCK_CHECK_HIP(hipStreamDestroy(S(stream))); });
}
int rck_stream_synchronize(void* stream) {
    return guard([&] { RCK_CHECK_HIP(hipStreamSynchronize(S(stream))); });
}
int rck_event_create(void** event) {
    return guard([&] {
        hipEvent_t e;
        RCK_CHECK_HIP(hipEventCreate(&e));
        *event = e;
    });
}
int rck_event_destroy(void* event) {
    return guard([&] { RCK_CHECK_HIP(hipEventDestroy(static_cast<hipEvent_t>(event))); });
}
int rck_event_record(void* event, void* stream) {
    return guard([&] { RCK_CHECK_HIP(hipEventRecord(static_cast<hipEvent_t>(event), S(stream))); });
}
int rck_event_synchronize(void* event) {
    return guard([&] { RCK_CHECK_HIP(hipEventSynchronize(static_cast<hipEvent_t>(event))); });
}
int rck_event_elapsed_ms(void* start, void* stop, float* ms) {
    return guard([&] {
        RCK_CHECK_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    });
}

int rck_context_create(uint32_t poly_degree, const uint64_t* primes, const uint64_t* roots, uint32_t n_primes,
                       const uint64_t* aux_primes, const uint64_t* aux_roots, uint32_t n_aux, uint32_t scale_bits,
                       rck_context** out) {
    return guard([&] {
        check_ptr(out, "output");
        *out = nullptr;
        RCK_REQUIRE(n_primes == 0 || (primes && roots), "null prime list");
        RCK_REQUIRE(n_aux == 0 || (aux_primes && aux_roots), "null aux prime list");
        Context* c = context_create(poly_degree, primes, roots, n_primes, aux_primes, aux_roots, n_aux, scale_bits);
        try {
            noalloc_setup_current_device();  // this device's allocation-free scratch address
        } catch (...) {
            delete c;
            throw;
        }
        *out = new rck_context{c};
    });
}

int rck_context_destroy(rck_context* ctx) {
    return guard([&] {
        if (!ctx) return;
        delete ctx->c;
        delete ctx;
    });
}

int rck_context_set_keyswitch_method(rck_context* ctx, int method) {
    return guard([&] {
        Context& c = C(ctx);
        RCK_REQUIRE(method == RCK_KS_EXACT_CONVOLUTION || method == RCK_KS_DIGIT_NTT || method == RCK_KS_AUTO,
                    "unknown keyswitch method");
        c.ks_method = method;
    });
}

int rck_context_set_arithmetic(rck_context* ctx, int arithmetic) {
    return guard([&] {
        Context& c = C(ctx);
        RCK_REQUIRE(arithmetic == RCK_ARITH_FP64 || arithmetic == RCK_ARITH_INT64, "unknown arithmetic");
        if (arithmetic == RCK_ARITH_INT64) return context_make_wide(c);
        for (uint32_t i = 0; i < c.L + c.K; i++)
            RCK_REQUIRE(c.hp[i].q < (1ull << 50) + (1ull << 46),
                        "moduli at or above 2^50 + 2^46 need the 64-bit integer arithmetic");
        c.wide = false;
    });
}

int rck_context_arithmetic(const rck_context* ctx, int* arithmetic) {
    return guard([&] {
        const Context& c = C(const_cast<rck_context*>(ctx));
        check_ptr(arithmetic, "arithmetic");
        *arithmetic = c.wide ? RCK_ARITH_INT64 : RCK_ARITH_FP64;
    });
}

int rck_keyswitch_exact_groups(rck_context* ctx, uint32_t limbs, uint32_t* digits_per_group, uint32_t* groups) {
    return guard([&] {
        Context& c = C(ctx);
        RCK_REQUIRE(limbs >= 1 && limbs <= c.L, "bad limb count");
        RCK_REQUIRE(exact_keyswitch_available(c, c.K), "the context has no exact-convolution keyswitch");
        const int D = (int)(c.qbits[limbs] + 29) / 30;
        int g = 0, G = 0;
        exact_grouping(c, (int)limbs, D, &g, &G);
        if (digits_per_group) *digits_per_group = (uint32_t)g;
        if (groups) *groups = (uint32_t)G;
    });
}

int rck_ksk_method(const rck_ksk* key, int* method) {
    return guard([&] {
        RCK_REQUIRE(key != nullptr && method != nullptr, "null argument");
        // 64-bit integer keys: digit NTT
        *method = key->mdata ? (key->data ? RCK_KS_AUTO : RCK_KS_EXACT_CONVOLUTION) : RCK_KS_DIGIT_NTT;
    });
}

int rck_num_decomp_digits(const rck_context* ctx, uint32_t limbs, uint32_t* digits) {
    return guard([&] {
        const Context& c = C(const_cast<rck_context*>(ctx));
        RCK_REQUIRE(limbs >= 1 && limbs <= c.L, "bad limb count");
        *digits = (c.qbits[limbs] + 29) / 30;
    });
}

uint64_t rck_galois_element_for_rotation(uint32_t poly_degree, int64_t steps) {
    const uint64_t m = 2ull * poly_degree;
    uint64_t g = 5 % m, e = steps >= 0 ? (uint64_t)steps : (uint64_t)(-steps);
    if (steps < 0) {
        uint64_t inv;
        if (!h_invmod(g, m, &inv)) return 0;
        g = inv;
    }
    uint64_t r = 1 % m;
    while (e) {
        if (e & 1) r = (uint64_t)((unsigned __int128)r * g % m);
        g = (uint64_t)((unsigned __int128)g * g % m);
        e >>= 1;
    }
    return r;
}

// ------------------------------------------------------------------ L0 cyclic NTT
int rck_ntt_forward(rck_context* ctx, uint64_t* data, uint32_t nq, uint32_t np, void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, nq, np);
        check_ptr(data, "data");
        const int limbs = nq + np;
        const long long n = c.n;
        PrimeMap pm{(int)nq, (int)c.L, 0};
        if (c.wide) return wide_ntt_cyclic(c, false, data, limbs, pm, S(stream));
        double* t = static_cast<double*>(c.workspace(8, (size_t)limbs * n * 8));
        launch_ntt_cyclic(c, false, data, t, limbs, pm, S(stream));
    });
}

int rck_ntt_inverse(rck_context* ctx, uint64_t* data, uint32_t nq, uint32_t np, void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, nq, np);
        check_ptr(data, "data");
        const int limbs = nq + np;
        const long long n = c.n;
        PrimeMap pm{(int)nq, (int)c.L, 0};
        if (c.wide) return wide_ntt_cyclic(c, true, data, limbs, pm, S(stream));
        double* t = static_cast<double*>(c.workspace(8, (size_t)limbs * n * 8));
        launch_ntt_cyclic(c, true, data, t, limbs, pm, S(stream));
    });
}

// ------------------------------------------------------------------ L1/L2 poly ops
int rck_poly_add(rck_context* ctx, const uint64_t* a, const uint64_t* b, uint64_t* out, uint32_t nq, uint32_t np,
                 void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, nq, np);
        const long long n = c.n;
        launch_addsub(c, {a, n, 0}, {b, n, 0}, {out, n, 0}, nq + np, 1, 0, PrimeMap{(int)nq, (int)c.L, 0}, S(stream));
    });
}

int rck_poly_sub(rck_context* ctx, const uint64_t* a, const uint64_t* b, uint64_t* out, uint32_t nq, uint32_t np,
                 void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, nq, np);
        const long long n = c.n;
        launch_addsub(c, {a, n, 0}, {b, n, 0}, {out, n, 0}, nq + np, 1, 1, PrimeMap{(int)nq, (int)c.L, 0}, S(stream));
    });
}

int rck_poly_multiply_ntt(rck_context* ctx, const uint64_t* a, const uint64_t* b, uint64_t* out, uint32_t nq,
                          uint32_t np, void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, nq, np);
        check_ptr(a, "a");
        check_ptr(b, "b");
        check_ptr(out, "out");
        const long long P = (long long)(nq + np) * c.n;
        poly_mul(c, a, P, b, P, out, P, nq + np, 1, PrimeMap{(int)nq, (int)c.L, 0}, S(stream));
    });
}

int rck_poly_galois(rck_context* ctx, const uint64_t* a, uint64_t* out, uint32_t nq, uint32_t np, uint64_t galois,
                    void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, nq, np);
        RCK_REQUIRE(a != out, "galois automorphism cannot run in place");
        const long long n = c.n;
        launch_galois(c, {a, n, 0}, {out, n, 0}, nq + np, 1, galois, PrimeMap{(int)nq, (int)c.L, 0}, S(stream));
    });
}

int rck_poly_mul_add(rck_context* ctx, const uint64_t* a, const uint64_t* b, const uint64_t* cc, uint64_t* out,
                     uint32_t nq, uint32_t np, void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, nq, np);
        const long long n = c.n;
        launch_mul_add_u64(c, {a, n, 0}, {b, n, 0}, {cc, n, 0}, cc != nullptr, {out, n, 0}, nq + np, 1,
                           PrimeMap{(int)nq, (int)c.L, 0}, S(stream));
    });
}

// ------------------------------------------------------------------ L3 scheme ops
int rck_ct_add(rck_context* ctx, const uint64_t* lhs, const uint64_t* rhs, uint32_t limbs, uint32_t lhs_scale,
               uint32_t rhs_scale, uint64_t* out, uint32_t* out_scale, void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, limbs, 0);
        RCK_REQUIRE(lhs_scale == rhs_scale, "ciphertexts must share the same scale to add");
        const long long n = c.n, P = (long long)limbs * n;
        launch_addsub(c, {lhs, n, P}, {rhs, n, P}, {out, n, P}, limbs, 2, 0, PrimeMap{(int)limbs, (int)c.L, 0},
                      S(stream));
        if (out_scale) *out_scale = lhs_scale;
    });
}

int rck_mult(rck_context* ctx, const uint64_t* lhs, const uint64_t* rhs, uint32_t limbs, uint32_t lhs_scale,
             uint32_t rhs_scale, uint64_t* out_quad, uint32_t* out_scale, void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, limbs, 0);
        check_ptr(lhs, "lhs");
        check_ptr(rhs, "rhs");
        check_ptr(out_quad, "out");
        RCK_REQUIRE(lhs_scale == rhs_scale, "ciphertexts must share the same scale to multiply");
        mult_impl(c, lhs, rhs, limbs, out_quad, S(stream));
        if (out_scale) *out_scale = lhs_scale + rhs_scale;
    });
}

int rck_mult_plain(rck_context* ctx, const uint64_t* ct, const uint64_t* plain, uint32_t limbs, uint32_t ct_scale,
                   uint64_t* out, uint32_t* out_scale, void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, limbs, 0);
        const long long P = (long long)limbs * c.n;
        poly_mul(c, ct, P, plain, 0, out, P, limbs, 2, PrimeMap{(int)limbs, (int)c.L, 0}, S(stream));
        if (out_scale) *out_scale = ct_scale + ct_scale;
    });
}

int rck_rescale(rck_context* ctx, const uint64_t* ct, uint32_t limbs, uint32_t scale, uint64_t* out,
                uint32_t* out_scale, void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, limbs, 0);
        RCK_REQUIRE(limbs > 1, "rescale requires at least two levels");
        RCK_REQUIRE((const void*)ct != (const void*)out, "rescale cannot run in place");
        rescale_impl(c, ct, limbs, 2, out, S(stream));
        if (out_scale) {
            uint32_t b = floor_log2(c.hp[limbs - 1].q);
            *out_scale = scale > b ? scale - b : 0;
        }
    });
}

int rck_rescale_quadratic(rck_context* ctx, const uint64_t* quad, uint32_t limbs, uint32_t scale, uint64_t* out,
                          uint32_t* out_scale, void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, limbs, 0);
        RCK_REQUIRE(limbs > 1, "quadratic rescale requires at least two levels");
        RCK_REQUIRE((const void*)quad != (const void*)out, "rescale cannot run in place");
        rescale_impl(c, quad, limbs, 3, out, S(stream));
        if (out_scale) {
            uint32_t b = floor_log2(c.hp[limbs - 1].q);
            *out_scale = scale > b ? scale - b : 0;
        }
    });
}

// ------------------------------------------------------------------ keys
static void ksk_create(rck_context* ctx, const uint64_t* key, bool on_device, uint32_t digits, uint32_t level,
                       uint32_t aux, uint64_t galois, rck_ksk** out) {
    Context& c = C(ctx);
    check_ptr(out, "output");
    check_ptr(key, "key");
    *out = nullptr;
    RCK_REQUIRE(level < c.L, "keyswitch key level out of range");
    RCK_REQUIRE(aux == 0 || aux == c.K, "aux_level_count must be 0 or the context's aux basis size");
    RCK_REQUIRE(digits >= 1, "keyswitch key must have digits");
    const uint32_t limbs = level + 1 + aux;
    const long long n = c.n, P = (long long)limbs * n;
    rck_ksk* k = new rck_ksk();
    k->owner = ctx;
    k->data = nullptr;
    k->digits = digits;
    k->level = level;
    k->aux = aux;
    k->limbs = limbs;
    k->galois = galois;
    try {
        if (c.wide) {
            uint64_t* dev = nullptr;
            if (!on_device) {
                RCK_CHECK_HIP(hipMalloc(&dev, (size_t)digits * 2 * P * 8));
                RCK_CHECK_HIP(hipMemcpy(dev, key, (size_t)digits * 2 * P * 8, hipMemcpyHostToDevice));
            }
            try {
                wide_key_create(c, k, on_device ? key : dev);
            } catch (...) {
                if (dev) (void)hipFree(dev);
                throw;
            }
            if (dev) RCK_CHECK_HIP(hipFree(dev));
            *out = k;
            return;
        }
        RCK_CHECK_HIP(hipMalloc(&k->data, (size_t)digits * 2 * P * 8));
        uint64_t* stage = nullptr;
        if (!on_device) RCK_CHECK_HIP(hipMalloc(&stage, (size_t)2 * P * 8));
        PrimeMap pm{(int)(level + 1), (int)c.L, 0};
        for (uint32_t d = 0; d < digits; d++) {
            const uint64_t* src = key + (size_t)d * 2 * P;
            if (!on_device) {
                RCK_CHECK_HIP(hipMemcpy(stage, src, (size_t)2 * P * 8, hipMemcpyHostToDevice));
                src = stage;
            }
            launch_fwd(c, kInU64, src, n, P, k->data + (size_t)d * 2 * P, n, P, limbs, 2, pm, nullptr);
        }
        RCK_CHECK_HIP(hipDeviceSynchronize());
        if (stage) RCK_CHECK_HIP(hipFree(stage));
        // automatic keys take the exact-convolution form where it applies (at most 64 digits)
        if (c.ks_method != RCK_KS_DIGIT_NTT && exact_keyswitch_available(c, aux) &&
            (c.ks_method != RCK_KS_AUTO || digits <= 64))
            key_make_exact(c, k, c.ks_method == RCK_KS_AUTO);
    } catch (...) {
        if (k->data) (void)hipFree(k->data);
        if (k->mdata) (void)hipFree(k->mdata);
        if (k->wdata) (void)hipFree(k->wdata);
        delete k;
        throw;
    }
    *out = k;
}

int rck_ksk_create(rck_context* ctx, const uint64_t* key, uint32_t digits, uint32_t level, uint32_t aux_level_count,
                   uint64_t galois_element, rck_ksk** out) {
    return guard([&] { ksk_create(ctx, key, false, digits, level, aux_level_count, galois_element, out); });
}

int rck_ksk_create_device(rck_context* ctx, const uint64_t* key, uint32_t digits, uint32_t level,
                          uint32_t aux_level_count, uint64_t galois_element, rck_ksk** out) {
    return guard([&] { ksk_create(ctx, key, true, digits, level, aux_level_count, galois_element, out); });
}

int rck_ksk_destroy(rck_ksk* key) {
    return guard([&] {
        if (!key) return;
        RCK_CHECK_HIP(hipDeviceSynchronize());
        if (key->data) RCK_CHECK_HIP(hipFree(key->data));
        if (key->mdata) RCK_CHECK_HIP(hipFree(key->mdata));
        if (key->wdata) RCK_CHECK_HIP(hipFree(key->wdata));
        delete key;
    });
}

int rck_keyswitch(rck_context* ctx, const uint64_t* poly, uint32_t limbs, const rck_ksk* key, uint64_t* out,
                  uint32_t* out_scale, void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, limbs, 0);
        check_ptr(poly, "poly");
        check_ptr(out, "out");
        keyswitch_impl(c, poly, limbs, key, nullptr, nullptr, out, S(stream));
        if (out_scale) *out_scale = limbs;
    });
}

int rck_relinearize(rck_context* ctx, const uint64_t* quad, uint32_t limbs, uint32_t scale, const rck_ksk* key,
                    uint64_t* out, uint32_t* out_scale, void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, limbs, 0);
        check_ptr(quad, "quad");
        check_ptr(out, "out");
        const long long P = (long long)limbs * c.n;
        keyswitch_impl(c, quad + 2 * P, limbs, key, quad, quad + P, out, S(stream));
        if (out_scale) *out_scale = scale;
    });
}

int rck_mult_relinearized(rck_context* ctx, const uint64_t* lhs, const uint64_t* rhs, uint32_t limbs,
                          uint32_t lhs_scale, uint32_t rhs_scale, const rck_ksk* key, uint64_t* out,
                          uint32_t* out_scale, void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, limbs, 0);
        RCK_REQUIRE(lhs_scale == rhs_scale, "ciphertexts must share the same scale to multiply");
        RCK_REQUIRE(limbs > 1, "rescale requires at least two levels");
        const long long P = (long long)limbs * c.n;
        uint64_t* rel = static_cast<uint64_t*>(c.workspace(3, (size_t)2 * P * 8));
        mult_relin_batch(c, lhs, rhs, 0, 1, limbs, key, {rel}, S(stream));
        rescale_impl(c, rel, limbs, 2, out, S(stream));
        if (out_scale) {
            uint32_t s = lhs_scale + rhs_scale, b = floor_log2(c.hp[limbs - 1].q);
            *out_scale = s > b ? s - b : 0;
        }
    });
}

int rck_ct_galois(rck_context* ctx, const uint64_t* ct, uint32_t limbs, uint64_t galois, uint64_t* out,
                  void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, limbs, 0);
        RCK_REQUIRE(ct != out, "galois automorphism cannot run in place");
        const long long n = c.n, P = (long long)limbs * n;
        launch_galois(c, {ct, n, P}, {out, n, P}, limbs, 2, galois, PrimeMap{(int)limbs, (int)c.L, 0}, S(stream));
    });
}

int rck_rotate(rck_context* ctx, const uint64_t* ct, uint32_t limbs, uint32_t scale, const rck_ksk* key,
               uint64_t* out, uint32_t* out_scale, void* stream) {
    return guard([&] {
        Context& c = C(ctx);
        check_limbs(c, limbs, 0);
        check_ptr(key, "key");
        check_ptr(ct, "ct");
        check_ptr(out, "out");
        RCK_REQUIRE((int)key->level == (int)limbs - 1, "rotation key level must match the ciphertext level");
        rotate_batch(c, {ct}, (int)limbs, key, {out}, S(stream));
        if (out_scale) *out_scale = scale;
    });
}

int rck_diag_copy(void* dst, const void* src, size_t bytes, void* stream) {
    return guard([&] { launch_copy(src, dst, bytes, S(stream)); });
}

int rck_profile_enable(int on) {
    return guard([&] { prof_set(on != 0); });
}

int rck_profile_report(char* buf, size_t len) {
    return guard([&] {
        std::string r = prof_report();
        RCK_REQUIRE(buf && len > r.size(), "report buffer too small");
        std::memcpy(buf, r.c_str(), r.size() + 1);
    });
}

}  // extern "C"
