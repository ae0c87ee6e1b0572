/*
 * Drives every function of the CPU oracle (ckks_oracle.c) at small sizes so that a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer sees each of them run — TEST INFRASTRUCTURE ONLY
 * (built and run by tests/test_oracle_sanitized.py; nothing else uses it).
 *
 * Besides running under the sanitizers it checks the exact invariants that need no second
 * implementation: the NTT round trip (ntt.rs:195-204), multiply_ntt == multiply_naive
 * (tests/math/poly_tests.rs:23-34), an automorphism followed by its inverse, the tensor
 * decryption identity Dec2(a (x) b) = Dec(a) Dec(b) (ops.rs:84-113), recomposition of the balanced
 * digits limb by limb (keyswitching.rs:196-236), mod-down of P*y lifted to the eval basis == y
 * (keyswitching.rs:278-322), and the constructors' rejections (ntt.rs:17-60, basis.rs:32-59).
 * Prints "sanitize_check ok" and exits 0 when everything holds.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ckks_oracle.h"

static int failures = 0;
#define CHECK(cond, what)                                                   \
    do {                                                                    \
        if (!(cond)) {                                                      \
            fprintf(stderr, "FAILED: %s (line %d)\n", what, __LINE__);      \
            failures++;                                                     \
        }                                                                   \
    } while (0)

static uint64_t* alloc_u64(size_t count) {
    uint64_t* p = (uint64_t*)calloc(count ? count : 1, sizeof(uint64_t));
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

static void uniform(const orc_basis* b, orc_rng* r, uint64_t* out, size_t polys) {
    for (size_t k = 0; k < polys; k++) orc_sample_uniform(b, r, out + k * b->levels * b->n);
}

static void plans_and_polys(void) {
    char err[256];
    orc_ntt_plan plan;
    CHECK(orc_ntt_plan_new(48, 998244353, 3, &plan, err, sizeof err) != 0, "size not a power of two rejected");
    CHECK(orc_ntt_plan_new(64, 1000003, 2, &plan, err, sizeof err) != 0, "2N not dividing q - 1 rejected");
    CHECK(orc_ntt_plan_new(64, 998244353, 3, &plan, err, sizeof err) == 0, "plan N=64");
    const size_t n = 64;
    const uint64_t q = 998244353;
    orc_rng r;
    orc_rng_seed(&r, 7);
    uint64_t *a = alloc_u64(n), *b = alloc_u64(n), *c = alloc_u64(n), *d = alloc_u64(n), *t = alloc_u64(n);
    for (size_t i = 0; i < n; i++) {
        a[i] = orc_rng_next(&r) % q;
        b[i] = orc_rng_next(&r) % q;
    }
    memcpy(t, a, n * sizeof *t);
    orc_ntt_forward(&plan, t);
    orc_ntt_inverse(&plan, t);
    CHECK(memcmp(t, a, n * sizeof *t) == 0, "NTT round trip");
    orc_poly_multiply_ntt(a, b, c, &plan);
    orc_poly_multiply_naive(a, b, d, n, q);
    CHECK(memcmp(c, d, n * sizeof *c) == 0, "multiply_ntt == multiply_naive");
    orc_poly_add(a, b, c, n, q);
    orc_poly_sub(c, b, d, n, q);
    CHECK(memcmp(d, a, n * sizeof *d) == 0, "(a + b) - b == a");
    const uint64_t g = 5, ginv = orc_mod_pow(5, n / 2 - 1, 2 * n);  /* 5 has order N/2 modulo 2N */
    CHECK((g * ginv) % (2 * n) == 1, "inverse Galois element");
    orc_poly_galois(a, c, n, q, g);
    orc_poly_galois(c, d, n, q, ginv);
    CHECK(memcmp(d, a, n * sizeof *d) == 0, "automorphism then its inverse");
    orc_poly_reduce(a, c, n, q);
    CHECK(memcmp(c, a, n * sizeof *c) == 0, "reduce of reduced values");
    int ok = 0;
    const uint64_t inv = orc_mod_inv(12345, q, &ok);
    CHECK(ok && orc_mod_mul(inv, 12345, q) == 1, "mod_inv");
    orc_mod_inv(0, q, &ok);
    CHECK(!ok, "mod_inv(0) reports no inverse");
    CHECK(orc_mod_sub(orc_mod_add(q - 1, 5, q), 5, q) == q - 1, "add/sub wrap");
    free(a), free(b), free(c), free(d), free(t);
}

/* A ring with nq Q primes and na aux primes: scheme ops, decomposition, lift/mod-down, keygen,
 * keyswitching (hybrid when na > 0, else Q-only), relinearization and rotation. */
static void ring(size_t n, const uint64_t* primes, const uint64_t* roots, size_t nq, size_t na) {
    char err[256];
    orc_basis qb, eb;
    CHECK(orc_basis_new(n, primes, roots, nq, &qb, err, sizeof err) == 0, "Q basis");
    if (na) CHECK(orc_basis_new(n, primes, roots, nq + na, &eb, err, sizeof err) == 0, "eval basis");
    const size_t P = nq * n, EP = (nq + na) * n;
    orc_rng r;
    orc_rng_seed(&r, 11 + n + nq + na);
    uint64_t *s = alloc_u64(P), *x = alloc_u64(P), *y = alloc_u64(P), *z = alloc_u64(P);
    uint64_t *ct_a = alloc_u64(2 * P), *ct_b = alloc_u64(2 * P), *quad = alloc_u64(3 * P);
    uint64_t *pk = alloc_u64(2 * P), *pke = alloc_u64(P), *u = alloc_u64(P), *e0 = alloc_u64(P), *e1 = alloc_u64(P);
    orc_sample_ternary(&qb, &r, s);
    orc_sample_sparse_ternary(&qb, &r, n / 4, z);
    orc_sample_error(&qb, &r, 3.19, e0);
    orc_keygen_public(&qb, &r, 3.19, s, pk, pke);
    uniform(&qb, &r, x, 1);
    orc_sample_ternary(&qb, &r, u);
    orc_sample_error(&qb, &r, 3.19, e1);
    orc_encrypt_with_samples(&qb, x, pk, u, e0, e1, ct_a);
    orc_decrypt(&qb, ct_a, s, y);
    uniform(&qb, &r, ct_a, 2);
    uniform(&qb, &r, ct_b, 2);
    /* tensor identity */
    orc_multiply_ciphertexts(&qb, ct_a, ct_b, quad);
    orc_decrypt_quadratic(&qb, quad, s, x);
    orc_decrypt(&qb, ct_a, s, y);
    orc_decrypt(&qb, ct_b, s, z);
    uint64_t* w = alloc_u64(P);
    orc_rns_multiply_ntt(&qb, y, z, w);
    CHECK(memcmp(w, x, P * sizeof *w) == 0, "Dec2(a x b) == Dec(a) Dec(b)");
    orc_multiply_plain(&qb, ct_a, y, ct_b);
    orc_rns_add(&qb, y, z, w);
    orc_rns_sub(&qb, w, z, x);
    CHECK(memcmp(x, y, P * sizeof *x) == 0, "RNS (y + z) - z == y");
    orc_rns_galois(&qb, y, w, 2 * n - 1);
    if (nq > 1) {
        uint64_t* res = alloc_u64(3 * P);
        orc_rescale(&qb, quad, 3, res);
        free(res);
    }
    /* decomposition: sum_i d_i 2^(30 i) == x in every limb */
    const size_t bits = orc_total_modulus_bits(primes, nq), D = orc_num_decomp_digits(bits, 30);
    uint64_t* dec = alloc_u64(D * P);
    int64_t* sd = (int64_t*)calloc(D * n, sizeof(int64_t));
    orc_decompose(&qb, y, 30, D, dec);
    orc_decompose_signed(&qb, y, 30, D, sd);
    int recomposed = 1, signed_match = 1;
    for (size_t l = 0; l < nq; l++) {
        const uint64_t q = primes[l], base = orc_mod_pow(2, 30, q);
        for (size_t c = 0; c < n; c++) {
            uint64_t acc = 0, pw = 1;
            for (size_t i = 0; i < D; i++) {
                const uint64_t di = dec[(i * nq + l) * n + c];
                acc = orc_mod_add(acc, orc_mod_mul(di, pw, q), q);
                pw = orc_mod_mul(pw, base, q);
                const int64_t v = sd[i * n + c];
                const uint64_t vm = v >= 0 ? (uint64_t)v % q : (q - (uint64_t)(-v) % q) % q;
                if (vm != di) signed_match = 0;
            }
            if (acc != y[l * n + c]) recomposed = 0;
        }
    }
    CHECK(recomposed, "balanced digits recompose to x");
    CHECK(signed_match, "signed digits == residues of the digits");
    free(dec);
    free(sd);
    if (na) {
        /* lift, and mod-down of P*y (zero on the aux limbs) == y */
        uint64_t* lifted = alloc_u64(EP);
        orc_lift(&qb, &eb, y, lifted);
        CHECK(memcmp(lifted, y, P * sizeof *y) == 0, "lift keeps the Q limbs");
        for (size_t l = 0; l < nq; l++) {
            uint64_t pm = 1;
            for (size_t k = 0; k < na; k++) pm = orc_mod_mul(pm, primes[nq + k] % primes[l], primes[l]);
            for (size_t c = 0; c < n; c++) lifted[l * n + c] = orc_mod_mul(y[l * n + c], pm, primes[l]);
        }
        memset(lifted + P, 0, (EP - P) * sizeof *lifted);
        orc_mod_down(&qb, &eb, lifted, w);
        CHECK(memcmp(w, y, P * sizeof *w) == 0, "mod_down(P y) == y");
        free(lifted);
    }
    /* keys and keyswitching */
    const size_t KL = na ? EP : P;
    uint64_t* ksk = alloc_u64(D * 2 * KL);
    uint64_t* ss = alloc_u64(P);
    uint64_t* out = alloc_u64(2 * P);
    orc_rns_multiply_ntt(&qb, s, s, ss);
    for (int noiseless = 0; noiseless < 2; noiseless++) {
        if (na)
            orc_keygen_hybrid(&qb, &eb, &r, 3.19, ss, s, D, noiseless, ksk);
        else
            orc_keygen_qonly(&qb, &r, 3.19, ss, pk, D, noiseless, ksk);
        CHECK(orc_relinearize(&qb, na ? &eb : NULL, quad, ksk, D, out) == 0, "relinearize");
        if (na)
            CHECK(orc_keyswitch_hybrid(&qb, &eb, quad + 2 * P, ksk, D, out) == 0, "hybrid keyswitch");
        else
            CHECK(orc_keyswitch_qonly(&qb, quad + 2 * P, ksk, D, out) == 0, "Q-only keyswitch");
    }
    if (!na) CHECK(orc_keyswitch_qonly(&qb, quad + 2 * P, ksk, D - 1, out) != 0, "too few digits rejected");
    const uint64_t gel = orc_galois_element_for_rotation(n, -1);
    orc_rns_galois(&qb, s, w, gel);
    if (na)
        orc_keygen_hybrid(&qb, &eb, &r, 3.19, w, s, D, 0, ksk);
    else
        orc_keygen_qonly(&qb, &r, 3.19, w, pk, D, 0, ksk);
    CHECK(orc_rotate(&qb, na ? &eb : NULL, ct_a, gel, ksk, D, out) == 0, "rotate");
    free(ksk), free(ss), free(out), free(w);
    free(s), free(x), free(y), free(z), free(ct_a), free(ct_b), free(quad);
    free(pk), free(pke), free(u), free(e0), free(e1);
    orc_basis_free(&qb);
    if (na) orc_basis_free(&eb);
}

int main(void) {
    orc_set_threads(2);
    plans_and_polys();
    {
        char err[256];
        orc_basis bad;
        const uint64_t dup[2] = {998244353, 998244353}, roots[2] = {3, 3};
        CHECK(orc_basis_new(64, dup, roots, 2, &bad, err, sizeof err) != 0, "repeated prime rejected");
    }
    /* the crate's sample_context primes (rns/tests.rs:14-24) plus two aux primes = 1 mod 128 */
    const uint64_t small[5] = {998244353, 1004535809, 469762049, 167772161, 754974721};
    const uint64_t small_g[5] = {3, 3, 3, 3, 11};
    ring(64, small, small_g, 3, 0);
    ring(64, small, small_g, 3, 2);
    ring(64, small, small_g, 1, 2);
    /* 50-bit primes = 1 mod 2^13 (the N=2^12 chain): the u128 paths at full width */
    const uint64_t big[3] = {1125899906826241ull, 1125899906629633ull, 1125899906424833ull};
    const uint64_t big_g[3] = {22, 10, 5};
    ring(4096, big, big_g, 3, 0);
    if (failures) {
        fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    printf("sanitize_check ok\n");
    return 0;
}
