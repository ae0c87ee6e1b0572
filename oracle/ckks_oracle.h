/*
 * CPU oracle for the RNS-CKKS arithmetic path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference crate's CPU algorithm (Akram275/rns-ckks-rust),
 * function by function, with the same u128 "%" modular arithmetic, the same radix-2 NTT
 * with bit reversal and twiddle recurrence, three NTTs per multiply_ntt, coefficient-form
 * keys, and the base-2^30 balanced gadget decomposition with mixed-radix lifting and
 * mod-down.  Every function cites the reference file:line it follows.
 *
 * Only the test suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * this library, and only as the checker (or as the timed CPU baseline).  The product path
 * never calls it.
 *
 * Data layout: an RNS polynomial with `levels` limbs of degree N is uint64_t[levels][N]
 * (limb-major, row i reduced modulo prime i).  A ciphertext is [c0][c1] back to back.
 *
 * Parity pinning (see tests/test_oracle.py): every function below is checked against the
 * reference's own known-answer tests, closed forms (brute-force DFT, schoolbook negacyclic
 * product, big-integer CRT formulas computed independently in Python), or exact invariants
 * (noiseless keyswitch, public-key relation, quadratic decryption identity).
 * The random samplers are seeded restatements of unseeded reference samplers: their
 * streams cannot match the reference (which uses thread_rng) and are "parity unpinned";
 * only their distributions are restated.
 */
#ifndef CKKS_ORACLE_H
#define CKKS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scalar modular arithmetic: src/math/ntt.rs:132-188 ---- */
uint64_t orc_mod_add(uint64_t a, uint64_t b, uint64_t q);
uint64_t orc_mod_sub(uint64_t a, uint64_t b, uint64_t q);
uint64_t orc_mod_mul(uint64_t a, uint64_t b, uint64_t q);
uint64_t orc_mod_pow(uint64_t base, uint64_t exp, uint64_t q);
/* returns 0 and sets *ok = 0 when not invertible */
uint64_t orc_mod_inv(uint64_t v, uint64_t q, int* ok);

/* ---- NttPlan: src/math/ntt.rs:4-114 ---- */
typedef struct {
    uint64_t modulus;
    uint64_t size;
    uint64_t root;
    uint64_t root_inv;
    uint64_t psi;      /* negacyclic_twiddle */
    uint64_t psi_inv;  /* negacyclic_twiddle_inv */
    uint64_t size_inv;
} orc_ntt_plan;

/* 0 on success; on failure writes the reference's error text into err (ntt.rs:17-60). */
int orc_ntt_plan_new(uint64_t size, uint64_t modulus, uint64_t primitive_root, orc_ntt_plan* out, char* err,
                     size_t err_len);
void orc_ntt_forward(const orc_ntt_plan* plan, uint64_t* values);
void orc_ntt_inverse(const orc_ntt_plan* plan, uint64_t* values);

/* ---- ModularPolynomial: src/poly/mod.rs ---- */
void orc_poly_reduce(const uint64_t* in, uint64_t* out, size_t n, uint64_t q);                 /* :14-18 */
void orc_poly_add(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n, uint64_t q);  /* :38-47 */
void orc_poly_sub(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n, uint64_t q);  /* :50-59 */
void orc_poly_multiply_naive(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n, uint64_t q); /* :62-107 */
void orc_poly_galois(const uint64_t* a, uint64_t* out, size_t n, uint64_t q, uint64_t galois);   /* :110-130 */
void orc_poly_multiply_ntt(const uint64_t* a, const uint64_t* b, uint64_t* out, const orc_ntt_plan* plan); /* :134-161 */

/* ---- RNS basis: a list of primes with their plans (src/rns/basis.rs:24-112) ---- */
typedef struct {
    size_t n;         /* poly degree */
    size_t levels;    /* number of primes */
    uint64_t* moduli; /* [levels] */
    uint64_t* roots;  /* [levels] */
    orc_ntt_plan* plans;
} orc_basis;

/* 0 ok; errors follow RnsContext::new (basis.rs:32-59). */
int orc_basis_new(size_t n, const uint64_t* moduli, const uint64_t* roots, size_t levels, orc_basis* out, char* err,
                  size_t err_len);
void orc_basis_free(orc_basis* b);
/* bits of the product of the first `levels` moduli (basis.rs:77-85, num-bigint bits()). */
uint64_t orc_total_modulus_bits(const uint64_t* moduli, size_t levels);
/* num_decomp_digits (keyswitching.rs:238-240) */
size_t orc_num_decomp_digits(uint64_t total_modulus_bits, size_t decomp_bits);

/* ---- RnsPolynomial limb ops: src/rns/polynomial.rs:99-145 (rayon over limbs -> OpenMP) ---- */
void orc_rns_add(const orc_basis* b, const uint64_t* x, const uint64_t* y, uint64_t* out);
void orc_rns_sub(const orc_basis* b, const uint64_t* x, const uint64_t* y, uint64_t* out);
void orc_rns_multiply_ntt(const orc_basis* b, const uint64_t* x, const uint64_t* y, uint64_t* out);
void orc_rns_galois(const orc_basis* b, const uint64_t* x, uint64_t* out, uint64_t galois);

/* ---- scheme ops: src/rns/ops.rs ---- */
/* multiply_ciphertexts_rns (:84-100): lhs/rhs = [c0][c1], out = [c0][c1][c2] */
void orc_multiply_ciphertexts(const orc_basis* b, const uint64_t* lhs, const uint64_t* rhs, uint64_t* out);
/* multiply_plain_rns (:54-66): ct = [c0][c1], plain = 1 poly, out = [c0][c1] */
void orc_multiply_plain(const orc_basis* b, const uint64_t* ct, const uint64_t* plain, uint64_t* out);
/* rescale_poly_rns (:135-173) applied to `polys` polynomials: in has b->levels limbs, out has b->levels-1 */
void orc_rescale(const orc_basis* b, const uint64_t* in, size_t polys, uint64_t* out);
/* decrypt_rns (:44-52) and decrypt_quadratic_rns (:102-113) */
void orc_decrypt(const orc_basis* b, const uint64_t* ct, const uint64_t* secret, uint64_t* out);
void orc_decrypt_quadratic(const orc_basis* b, const uint64_t* ct, const uint64_t* secret, uint64_t* out);
/* encrypt_rns_with_samples (:25-42): pk = [p0][p1] */
void orc_encrypt_with_samples(const orc_basis* b, const uint64_t* plain, const uint64_t* pk, const uint64_t* u,
                              const uint64_t* e0, const uint64_t* e1, uint64_t* out);

/* ---- keyswitching: src/rns/keyswitching.rs ---- */
/* decompose_poly_rns_with_levels (:196-236): out = [levels][b->levels][N] residues */
void orc_decompose(const orc_basis* b, const uint64_t* poly, size_t decomp_bits, size_t levels, uint64_t* out);
/* same decomposition, but the signed balanced digits themselves: out = [levels][N] int64 */
void orc_decompose_signed(const orc_basis* b, const uint64_t* poly, size_t decomp_bits, size_t levels, int64_t* out);
/* lift_poly_to_context (:242-276): src has s->levels limbs; out has t->levels limbs */
void orc_lift(const orc_basis* s, const orc_basis* t, const uint64_t* poly, uint64_t* out);
/* mod_down_eval_to_q (:278-322): poly over e (q prefix then aux), out over q */
void orc_mod_down(const orc_basis* q, const orc_basis* e, const uint64_t* poly, uint64_t* out);
/* keyswitch_rns (:130-157). ksk = [digits][2][q->levels][N]; out = [c0][c1]. Returns 0 or -1 (assertion). */
int orc_keyswitch_qonly(const orc_basis* q, const uint64_t* poly, const uint64_t* ksk, size_t ksk_digits,
                        uint64_t* out);
/* keyswitch_rns_hybrid (:159-194). ksk = [digits][2][e->levels][N]; out = [c0][c1] over q. */
int orc_keyswitch_hybrid(const orc_basis* q, const orc_basis* e, const uint64_t* poly, const uint64_t* ksk,
                         size_t ksk_digits, uint64_t* out);

/* ---- seeded sampling (restated distributions of src/rns/keys.rs:18-151; streams unpinned) ---- */
typedef struct {
    uint64_t s[4];
    int have_spare;
    double spare;
} orc_rng;
void orc_rng_seed(orc_rng* r, uint64_t seed);
uint64_t orc_rng_next(orc_rng* r);
void orc_sample_uniform(const orc_basis* b, orc_rng* r, uint64_t* out);                  /* keys.rs:88-100 */
void orc_sample_error(const orc_basis* b, orc_rng* r, double sigma, uint64_t* out);      /* keys.rs:102-124 */
void orc_sample_ternary(const orc_basis* b, orc_rng* r, uint64_t* out);                  /* keys.rs:126-151 */
void orc_sample_sparse_ternary(const orc_basis* b, orc_rng* r, size_t hamming, uint64_t* out); /* keys.rs:18-52 */
/* rns_keygen (keys.rs:153-167) with a given secret: pk = [p0][p1], error = e */
void orc_keygen_public(const orc_basis* b, orc_rng* r, double sigma, const uint64_t* secret, uint64_t* pk,
                       uint64_t* error);
/* generate_keyswitch_key_rns_hybrid (keyswitching.rs:81-128), RNS-native restatement.
 * source/secret over q (secret may have more limbs: only q->levels are used). out = [digits][2][e->levels][N] */
void orc_keygen_hybrid(const orc_basis* q, const orc_basis* e, orc_rng* r, double sigma, const uint64_t* source,
                       const uint64_t* secret, size_t digits, int noiseless, uint64_t* out);
/* generate_keyswitch_key_rns (keyswitching.rs:44-79): out = [digits][2][q->levels][N] */
void orc_keygen_qonly(const orc_basis* q, orc_rng* r, double sigma, const uint64_t* source, const uint64_t* pk,
                      size_t digits, int noiseless, uint64_t* out);

/* galois_element_for_rotation (src/rns/rotation.rs:14-26) */
uint64_t orc_galois_element_for_rotation(uint64_t poly_degree, int64_t steps);

/* ---- context-level compositions: src/rns/context/{ops,rotation}.rs ---- */
/* relinearize (context/ops.rs:111-122): quad = [c0][c1][c2]; out = [c0][c1] */
int orc_relinearize(const orc_basis* q, const orc_basis* e_or_null, const uint64_t* quad, const uint64_t* ksk,
                    size_t ksk_digits, uint64_t* out);
/* rotate (context/rotation.rs:104-127): ct = [c0][c1] */
int orc_rotate(const orc_basis* q, const orc_basis* e_or_null, const uint64_t* ct, uint64_t galois,
               const uint64_t* ksk, size_t ksk_digits, uint64_t* out);

void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
