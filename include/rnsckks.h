/*
 * rnsckks.h — C ABI of the MI355X (gfx950) RNS-CKKS arithmetic hot path.
 *
 * Drop-in boundary for the NTT, poly-multiply, Hadamard multiply-add, ct x ct multiply,
 * relinearize / keyswitch, rescale and rotate calls of the Rust crate Akram275/rns-ckks-rust.
 * Each entry point names the crate function it replaces (file:line in that crate).
 *
 * Conventions
 *  - Residues are canonical uint64 in [0, q).  A polynomial with nq Q-limbs and np aux limbs is
 *    uint64_t[nq + np][N] in DEVICE memory, limb-major: limb l < nq is reduced mod Q-prime l,
 *    limb nq + k mod aux prime k (the eval-key basis Q[0..nq) || P of
 *    src/rns/context/construction.rs:76-90).  A ciphertext is [c0][c1] (2 polys back to back),
 *    a quadratic ciphertext [c0][c1][c2].  The ciphertext level is limbs - 1.
 *  - All operations are out-of-place unless stated; outputs may alias inputs only where stated.
 *  - `stream` is a hipStream_t (NULL = default stream).  Calls are asynchronous on it.
 *  - Return value: RCK_OK (0) or a negative code; rck_last_error() returns the message of the
 *    last failure on the calling thread.  Where the crate panics on a mismatch (assert!), these
 *    calls return RCK_ERR_INVALID with the crate's assertion message.
 *  - A context owns device scratch reused across calls: calls on one context must not overlap
 *    in time (use one context per concurrently used stream).
 */
#ifndef RNSCKKS_H
#define RNSCKKS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RCK_OK 0
#define RCK_ERR_INVALID (-1)  /* argument / assertion failure (the crate panics or returns Err) */
#define RCK_ERR_DEVICE (-2)   /* HIP runtime error */
#define RCK_ERR_INTERNAL (-3) /* host allocation or other failure */

typedef struct rck_context rck_context;
typedef struct rck_ksk rck_ksk;

/* Message of the last failing call on this thread ("" if none). */
const char* rck_last_error(void);
/* ABI version (major * 100 + minor). */
int rck_version(void);

/* ---- device memory, streams and timing (plumbing for callers without a GPU runtime) ---- */
int rck_malloc(void** ptr, size_t bytes);
int rck_free(void* ptr);
int rck_memcpy_htod(void* dst, const void* src, size_t bytes);
int rck_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int rck_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
int rck_device_synchronize(void);
int rck_stream_create(void** stream);
int rck_stream_destroy(void* stream);
int rck_stream_synchronize(void* stream);
int rck_event_create(void** event);
int rck_event_destroy(void* event);
int rck_event_record(void* event, void* stream);
int rck_event_synchronize(void* event);
int rck_event_elapsed_ms(void* start, void* stop, float* ms);

/* ---- context ----
 * Replaces RnsCkksContext::new(RnsCkksParams) (src/rns/context/construction.rs:9-30) with its
 * validation (src/rns/context/params.rs:48-73, src/rns/basis.rs:32-59, src/math/ntt.rs:17-60;
 * same error messages).  primes/roots: the Q chain (RnsPrime{modulus, primitive_root}),
 * aux_primes/aux_roots: the P chain (may be empty: Q-only keyswitching).  Every odd prime below 2^64
 * is accepted, as in the crate.  Moduli below 2^50 + 2^46 run on the FP64-exact kernels; a larger
 * modulus selects the 64-bit integer arithmetic for the whole context (see rck_context_set_arithmetic). */
int rck_context_create(uint32_t poly_degree, const uint64_t* primes, const uint64_t* roots, uint32_t n_primes,
                       const uint64_t* aux_primes, const uint64_t* aux_roots, uint32_t n_aux, uint32_t scale_bits,
                       rck_context** out);
int rck_context_destroy(rck_context* ctx);
/* num_decomp_digits(total_modulus_bits(Q[0..limbs)), 30) (src/rns/keyswitching.rs:238-240) */
int rck_num_decomp_digits(const rck_context* ctx, uint32_t limbs, uint32_t* digits);
/* Keyswitch method of the keys created (or generated) from now on in this context; no counterpart in
 * the crate, whose single method both produce bit for bit (keyswitching.rs:159-194).
 * RCK_KS_EXACT_CONVOLUTION: hybrid keys at N >= 2^12 are kept as the NTT forms, under two
 * auxiliary primes, of their centred limbs; a digit times a key limb is then an exact integer
 * convolution modulo the two primes' product, summed on the int8 matrix cores (every digit transformed
 * twice instead of once per eval prime; the key, 7 signed bytes per value, takes 1.75 times the
 * memory).  RCK_KS_DIGIT_NTT: keys kept in NTT form mod every eval prime, every digit transformed
 * under every eval prime.  Q-only keys and N <= 2^11 always use the latter.  RCK_KS_AUTO (the default): such keys keep both forms (2.75 times the memory of the
 * digit-NTT form) and every keyswitch takes the faster method for its batch size (the tuned times of
 * both when the tuning table has them, else the exact convolution from four ciphertexts up).
 * rck_ksk_method reports a key's method (RCK_KS_AUTO when it holds both forms). */
#define RCK_KS_EXACT_CONVOLUTION 0
#define RCK_KS_DIGIT_NTT 1
#define RCK_KS_AUTO 2
int rck_context_set_keyswitch_method(rck_context* ctx, int method);
int rck_ksk_method(const rck_ksk* key, int* method);
/* Exact-convolution keyswitch at `limbs` Q limbs: the digit grouping it runs (digits per group, groups).
 * A group's exact sums must stay below half the auxiliary primes' product.  When the digits of typical
 * (uniform) ciphertexts allow it this is one group: the decomposition then measures every ciphertext's
 * digits, and a batch with one whose digits could leave that range is recomputed on the device with
 * the worst-case grouping (every digit at 2^29), so the results are exact for every input. */
int rck_keyswitch_exact_groups(rck_context* ctx, uint32_t limbs, uint32_t* digits_per_group, uint32_t* groups);
/* Arithmetic of the context's kernels; no counterpart in the crate, whose u128 products both reproduce
 * bit for bit (src/math/ntt.rs:132-150).  RCK_ARITH_FP64: FP64-exact modular products (every modulus
 * below 2^50 + 2^46; the default then).  RCK_ARITH_INT64: Montgomery products on 64-bit integers, any
 * modulus below 2^64 (the default when a modulus is at or above 2^50 + 2^46; requesting it on any
 * context is allowed).  Keys belong to the arithmetic they were created under.  The 64-bit integer
 * arithmetic serves the NTT, polynomial, ciphertext, key and keyswitch entry points (rck_ntt_* through
 * rck_rotate, rck_ksk_*, sampling, rck_keygen_public, rck_encrypt / rck_decrypt); the other entry
 * points return RCK_ERR_INVALID on such a context. */
#define RCK_ARITH_FP64 0
#define RCK_ARITH_INT64 1
int rck_context_set_arithmetic(rck_context* ctx, int arithmetic);
int rck_context_arithmetic(const rck_context* ctx, int* arithmetic);
/* galois_element_for_rotation (src/rns/rotation.rs:14-26): 5^steps mod 2N, (5^-1)^|steps| if < 0 */
uint64_t rck_galois_element_for_rotation(uint32_t poly_degree, int64_t steps);

/* ---- L0: NttPlan::forward / inverse (src/math/ntt.rs:78-89) over every limb, in place ----
 * Cyclic transform in natural order with w = g^((q-1)/N): X_k = sum_j x_j w^{jk}; inverse
 * includes N^{-1}.  Bit-identical to the crate's radix-2 NTT. */
int rck_ntt_forward(rck_context* ctx, uint64_t* data, uint32_t nq, uint32_t np, void* stream);
int rck_ntt_inverse(rck_context* ctx, uint64_t* data, uint32_t nq, uint32_t np, void* stream);

/* ---- L1/L2: ModularPolynomial / RnsPolynomial limb ops (src/poly/mod.rs, src/rns/polynomial.rs) ---- */
/* RnsPolynomial::add / sub (polynomial.rs:99-119); out may alias a or b */
int rck_poly_add(rck_context* ctx, const uint64_t* a, const uint64_t* b, uint64_t* out, uint32_t nq, uint32_t np,
                 void* stream);
int rck_poly_sub(rck_context* ctx, const uint64_t* a, const uint64_t* b, uint64_t* out, uint32_t nq, uint32_t np,
                 void* stream);
/* RnsPolynomial::multiply_ntt (polynomial.rs:121-136 -> poly/mod.rs:134-161): negacyclic product */
int rck_poly_multiply_ntt(rck_context* ctx, const uint64_t* a, const uint64_t* b, uint64_t* out, uint32_t nq,
                          uint32_t np, void* stream);
/* No counterpart in the crate: the same negacyclic product as rck_poly_multiply_ntt, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops).  It calls no
 * device allocator: the operands' NTT forms go to a fixed 64 MiB scratch array that every GPU the
 * library is loaded on holds until the process ends, so 2 x (nq + np) x N x 8 bytes must be at most
 * 64 MiB (e.g. up to 64 limbs at N = 2^16), else RCK_ERR_INVALID.  All calls on one GPU share that
 * scratch: calls that overlap on one GPU (other host threads, unordered streams) race on it. */
int rck_poly_multiply_ntt_noalloc(rck_context* ctx, const uint64_t* a, const uint64_t* b, uint64_t* out,
                                  uint32_t nq, uint32_t np, void* stream);
/* RnsPolynomial::apply_galois_automorphism (polynomial.rs:138-145 -> poly/mod.rs:110-130) */
int rck_poly_galois(rck_context* ctx, const uint64_t* a, uint64_t* out, uint32_t nq, uint32_t np, uint64_t galois,
                    void* stream);
/* pointwise (Hadamard) modular multiply-add out = a*b + c per coefficient (the pointwise step of
 * poly/mod.rs:152-154 plus the accumulation of keyswitching.rs:182-187); c may be NULL; out may
 * alias any input */
int rck_poly_mul_add(rck_context* ctx, const uint64_t* a, const uint64_t* b, const uint64_t* c, uint64_t* out,
                     uint32_t nq, uint32_t np, void* stream);

/* ---- L3: RnsCkksContext scheme ops (src/rns/context/ops.rs, src/rns/context/rotation.rs) ---- */
/* RnsCiphertext::add (src/rns/ciphertexts.rs:60-69) */
int rck_ct_add(rck_context* ctx, const uint64_t* lhs, const uint64_t* rhs, uint32_t limbs, uint32_t lhs_scale,
               uint32_t rhs_scale, uint64_t* out, uint32_t* out_scale, void* stream);
/* RnsCkksContext::mult (context/ops.rs:63-70 -> ops.rs:84-100): quadratic out, scale = sum */
int rck_mult(rck_context* ctx, const uint64_t* lhs, const uint64_t* rhs, uint32_t limbs, uint32_t lhs_scale,
             uint32_t rhs_scale, uint64_t* out_quad, uint32_t* out_scale, void* stream);
/* RnsCkksContext::mult_plain (context/ops.rs:20-30 -> ops.rs:54-66): scale := 2 * ct scale */
int rck_mult_plain(rck_context* ctx, const uint64_t* ct, const uint64_t* plain, uint32_t limbs, uint32_t ct_scale,
                   uint64_t* out, uint32_t* out_scale, void* stream);
/* RnsCkksContext::rescale (context/ops.rs:32-37 -> ops.rs:68-82, 135-186): out has limbs-1 limbs,
 * scale -= floor(log2 q_last) (saturating) */
int rck_rescale(rck_context* ctx, const uint64_t* ct, uint32_t limbs, uint32_t scale, uint64_t* out,
                uint32_t* out_scale, void* stream);
/* RnsCkksContext::rescale_quadratic (context/ops.rs:85-90 -> ops.rs:115-133) */
int rck_rescale_quadratic(rck_context* ctx, const uint64_t* quad, uint32_t limbs, uint32_t scale, uint64_t* out,
                          uint32_t* out_scale, void* stream);

/* Keyswitching key upload: RnsKeySwitchingKey {ksk, decomp_bits=30, level, aux_level_count}
 * (src/rns/keyswitching.rs:12-42) plus the RnsRotationKey galois element (src/rns/rotation.rs:8-12;
 * pass 0 for a relinearization key).  key: uint64_t[digits][2][level+1+aux][N] in HOST memory,
 * coefficient form (ksk[i].c0, ksk[i].c1).  Converted once to the NTT domain and kept in HBM. */
int rck_ksk_create(rck_context* ctx, const uint64_t* key, uint32_t digits, uint32_t level, uint32_t aux_level_count,
                   uint64_t galois_element, rck_ksk** out);
/* same, from DEVICE memory */
int rck_ksk_create_device(rck_context* ctx, const uint64_t* key, uint32_t digits, uint32_t level,
                          uint32_t aux_level_count, uint64_t galois_element, rck_ksk** out);
int rck_ksk_destroy(rck_ksk* key);

/* RnsCkksContext::keyswitch (context/ops.rs:97-109 -> keyswitching.rs:130-157 Q-only,
 * :159-194 hybrid): poly has `limbs` Q-limbs; out = [c0][c1] over the same limbs;
 * *out_scale = limbs (the crate's keyswitch returns scale_bits = level count) */
int rck_keyswitch(rck_context* ctx, const uint64_t* poly, uint32_t limbs, const rck_ksk* key, uint64_t* out,
                  uint32_t* out_scale, void* stream);
/* RnsCkksContext::relinearize (context/ops.rs:111-122): (c0 + ks.c0, c1 + ks.c1), scale kept */
int rck_relinearize(rck_context* ctx, const uint64_t* quad, uint32_t limbs, uint32_t scale, const rck_ksk* key,
                    uint64_t* out, uint32_t* out_scale, void* stream);
/* RnsCkksContext::mult_relinearized (context/ops.rs:124-133): mult, relinearize, rescale;
 * out has limbs-1 limbs */
int rck_mult_relinearized(rck_context* ctx, const uint64_t* lhs, const uint64_t* rhs, uint32_t limbs,
                          uint32_t lhs_scale, uint32_t rhs_scale, const rck_ksk* key, uint64_t* out,
                          uint32_t* out_scale, void* stream);
/* RnsCiphertext::apply_galois_automorphism (ciphertexts.rs:71-78) */
int rck_ct_galois(rck_context* ctx, const uint64_t* ct, uint32_t limbs, uint64_t galois, uint64_t* out,
                  void* stream);
/* RnsCkksContext::rotate / conjugate (context/rotation.rs:104-135) with the key's galois element:
 * (sigma(c0) + ks.c0, ks.c1), scale kept; key level must equal the ciphertext level */
int rck_rotate(rck_context* ctx, const uint64_t* ct, uint32_t limbs, uint32_t scale, const rck_ksk* key,
               uint64_t* out, uint32_t* out_scale, void* stream);

/* ---- seeded sampling, key generation, encryption (client side of the hot path) ----
 * The crate draws from rand::thread_rng (unseeded).  Here one seeded stream (xoshiro256** seeded
 * through splitmix64) is consumed in the crate's order, so a seed fixes every key and ciphertext:
 *   uniform: for each limb, for each coefficient, next() % q          (src/rns/keys.rs:88-100)
 *   error:   N draws of round(sigma * z), z ~ N(0,1) by Box-Muller with the pair's second normal
 *            kept for the next draw, lifted to every limb              (src/rns/keys.rs:102-124)
 *   ternary: N draws of next() % 3 -> 0, 1, -1                         (src/rns/keys.rs:126-151)
 *   sparse ternary: idx = next() % N; if unused: next() & 1 ? 1 : -1   (src/rns/keys.rs:18-52)
 * Sampling runs on the host; the arithmetic (a*s, pk*u, key plaintexts, NTT-form upload) on the
 * device.  Polynomials are DEVICE buffers in the layout above.  These calls synchronize. */
typedef struct rck_rng rck_rng;
int rck_rng_create(uint64_t seed, rck_rng** out);
int rck_rng_destroy(rck_rng* rng);
uint64_t rck_rng_next(rck_rng* rng);
int rck_sample_uniform(rck_context* ctx, rck_rng* rng, uint64_t* out, uint32_t nq, uint32_t np);
int rck_sample_error(rck_context* ctx, rck_rng* rng, double sigma, uint64_t* out, uint32_t nq, uint32_t np);
int rck_sample_ternary(rck_context* ctx, rck_rng* rng, uint64_t* out, uint32_t nq, uint32_t np);
/* RnsSecretKey::sample_sparse_ternary (src/rns/keys.rs:18-52) */
int rck_sample_sparse_ternary(rck_context* ctx, rck_rng* rng, uint32_t hamming_weight, uint64_t* out, uint32_t nq,
                              uint32_t np);
/* rns_keygen's public key (src/rns/keys.rs:153-167): pk1 = uniform, pk0 = e - pk1 * s; pk: [pk0][pk1]
 * over `limbs` limbs. */
int rck_keygen_public(rck_context* ctx, rck_rng* rng, double sigma, const uint64_t* secret, uint32_t limbs,
                      uint64_t* pk);
/* generate_relinearization_key_at_level (src/rns/context/keys.rs:46-63): source s^2 at `level`;
 * hybrid (keyswitching.rs:81-128) when the context has aux primes, else Q-only encryption under pk
 * (keyswitching.rs:44-79; pk has pk_limbs limbs per component).  The secret (L limbs) must be a
 * small polynomial (|coefficients| < q_0 / 2; the crate's secrets are ternary).  The key is created
 * NTT-resident, as by rck_ksk_create. */
int rck_keygen_relin(rck_context* ctx, rck_rng* rng, double sigma, const uint64_t* secret, const uint64_t* pk,
                     uint32_t pk_limbs, uint32_t level, rck_ksk** out);
/* generate_rotation_key_at_level (src/rns/context/rotation.rs:43-60): source s(X^g),
 * g = galois_element_for_rotation(N, steps). */
int rck_keygen_rotation(rck_context* ctx, rck_rng* rng, double sigma, const uint64_t* secret, const uint64_t* pk,
                        uint32_t pk_limbs, int64_t steps, uint32_t level, rck_ksk** out);
/* The key in the crate's coefficient form, uint64_t[digits][2][level+1+aux][N] in HOST memory. */
int rck_ksk_export(rck_context* ctx, const rck_ksk* key, uint64_t* out);
int rck_ksk_info(const rck_ksk* key, uint32_t* digits, uint32_t* level, uint32_t* aux, uint64_t* galois);
/* encrypt_rns (src/rns/ops.rs:8-42) at `limbs` limbs: draws u (ternary), e0, e1 (errors);
 * out = [pk0 u + e0 + m][pk1 u + e1]; pk is truncated to `limbs` (pk_limbs limbs per component). */
int rck_encrypt(rck_context* ctx, rck_rng* rng, double sigma, const uint64_t* plain, const uint64_t* pk,
                uint32_t pk_limbs, uint32_t limbs, uint64_t* out);
/* decrypt_rns (src/rns/ops.rs:44-52): out = c0 + c1 * s over `limbs` limbs */
int rck_decrypt(rck_context* ctx, const uint64_t* ct, uint32_t limbs, const uint64_t* secret, uint64_t* out,
                void* stream);

/* ---- encode / decode (host FFT; client side of the hot path) ----
 * RnsCkksContext::encode_at_level_and_scale (src/rns/context/encoding.rs:24-35 ->
 * encode_integer_coeffs, src/rns/encoding.rs:9-31): slots (re[i] + j im[i], i < N/2; im may be NULL)
 * -> integer coefficients round(2^scale_bits * Re(twisted inverse embedding)) as i128 -> residues of
 * the first `limbs` primes, written to DEVICE memory.  Bit-exact parity with the crate depends on
 * num-complex 0.4's rounding and is not pinned; the CKKS error bounds are. */
int rck_encode(rck_context* ctx, const double* re, const double* im, uint32_t slots, uint32_t limbs,
               uint32_t scale_bits, uint64_t* out);
/* RnsCkksContext::decode_at_scale (src/rns/context/encoding.rs:37-48 -> decode_integer_coeffs,
 * src/rns/encoding.rs:33-66): CRT-reconstruct the `limbs`-limb DEVICE polynomial, centre against
 * floor(Q/2), scale by 2^-scale_bits and evaluate at the slots; re/im (HOST, N/2 each; im may be NULL). */
int rck_decode(rck_context* ctx, const uint64_t* poly, uint32_t limbs, uint32_t scale_bits, double* re, double* im);

/* ---- Paterson-Stockmeyer polynomial evaluation (src/algorithms/paterson_stockmeyer.rs:28-350) ----
 * coeffs[i] is the real coefficient of x^i (count >= 1).  relin_keys[level] (n_keys entries, NULL =
 * missing) are relinearization keys at those levels, as the crate's BTreeMap<usize, key>; the keys
 * needed are those of paterson_stockmeyer_required_levels.  The result has *out_limbs limbs and scale
 * *out_scale (written to out: [c0][c1]).  The crate's panics (degree too large for the level budget,
 * missing key, working scale below the input scale) return RCK_ERR_INVALID with its message.  The
 * plaintexts the algorithm builds are constants: multiplying by / adding them is done exactly as
 * scalar products / a coefficient-0 add, which equals the crate's encoded multiply_plain / add. */
int rck_ps_block_size(uint32_t term_count, uint32_t* block_size);
int rck_ps_required_levels(uint32_t ciphertext_level, uint32_t term_count, uint32_t* levels, uint32_t cap,
                           uint32_t* count);
/* Result metadata, ct x ct multiplications and device working-set bytes of an evaluation, without running it. */
int rck_ps_plan(rck_context* ctx, uint32_t limbs, uint32_t scale_bits, const double* coeffs, uint32_t count,
                uint32_t* out_limbs, uint32_t* out_scale, uint32_t* mults, uint64_t* footprint);
int rck_ps_eval(rck_context* ctx, const uint64_t* ct, uint32_t limbs, uint32_t scale_bits, const double* coeffs,
                uint32_t count, const rck_ksk* const* relin_keys, uint32_t n_keys, uint64_t* out,
                uint32_t* out_limbs, uint32_t* out_scale, void* stream);

/* ---- baby-step giant-step diagonal transform (src/ckks/bootstrapping/linear_transform.rs:22-253,
 *      src/algorithms/packed_diagonal.rs:29-77) ----
 * diagonal_transform_bsgs_rotation_steps (:27-47): the rotation steps whose keys the transform uses. */
int rck_bsgs_rotation_steps(uint32_t dim, uint32_t* steps, uint32_t cap, uint32_t* count);
/* pack_diagonal_transform_plaintexts' giant-step diagonals (:102-116, 161-280) of a real dim x dim
 * matrix (HOST, row-major), encoded at (level, scale_bits) with rck_encode's encoder and kept in HBM
 * in NTT form.  dim: a power of two, at most N/2 and 4096. */
typedef struct rck_diag rck_diag;
int rck_diag_pack(rck_context* ctx, const double* matrix, uint32_t dim, uint32_t level, uint32_t scale_bits,
                  rck_diag** out);
int rck_diag_info(const rck_diag* diag, uint32_t* dim, uint32_t* baby_step, uint32_t* giant_steps, uint32_t* level,
                  uint32_t* scale_bits);
/* The encoded diagonals in coefficient form, uint64_t[giant_steps][baby_step][level+1][N] (HOST;
 * diagonal steps >= dim are zero). */
int rck_diag_export(const rck_diag* diag, uint64_t* out);
int rck_diag_destroy(rck_diag* diag);
/* apply_diagonal_linear_transform (:118-144) on `batch` ciphertexts cts[b] = [c0][c1] (DEVICE,
 * contiguous, `limbs` limbs, scale scale_bits) with keys_by_step[step] (n_keys entries, NULL = missing;
 * rotation keys at level limbs-1): out[b] (limbs-1 limbs, contiguous), *out_scale = 2 * scale_bits -
 * floor(log2 q_{limbs-1}) (mult_plain doubles the scale, then one rescale).  Synchronizes `stream`. */
int rck_bsgs_matvec(rck_context* ctx, const uint64_t* cts, uint32_t batch, uint32_t limbs, uint32_t scale_bits,
                    const rck_diag* diag, const rck_ksk* const* keys_by_step, uint32_t n_keys, uint64_t* out,
                    uint32_t* out_scale, void* stream);

/* ---- request server (no counterpart in the crate: it serves many paterson_stockmeyer_eval calls) ----
 * Requests are Paterson-Stockmeyer evaluations (same arguments and result as rck_ps_eval).  A worker
 * thread admits them in arrival order while their device working sets (rck_ps_plan's footprint) fit
 * hbm_budget_bytes (and at most max_active run, 0 = unlimited), advances all admitted requests
 * together, and runs the ct x ct multiplications that share a ring, a level and a key as one
 * batched launch.  A result is bit-identical to rck_ps_eval of the same request.  `ct` and `out`
 * are caller-owned DEVICE buffers that must stay valid until the request completes.  Register every
 * ring (context + relinearization keys by level) before the first submit. */
typedef struct rck_server rck_server;
/* lanes (1..8): streams on which independent multiplication groups of a round run concurrently,
 * each with its own scratch workspace per ring. */
int rck_server_create(uint64_t hbm_budget_bytes, uint32_t max_active, uint32_t lanes, rck_server** out);
int rck_server_register_ring(rck_server* server, rck_context* ctx, const rck_ksk* const* relin_keys,
                             uint32_t n_keys);
int rck_server_submit(rck_server* server, rck_context* ctx, const uint64_t* ct, uint32_t limbs, uint32_t scale_bits,
                      const double* coeffs, uint32_t count, uint64_t* out, uint64_t* request_id);
/* Waits up to timeout_s for a request; *done = 1 when it completed (then out_limbs / out_scale /
 * latency_ms = submit-to-completion time are set). */
int rck_server_wait(rck_server* server, uint64_t request_id, double timeout_s, int* done, uint32_t* out_limbs,
                    uint32_t* out_scale, double* latency_ms);
/* Waits until every submitted request completed (error on time-out). */
int rck_server_drain(rck_server* server, double timeout_s);
/* JSON: completed, p50/p99/mean/max latency (ms), queueing, throughput, rounds, batched launches,
 * peak working-set bytes, histogram of multiplications per launch. */
int rck_server_stats(rck_server* server, char* buf, size_t len);
int rck_server_destroy(rck_server* server);

/* ---- versioned binary format (no counterpart in the crate, which has no serialization) ----
 * "RCK1" header (kind, ring degree, counts, limbs, scale or key level, galois element, a fingerprint
 * of the prime chain, payload size) followed by canonical u64 residues, limb-major.  Buffers are HOST
 * memory; polys are DEVICE memory.  Passing buf = NULL returns the needed size.  Deserialization
 * checks the magic, version, kind, ring degree, limb counts, prime-chain fingerprint and sizes. */
int rck_serialize_polys(rck_context* ctx, const uint64_t* polys, uint32_t count, uint32_t nq, uint32_t np,
                        uint32_t scale_bits, uint8_t* buf, size_t len, size_t* needed);
int rck_deserialize_polys(rck_context* ctx, const uint8_t* buf, size_t len, uint64_t* polys, size_t capacity,
                          uint32_t* count, uint32_t* nq, uint32_t* np, uint32_t* scale_bits);
/* Keys travel in the crate's coefficient form (rck_ksk_export / rck_ksk_create). */
int rck_ksk_serialize(rck_context* ctx, const rck_ksk* key, uint8_t* buf, size_t len, size_t* needed);
int rck_ksk_deserialize(rck_context* ctx, const uint8_t* buf, size_t len, rck_ksk** out);

/* ---- batched executor (no single counterpart in the crate: it runs many of the calls above) ----
 * Independent operations at any levels, of mixed kinds, are planned into grouped launches
 * (same kind, level and key -> one group whose kernels process all its ciphertexts together and
 * stream the key once) and captured in one hipGraph.  Results are bit-identical to the per-op
 * calls.  Operands stay caller-owned device buffers; ops of one batch must not alias each other's
 * outputs.  Output scales follow the per-op rules (mult: sum of scales, relinearize/rotate: kept,
 * mult_plain: doubled, rescale: minus floor(log2 q_last)). */
#define RCK_OP_MULT_RELIN 1         /* out = relinearize(mult(in0, in1), key)        (context/ops.rs:63-122) */
#define RCK_OP_ROTATE 2             /* out = rotate(in0, key) with the key's galois  (context/rotation.rs:104-127) */
#define RCK_OP_MULT_PLAIN 3         /* out = mult_plain(in0, in1)                    (context/ops.rs:20-30) */
#define RCK_OP_MULT_RELIN_RESCALE 4 /* out = mult_relinearized(in0, in1, key)        (context/ops.rs:124-133) */
typedef struct rck_batch rck_batch;
int rck_batch_create(rck_context* ctx, rck_batch** out);
/* limbs: the operands' Q-limb count; in1 unused for RCK_OP_ROTATE; key unused for RCK_OP_MULT_PLAIN */
int rck_batch_add(rck_batch* batch, int op, uint32_t limbs, const uint64_t* in0, const uint64_t* in1,
                  const rck_ksk* key, uint64_t* out);
/* Plans the groups, optionally autotunes each group's digit-MAC launch shape (tune != 0) by timing
 * candidates on the group's operands, runs the batch once eagerly and captures it in a hipGraph.
 * A batch with no operations is rejected (RCK_ERR_INVALID). */
int rck_batch_compile(rck_batch* batch, int tune);
/* Launches the captured graph on `stream` (NULL: the batch's own stream). */
int rck_batch_run(rck_batch* batch, void* stream);
/* Runs the planned launches directly on the batch's stream, without the graph (same results;
 * lets per-kernel timing see the individual launches). */
int rck_batch_run_eager(rck_batch* batch);
int rck_batch_synchronize(rck_batch* batch);
/* JSON: groups, their tuned launch shapes and timings. */
int rck_batch_report(rck_batch* batch, char* buf, size_t len);
int rck_batch_destroy(rck_batch* batch);

/* Autotuner for the per-op path: times the candidate digit-MAC launch shapes (eval-limb grouping,
 * digit splits, rows per workgroup, key prefetch) of a keyswitch of `batch` ciphertexts at `limbs`
 * limbs with `key`, and records the fastest in the process-wide tuning table that rck_keyswitch,
 * rck_relinearize, rck_rotate and batches consult.  buf (optional) receives the choice as JSON.
 * Results are identical for every shape; only the launch geometry changes. */
int rck_tune_keyswitch(rck_context* ctx, uint32_t limbs, const rck_ksk* key, uint32_t batch, char* buf, size_t len);
/* The process-wide tuning table that per-op calls and rck_batch_compile consult (entries keyed by
 * ring degree, limbs, eval limbs, digits and batch size), as text: one line per entry,
 * "log2N limbs eval_limbs digits batch group digit_splits mac_rows load_schedule mac_width ms" (an
 * exact-convolution entry has load_schedule 7, mac_rows = slots per wave of its MAC, mac_width 16).  Export writes
 * the table into buf (needed = bytes including the terminating NUL; buf may be NULL to query);
 * import validates every line, then merges the entries (a tuned group is not re-timed).  No
 * counterpart in the crate (it has no GPU launch shapes); lets a process reuse earlier tuning. */
int rck_tune_table_export(char* buf, size_t len, size_t* needed);
int rck_tune_table_import(const char* text, uint32_t* entries);

/* ---- diagnostics (no counterpart in the crate) ----
 * Per-kernel timing: while enabled, every kernel launch is bracketed by HIP events on the stream it
 * is launched on.  rck_profile_report synchronizes the device and writes a JSON object
 * {"kernel": [launches, total_ms], ...} for the launches since the previous report. */
int rck_profile_enable(int on);
/* A plain streaming copy kernel of `bytes` (multiple of 32): the known-volume reference for
 * calibrating hardware byte counters (FETCH_SIZE / WRITE_SIZE). */
int rck_diag_copy(void* dst, const void* src, size_t bytes, void* stream);
int rck_profile_report(char* buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif /* RNSCKKS_H */
