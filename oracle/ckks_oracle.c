/*
 * CPU oracle for the RNS-CKKS arithmetic path — TEST INFRASTRUCTURE ONLY.
 * See ckks_oracle.h for scope and pinning.  Citations are files of the reference crate
 * Akram275/rns-ckks-rust.
 */
#include "ckks_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;
typedef __int128 i128;

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

static void set_err(char* err, size_t len, const char* msg) {
    if (err && len) {
        snprintf(err, len, "%s", msg);
    }
}

/* ------------------------------------------------------------------ */
/* src/math/ntt.rs:132-188                                             */
/* ------------------------------------------------------------------ */
uint64_t orc_mod_add(uint64_t a, uint64_t b, uint64_t q) {
    u128 s = (u128)a + b;
    return s >= q ? (uint64_t)(s - q) : (uint64_t)s;
}

uint64_t orc_mod_sub(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : q - (b - a); }

uint64_t orc_mod_mul(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

uint64_t orc_mod_pow(uint64_t base, uint64_t exp, uint64_t q) {
    uint64_t result = 1;
    while (exp > 0) {
        if (exp & 1) result = orc_mod_mul(result, base, q);
        base = orc_mod_mul(base, base, q);
        exp >>= 1;
    }
    return result;
}

uint64_t orc_mod_inv(uint64_t v, uint64_t q, int* ok) {
    if (v == 0) {
        *ok = 0;
        return 0;
    }
    i128 t = 0, nt = 1, r = q, nr = v;
    while (nr != 0) {
        i128 quo = r / nr;
        i128 tmp = t - quo * nt;
        t = nt;
        nt = tmp;
        tmp = r - quo * nr;
        r = nr;
        nr = tmp;
    }
    if (r != 1) {
        *ok = 0;
        return 0;
    }
    if (t < 0) t += q;
    *ok = 1;
    return (uint64_t)(t % (i128)q);
}

/* ------------------------------------------------------------------ */
/* NttPlan::new  src/math/ntt.rs:17-60                                 */
/* ------------------------------------------------------------------ */
int orc_ntt_plan_new(uint64_t size, uint64_t modulus, uint64_t g, orc_ntt_plan* out, char* err, size_t err_len) {
    if (size == 0 || (size & (size - 1)) != 0) {
        set_err(err, err_len, "NTT size must be a non-zero power of two");
        return -1;
    }
    if (modulus < 3) {
        set_err(err, err_len, "modulus must be an odd prime larger than 2");
        return -1;
    }
    if ((modulus - 1) % size != 0) {
        set_err(err, err_len, "size must divide modulus - 1");
        return -1;
    }
    if ((modulus - 1) % (2 * size) != 0) {
        set_err(err, err_len, "2 * size must divide modulus - 1 for negacyclic NTTs");
        return -1;
    }
    uint64_t exponent = (modulus - 1) / size;
    uint64_t root = orc_mod_pow(g % modulus, exponent, modulus);
    if (orc_mod_pow(root, size, modulus) != 1) {
        set_err(err, err_len, "primitive_root does not generate a size-th root of unity");
        return -1;
    }
    if (size > 1 && orc_mod_pow(root, size / 2, modulus) == 1) {
        set_err(err, err_len, "root is not primitive for the requested size");
        return -1;
    }
    int ok;
    uint64_t root_inv = orc_mod_inv(root, modulus, &ok);
    if (!ok) {
        set_err(err, err_len, "root is not invertible");
        return -1;
    }
    uint64_t size_inv = orc_mod_inv(size % modulus, modulus, &ok);
    if (!ok) {
        set_err(err, err_len, "size is not invertible modulo modulus");
        return -1;
    }
    uint64_t psi = orc_mod_pow(g % modulus, (modulus - 1) / (2 * size), modulus);
    if (orc_mod_pow(psi, 2 * size, modulus) != 1 || orc_mod_pow(psi, size, modulus) != modulus - 1) {
        set_err(err, err_len, "primitive_root does not generate a valid negacyclic twiddle");
        return -1;
    }
    uint64_t psi_inv = orc_mod_inv(psi, modulus, &ok);
    if (!ok) {
        set_err(err, err_len, "negacyclic twiddle is not invertible");
        return -1;
    }
    out->modulus = modulus;
    out->size = size;
    out->root = root;
    out->root_inv = root_inv;
    out->psi = psi;
    out->psi_inv = psi_inv;
    out->size_inv = size_inv;
    return 0;
}

/* bit_reverse_permute  src/math/ntt.rs:116-130 */
static void bit_reverse_permute(uint64_t* v, size_t n) {
    size_t j = 0;
    for (size_t i = 1; i < n; i++) {
        size_t bit = n >> 1;
        while (j & bit) {
            j ^= bit;
            bit >>= 1;
        }
        j ^= bit;
        if (i < j) {
            uint64_t t = v[i];
            v[i] = v[j];
            v[j] = t;
        }
    }
}

/* ntt_in_place  src/math/ntt.rs:92-114 (radix-2 DIT, twiddle recurrence w *= w_len) */
static void ntt_in_place(uint64_t* v, size_t n, uint64_t q, uint64_t root) {
    bit_reverse_permute(v, n);
    for (size_t len = 2; len <= n; len <<= 1) {
        size_t step = n / len;
        uint64_t w_len = orc_mod_pow(root, step, q);
        for (size_t start = 0; start < n; start += len) {
            uint64_t w = 1;
            for (size_t i = 0; i < len / 2; i++) {
                uint64_t u = v[start + i];
                uint64_t t = orc_mod_mul(v[start + i + len / 2], w, q);
                v[start + i] = orc_mod_add(u, t, q);
                v[start + i + len / 2] = orc_mod_sub(u, t, q);
                w = orc_mod_mul(w, w_len, q);
            }
        }
    }
}

/* NttPlan::forward / inverse  src/math/ntt.rs:78-89 */
void orc_ntt_forward(const orc_ntt_plan* p, uint64_t* values) { ntt_in_place(values, p->size, p->modulus, p->root); }

void orc_ntt_inverse(const orc_ntt_plan* p, uint64_t* values) {
    ntt_in_place(values, p->size, p->modulus, p->root_inv);
    for (size_t i = 0; i < p->size; i++) values[i] = orc_mod_mul(values[i], p->size_inv, p->modulus);
}

/* ------------------------------------------------------------------ */
/* ModularPolynomial  src/poly/mod.rs                                  */
/* ------------------------------------------------------------------ */
void orc_poly_reduce(const uint64_t* in, uint64_t* out, size_t n, uint64_t q) {
    for (size_t i = 0; i < n; i++) out[i] = in[i] % q;
}

void orc_poly_add(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n, uint64_t q) {
    for (size_t i = 0; i < n; i++) out[i] = orc_mod_add(a[i], b[i], q) % q;
}

void orc_poly_sub(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n, uint64_t q) {
    for (size_t i = 0; i < n; i++) out[i] = orc_mod_sub(a[i], b[i], q) % q;
}

/* multiply_naive  src/poly/mod.rs:62-107 (schoolbook; the nonzero filtering only skips zero terms) */
void orc_poly_multiply_naive(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n, uint64_t q) {
    uint64_t* r = (uint64_t*)calloc(n, sizeof(uint64_t));
    for (size_t i = 0; i < n; i++) {
        if (a[i] == 0) continue;
        for (size_t j = 0; j < n; j++) {
            if (b[j] == 0) continue;
            uint64_t prod = orc_mod_mul(a[i], b[j], q);
            size_t k = i + j;
            if (k < n)
                r[k] = orc_mod_add(r[k], prod, q);
            else
                r[k - n] = orc_mod_sub(r[k - n], prod, q);
        }
    }
    for (size_t i = 0; i < n; i++) out[i] = r[i] % q;
    free(r);
}

/* apply_galois_automorphism  src/poly/mod.rs:110-130 */
void orc_poly_galois(const uint64_t* a, uint64_t* out, size_t n, uint64_t q, uint64_t galois) {
    uint64_t* r = (uint64_t*)calloc(n, sizeof(uint64_t));
    uint64_t two_n = 2 * (uint64_t)n;
    for (size_t i = 0; i < n; i++) {
        if (a[i] == 0) continue;
        uint64_t mapped = (uint64_t)(((u128)i * galois) % two_n);
        if (mapped < n)
            r[mapped] = orc_mod_add(r[mapped], a[i], q);
        else
            r[mapped - n] = orc_mod_sub(r[mapped - n], a[i], q);
    }
    for (size_t i = 0; i < n; i++) out[i] = r[i] % q;
    free(r);
}

/* twist_in_place  src/poly/mod.rs:169-175 */
static void twist_in_place(uint64_t* c, size_t n, uint64_t tw, uint64_t q) {
    uint64_t power = 1;
    for (size_t i = 0; i < n; i++) {
        c[i] = orc_mod_mul(c[i], power, q);
        power = orc_mod_mul(power, tw, q);
    }
}

/* multiply_ntt  src/poly/mod.rs:134-161 */
void orc_poly_multiply_ntt(const uint64_t* a, const uint64_t* b, uint64_t* out, const orc_ntt_plan* p) {
    size_t n = p->size;
    uint64_t q = p->modulus;
    uint64_t* lhs = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint64_t* rhs = (uint64_t*)malloc(n * sizeof(uint64_t));
    memcpy(lhs, a, n * sizeof(uint64_t));
    memcpy(rhs, b, n * sizeof(uint64_t));
    twist_in_place(lhs, n, p->psi, q);
    twist_in_place(rhs, n, p->psi, q);
    orc_ntt_forward(p, lhs);
    orc_ntt_forward(p, rhs);
    for (size_t i = 0; i < n; i++) lhs[i] = orc_mod_mul(lhs[i], rhs[i], q);
    orc_ntt_inverse(p, lhs);
    twist_in_place(lhs, n, p->psi_inv, q);
    for (size_t i = 0; i < n; i++) out[i] = lhs[i] % q;
    free(lhs);
    free(rhs);
}

/* ------------------------------------------------------------------ */
/* RnsContext  src/rns/basis.rs:32-59                                  */
/* ------------------------------------------------------------------ */
int orc_basis_new(size_t n, const uint64_t* moduli, const uint64_t* roots, size_t levels, orc_basis* out, char* err,
                  size_t err_len) {
    memset(out, 0, sizeof(*out));
    if (n == 0 || (n & (n - 1)) != 0) {
        set_err(err, err_len, "poly_degree must be a non-zero power of two");
        return -1;
    }
    if (levels == 0) {
        set_err(err, err_len, "RNS basis must contain at least one prime");
        return -1;
    }
    for (size_t i = 0; i < levels; i++) {
        for (size_t j = 0; j < i; j++) {
            if (moduli[j] == moduli[i]) {
                char buf[96];
                snprintf(buf, sizeof buf, "duplicate RNS modulus %llu", (unsigned long long)moduli[i]);
                set_err(err, err_len, buf);
                return -1;
            }
        }
    }
    out->n = n;
    out->levels = levels;
    out->moduli = (uint64_t*)malloc(levels * sizeof(uint64_t));
    out->roots = (uint64_t*)malloc(levels * sizeof(uint64_t));
    out->plans = (orc_ntt_plan*)malloc(levels * sizeof(orc_ntt_plan));
    memcpy(out->moduli, moduli, levels * sizeof(uint64_t));
    memcpy(out->roots, roots, levels * sizeof(uint64_t));
    for (size_t i = 0; i < levels; i++) {
        char perr[160];
        if (orc_ntt_plan_new(n, moduli[i], roots[i], &out->plans[i], perr, sizeof perr) != 0) {
            char buf[256];
            snprintf(buf, sizeof buf, "invalid RNS prime %llu: %s", (unsigned long long)moduli[i], perr);
            set_err(err, err_len, buf);
            orc_basis_free(out);
            return -1;
        }
    }
    return 0;
}

void orc_basis_free(orc_basis* b) {
    free(b->moduli);
    free(b->roots);
    free(b->plans);
    memset(b, 0, sizeof(*b));
}

/* ------------------------------------------------------------------ */
/* base-2^k digit vectors: keyswitching.rs:469-538                     */
/* ------------------------------------------------------------------ */
#define ORC_MAXD 160 /* up to 4800 bits at base 2^30 */
typedef struct {
    size_t len;
    uint64_t d[ORC_MAXD];
} digits_t;

/* mul_add_small_base  keyswitching.rs:469-480 */
static void mul_add_small_base(digits_t* v, uint64_t mult, uint64_t addend, uint64_t base) {
    u128 carry = addend;
    for (size_t i = 0; i < v->len; i++) {
        u128 value = (u128)v->d[i] * mult + carry;
        v->d[i] = (uint64_t)(value % base);
        carry = value / base;
    }
    while (carry > 0) {
        if (v->len >= ORC_MAXD) abort();
        v->d[v->len++] = (uint64_t)(carry % base);
        carry /= base;
    }
}

/* div2_base_digits  keyswitching.rs:482-489 */
static void div2_base_digits(digits_t* v, uint64_t base) {
    uint64_t carry = 0;
    for (size_t k = v->len; k-- > 0;) {
        u128 value = (u128)carry * base + v->d[k];
        v->d[k] = (uint64_t)(value / 2);
        carry = (uint64_t)(value % 2);
    }
}

/* trim_base_digits  keyswitching.rs:491-495 */
static void trim_base_digits(digits_t* v) {
    while (v->len > 1 && v->d[v->len - 1] == 0) v->len--;
}

/* cmp_base_digits  keyswitching.rs:497-509 */
static int cmp_base_digits(const digits_t* a, const digits_t* b) {
    size_t la = 0, lb = 0;
    for (size_t i = a->len; i-- > 0;)
        if (a->d[i] != 0) {
            la = i + 1;
            break;
        }
    for (size_t i = b->len; i-- > 0;)
        if (b->d[i] != 0) {
            lb = i + 1;
            break;
        }
    if (la != lb) return la < lb ? -1 : 1;
    for (size_t i = la; i-- > 0;) {
        if (a->d[i] != b->d[i]) return a->d[i] < b->d[i] ? -1 : 1;
    }
    return 0;
}

/* sub_base_digits  keyswitching.rs:511-529 */
static void sub_base_digits(const digits_t* a, const digits_t* b, uint64_t base, digits_t* out) {
    i128 borrow = 0;
    out->len = a->len;
    for (size_t i = 0; i < a->len; i++) {
        i128 ad = a->d[i];
        i128 bd = i < b->len ? b->d[i] : 0;
        i128 value = ad - bd - borrow;
        if (value < 0) {
            value += base;
            borrow = 1;
        } else {
            borrow = 0;
        }
        out->d[i] = (uint64_t)value;
    }
    trim_base_digits(out);
}

static uint64_t add_mod_u64(uint64_t a, uint64_t b, uint64_t q) { return orc_mod_add(a, b, q); }

/* sub_mod_u64  keyswitching.rs:549-557 (pre-reduces both operands) */
static uint64_t sub_mod_u64(uint64_t a, uint64_t b, uint64_t q) {
    a %= q;
    b %= q;
    return a >= b ? a - b : q - (b - a);
}

/* base_digits_mod_modulus  keyswitching.rs:531-538 */
static uint64_t base_digits_mod_modulus(const digits_t* v, uint64_t base, uint64_t q) {
    uint64_t acc = 0;
    for (size_t k = v->len; k-- > 0;) {
        acc = (uint64_t)(((u128)acc * (base % q)) % q);
        acc = add_mod_u64(acc, v->d[k] % q, q);
    }
    return acc;
}

/* BasisConversionPrecomp  keyswitching.rs:324-363 */
typedef struct {
    size_t count;
    const uint64_t* moduli;
    uint64_t* inverses; /* [count][count], lower triangle: inverses[i*count + j], j < i */
    digits_t modulus_digits;
    digits_t half_modulus_digits;
    uint64_t base;
} bconv_t;

static void bconv_new(bconv_t* p, const uint64_t* moduli, size_t count, uint64_t base) {
    p->count = count;
    p->moduli = moduli;
    p->base = base;
    p->inverses = (uint64_t*)calloc(count * count + 1, sizeof(uint64_t));
    for (size_t i = 0; i < count; i++) {
        for (size_t j = 0; j < i; j++) {
            int ok;
            p->inverses[i * count + j] = orc_mod_inv(moduli[j] % moduli[i], moduli[i], &ok);
            if (!ok) {
                fprintf(stderr, "RNS moduli must be pairwise coprime\n");
                abort();
            }
        }
    }
    p->modulus_digits.len = 1;
    p->modulus_digits.d[0] = 1;
    for (size_t i = 0; i < count; i++) mul_add_small_base(&p->modulus_digits, moduli[i], 0, base);
    trim_base_digits(&p->modulus_digits);
    p->half_modulus_digits = p->modulus_digits;
    div2_base_digits(&p->half_modulus_digits, base);
    trim_base_digits(&p->half_modulus_digits);
}

static void bconv_free(bconv_t* p) { free(p->inverses); }

/* mixed_radix_digits  keyswitching.rs:397-410 (Garner) */
static void mixed_radix_digits(const uint64_t* residues, const bconv_t* p, uint64_t* out) {
    size_t n = p->count;
    for (size_t i = 0; i < n; i++) out[i] = residues[i];
    for (size_t i = 1; i < n; i++) {
        uint64_t q = p->moduli[i];
        uint64_t value = out[i];
        for (size_t j = 0; j < i; j++) {
            value = sub_mod_u64(value, out[j] % q, q);
            value = orc_mod_mul(value, p->inverses[i * n + j], q);
        }
        out[i] = value;
    }
}

/* evaluate_mixed_radix_mod  keyswitching.rs:412-423 */
static uint64_t evaluate_mixed_radix_mod(const uint64_t* mixed, const uint64_t* moduli, size_t n, uint64_t target) {
    if (n == 0) return 0;
    uint64_t acc = mixed[n - 1] % target;
    for (size_t k = n - 1; k-- > 0;) {
        acc = (uint64_t)(((u128)acc * (moduli[k] % target)) % target);
        acc = add_mod_u64(acc, mixed[k] % target, target);
    }
    return acc;
}

/* canonical_digits_from_mixed_radix  keyswitching.rs:425-432 */
static void canonical_digits_from_mixed_radix(const uint64_t* mixed, const uint64_t* moduli, size_t n, uint64_t base,
                                              digits_t* out) {
    out->len = 1;
    out->d[0] = 0;
    for (size_t k = n; k-- > 0;) mul_add_small_base(out, moduli[k], mixed[k], base);
    trim_base_digits(out);
}

/* convert_centered_residues_to_target  keyswitching.rs:365-379 */
static uint64_t convert_centered_residues_to_target(const uint64_t* residues, const bconv_t* p, uint64_t target,
                                                    uint64_t* scratch) {
    mixed_radix_digits(residues, p, scratch);
    digits_t canonical;
    canonical_digits_from_mixed_radix(scratch, p->moduli, p->count, p->base, &canonical);
    uint64_t canon_mod_target = evaluate_mixed_radix_mod(scratch, p->moduli, p->count, target);
    if (cmp_base_digits(&canonical, &p->half_modulus_digits) > 0) {
        uint64_t m = base_digits_mod_modulus(&p->modulus_digits, p->base, target);
        return sub_mod_u64(canon_mod_target, m, target);
    }
    return canon_mod_target;
}

/* normalize_balanced_digits  keyswitching.rs:434-454 (final carry dropped, as in release builds) */
static void normalize_balanced_digits(const int64_t* digits, size_t ndig, size_t levels, int64_t base, int64_t* out) {
    int64_t half = base >> 1;
    int64_t carry = 0;
    for (size_t i = 0; i < levels; i++) {
        int64_t di = i < ndig ? digits[i] : 0;
        int64_t value = di + carry;
        int64_t digit = value % base; /* C '%' truncates like Rust's */
        if (digit >= half)
            digit -= base;
        else if (digit < -half)
            digit += base;
        carry = (value - digit) / base;
        out[i] = digit;
    }
}

/* decompose_residues_balanced  keyswitching.rs:381-395 */
static void decompose_residues_balanced(const uint64_t* residues, const bconv_t* p, size_t levels, uint64_t* scratch,
                                        int64_t* out) {
    mixed_radix_digits(residues, p, scratch);
    digits_t canonical;
    canonical_digits_from_mixed_radix(scratch, p->moduli, p->count, p->base, &canonical);
    int64_t signed_digits[ORC_MAXD];
    size_t nd;
    if (cmp_base_digits(&canonical, &p->half_modulus_digits) > 0) {
        digits_t mag;
        sub_base_digits(&p->modulus_digits, &canonical, p->base, &mag);
        nd = mag.len;
        for (size_t i = 0; i < nd; i++) signed_digits[i] = -(int64_t)mag.d[i];
    } else {
        nd = canonical.len;
        for (size_t i = 0; i < nd; i++) signed_digits[i] = (int64_t)canonical.d[i];
    }
    /* digits.resize(levels + 1, 0) truncates or pads; only the first `levels` enter the loop */
    normalize_balanced_digits(signed_digits, nd < levels + 1 ? nd : levels + 1, levels, (int64_t)p->base, out);
}

/* signed_digit_mod_q  keyswitching.rs:456-467 */
static uint64_t signed_digit_mod_q(int64_t d, uint64_t q) {
    if (d >= 0) return (uint64_t)d % q;
    uint64_t mag = ((uint64_t)(-d)) % q;
    return mag == 0 ? 0 : q - mag;
}

uint64_t orc_total_modulus_bits(const uint64_t* moduli, size_t levels) {
    digits_t v;
    v.len = 1;
    v.d[0] = 1;
    for (size_t i = 0; i < levels; i++) mul_add_small_base(&v, moduli[i], 0, (uint64_t)1 << 32);
    trim_base_digits(&v);
    uint64_t top = v.d[v.len - 1];
    uint64_t bits = 0;
    while (top) {
        bits++;
        top >>= 1;
    }
    return (uint64_t)(v.len - 1) * 32 + bits;
}

size_t orc_num_decomp_digits(uint64_t total_bits, size_t decomp_bits) {
    return ((size_t)total_bits + decomp_bits - 1) / decomp_bits;
}

/* decompose_poly_rns_with_levels  keyswitching.rs:196-236 */
void orc_decompose_signed(const orc_basis* b, const uint64_t* poly, size_t decomp_bits, size_t levels, int64_t* out) {
    uint64_t base = (uint64_t)1 << decomp_bits;
    bconv_t p;
    bconv_new(&p, b->moduli, b->levels, base);
    size_t n = b->n, L = b->levels;
#pragma omp parallel
    {
        uint64_t* res = (uint64_t*)malloc(L * sizeof(uint64_t));
        uint64_t* scratch = (uint64_t*)malloc(L * sizeof(uint64_t));
        int64_t* dig = (int64_t*)malloc((levels + 1) * sizeof(int64_t));
#pragma omp for schedule(static)
        for (size_t c = 0; c < n; c++) {
            for (size_t l = 0; l < L; l++) res[l] = poly[l * n + c];
            decompose_residues_balanced(res, &p, levels, scratch, dig);
            for (size_t i = 0; i < levels; i++) out[i * n + c] = dig[i];
        }
        free(res);
        free(scratch);
        free(dig);
    }
    bconv_free(&p);
}

void orc_decompose(const orc_basis* b, const uint64_t* poly, size_t decomp_bits, size_t levels, uint64_t* out) {
    size_t n = b->n, L = b->levels;
    int64_t* sd = (int64_t*)malloc(levels * n * sizeof(int64_t));
    orc_decompose_signed(b, poly, decomp_bits, levels, sd);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < levels; i++)
        for (size_t l = 0; l < L; l++)
            for (size_t c = 0; c < n; c++) out[(i * L + l) * n + c] = signed_digit_mod_q(sd[i * n + c], b->moduli[l]);
    free(sd);
}

/* lift_poly_to_context  keyswitching.rs:242-276 */
void orc_lift(const orc_basis* s, const orc_basis* t, const uint64_t* poly, uint64_t* out) {
    size_t n = s->n, Ls = s->levels, Lt = t->levels;
    memcpy(out, poly, Ls * n * sizeof(uint64_t));
    if (Lt == Ls) return;
    bconv_t p;
    bconv_new(&p, s->moduli, Ls, (uint64_t)1 << 30);
#pragma omp parallel
    {
        uint64_t* res = (uint64_t*)malloc(Ls * sizeof(uint64_t));
        uint64_t* scratch = (uint64_t*)malloc(Ls * sizeof(uint64_t));
#pragma omp for schedule(static)
        for (size_t c = 0; c < n; c++) {
            for (size_t l = 0; l < Ls; l++) res[l] = poly[l * n + c];
            for (size_t k = Ls; k < Lt; k++)
                out[k * n + c] = convert_centered_residues_to_target(res, &p, t->moduli[k], scratch) % t->moduli[k];
        }
        free(res);
        free(scratch);
    }
    bconv_free(&p);
}

/* mod_down_eval_to_q  keyswitching.rs:278-322 */
void orc_mod_down(const orc_basis* q, const orc_basis* e, const uint64_t* poly, uint64_t* out) {
    size_t n = q->n, Lq = q->levels, K = e->levels - q->levels;
    const uint64_t* aux = e->moduli + Lq;
    bconv_t p;
    bconv_new(&p, aux, K, (uint64_t)1 << 30);
    for (size_t i = 0; i < Lq; i++) {
        uint64_t qi = q->moduli[i];
        /* P mod q_i (the reference reduces the BigUint P) */
        uint64_t p_mod_q = 1 % qi;
        for (size_t k = 0; k < K; k++) p_mod_q = orc_mod_mul(p_mod_q, aux[k] % qi, qi);
        int ok;
        uint64_t p_inv = orc_mod_inv(p_mod_q, qi, &ok);
        if (!ok) {
            fprintf(stderr, "P must be invertible modulo each q_i\n");
            abort();
        }
#pragma omp parallel
        {
            uint64_t res[64];
            uint64_t scratch[64];
#pragma omp for schedule(static)
            for (size_t c = 0; c < n; c++) {
                for (size_t k = 0; k < K; k++) res[k] = poly[(Lq + k) * n + c];
                uint64_t cp = convert_centered_residues_to_target(res, &p, qi, scratch);
                uint64_t num = sub_mod_u64(poly[i * n + c], cp, qi);
                out[i * n + c] = (uint64_t)(((u128)num * p_inv) % qi);
            }
        }
    }
    bconv_free(&p);
}

/* ------------------------------------------------------------------ */
/* RnsPolynomial ops  src/rns/polynomial.rs:99-145                     */
/* ------------------------------------------------------------------ */
void orc_rns_add(const orc_basis* b, const uint64_t* x, const uint64_t* y, uint64_t* out) {
#pragma omp parallel for schedule(static)
    for (size_t l = 0; l < b->levels; l++)
        orc_poly_add(x + l * b->n, y + l * b->n, out + l * b->n, b->n, b->moduli[l]);
}

void orc_rns_sub(const orc_basis* b, const uint64_t* x, const uint64_t* y, uint64_t* out) {
#pragma omp parallel for schedule(static)
    for (size_t l = 0; l < b->levels; l++)
        orc_poly_sub(x + l * b->n, y + l * b->n, out + l * b->n, b->n, b->moduli[l]);
}

void orc_rns_multiply_ntt(const orc_basis* b, const uint64_t* x, const uint64_t* y, uint64_t* out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t l = 0; l < b->levels; l++)
        orc_poly_multiply_ntt(x + l * b->n, y + l * b->n, out + l * b->n, &b->plans[l]);
}

void orc_rns_galois(const orc_basis* b, const uint64_t* x, uint64_t* out, uint64_t galois) {
#pragma omp parallel for schedule(static)
    for (size_t l = 0; l < b->levels; l++) orc_poly_galois(x + l * b->n, out + l * b->n, b->n, b->moduli[l], galois);
}

/* ------------------------------------------------------------------ */
/* scheme ops  src/rns/ops.rs                                          */
/* ------------------------------------------------------------------ */
void orc_multiply_ciphertexts(const orc_basis* b, const uint64_t* lhs, const uint64_t* rhs, uint64_t* out) {
    size_t P = b->levels * b->n;
    const uint64_t *l0 = lhs, *l1 = lhs + P, *r0 = rhs, *r1 = rhs + P;
    uint64_t* t = (uint64_t*)malloc(2 * P * sizeof(uint64_t));
    orc_rns_multiply_ntt(b, l0, r0, out);             /* c0 = l0*r0 (:92) */
    orc_rns_multiply_ntt(b, l0, r1, t);               /* l0*r1 (:93-95) */
    orc_rns_multiply_ntt(b, l1, r0, t + P);           /* l1*r0 (:96) */
    orc_rns_add(b, t, t + P, out + P);                /* c1 (:93-96) */
    orc_rns_multiply_ntt(b, l1, r1, out + 2 * P);     /* c2 (:97) */
    free(t);
}

void orc_multiply_plain(const orc_basis* b, const uint64_t* ct, const uint64_t* plain, uint64_t* out) {
    size_t P = b->levels * b->n;
    orc_rns_multiply_ntt(b, ct, plain, out);
    orc_rns_multiply_ntt(b, ct + P, plain, out + P);
}

/* centered_residue_mod_prime  ops.rs:175-186 */
static uint64_t centered_residue_mod_prime(uint64_t v, uint64_t modulus, uint64_t half, uint64_t target) {
    if (v <= half) return v % target;
    uint64_t dist = (modulus - v) % target;
    return dist == 0 ? 0 : target - dist;
}

/* rescale_poly_rns  ops.rs:135-173, for `polys` consecutive polynomials */
void orc_rescale(const orc_basis* b, const uint64_t* in, size_t polys, uint64_t* out) {
    size_t n = b->n, L = b->levels;
    uint64_t dropped = b->moduli[L - 1];
    uint64_t half = dropped >> 1;
    for (size_t pi = 0; pi < polys; pi++) {
        const uint64_t* src = in + pi * L * n;
        const uint64_t* dc = src + (L - 1) * n;
        uint64_t* dst = out + pi * (L - 1) * n;
#pragma omp parallel for schedule(static)
        for (size_t l = 0; l < L - 1; l++) {
            uint64_t qi = b->moduli[l];
            int ok;
            uint64_t inv = orc_mod_inv(dropped % qi, qi, &ok);
            if (!ok) abort();
            for (size_t c = 0; c < n; c++) {
                uint64_t cen = centered_residue_mod_prime(dc[c], dropped, half, qi);
                uint64_t x = src[l * n + c];
                uint64_t num = x >= cen ? x - cen : qi - (cen - x); /* ops.rs:188-193 */
                dst[l * n + c] = (uint64_t)(((u128)num * inv) % qi);
            }
        }
    }
}

void orc_decrypt(const orc_basis* b, const uint64_t* ct, const uint64_t* secret, uint64_t* out) {
    size_t P = b->levels * b->n;
    uint64_t* t = (uint64_t*)malloc(P * sizeof(uint64_t));
    orc_rns_multiply_ntt(b, ct + P, secret, t);
    orc_rns_add(b, ct, t, out);
    free(t);
}

void orc_decrypt_quadratic(const orc_basis* b, const uint64_t* ct, const uint64_t* secret, uint64_t* out) {
    size_t P = b->levels * b->n;
    uint64_t* s2 = (uint64_t*)malloc(P * sizeof(uint64_t));
    uint64_t* t1 = (uint64_t*)malloc(P * sizeof(uint64_t));
    uint64_t* t2 = (uint64_t*)malloc(P * sizeof(uint64_t));
    orc_rns_multiply_ntt(b, secret, secret, s2);
    orc_rns_multiply_ntt(b, ct + P, secret, t1);
    orc_rns_multiply_ntt(b, ct + 2 * P, s2, t2);
    orc_rns_add(b, ct, t1, out);
    orc_rns_add(b, out, t2, out);
    free(s2);
    free(t1);
    free(t2);
}

void orc_encrypt_with_samples(const orc_basis* b, const uint64_t* plain, const uint64_t* pk, const uint64_t* u,
                              const uint64_t* e0, const uint64_t* e1, uint64_t* out) {
    size_t P = b->levels * b->n;
    uint64_t* t = (uint64_t*)malloc(P * sizeof(uint64_t));
    orc_rns_multiply_ntt(b, pk, u, t);
    orc_rns_add(b, t, e0, out);
    orc_rns_add(b, out, plain, out);
    orc_rns_multiply_ntt(b, pk + P, u, t);
    orc_rns_add(b, t, e1, out + P);
    free(t);
}

/* ------------------------------------------------------------------ */
/* keyswitch  keyswitching.rs:130-194                                  */
/* ------------------------------------------------------------------ */
int orc_keyswitch_qonly(const orc_basis* q, const uint64_t* poly, const uint64_t* ksk, size_t ksk_digits,
                        uint64_t* out) {
    size_t n = q->n, L = q->levels, P = L * n;
    size_t active = orc_num_decomp_digits(orc_total_modulus_bits(q->moduli, L), 30);
    if (ksk_digits < active) return -1; /* keyswitch key must cover the active modulus chain */
    uint64_t* digits = (uint64_t*)malloc(ksk_digits * P * sizeof(uint64_t));
    orc_decompose(q, poly, 30, ksk_digits, digits); /* levels = ksk.ksk.len() (:138) */
    uint64_t* t = (uint64_t*)malloc(P * sizeof(uint64_t));
    memset(out, 0, 2 * P * sizeof(uint64_t));
    for (size_t i = 0; i < ksk_digits; i++) {
        const uint64_t* k0 = ksk + (i * 2 + 0) * P;
        const uint64_t* k1 = ksk + (i * 2 + 1) * P;
        orc_rns_multiply_ntt(q, digits + i * P, k0, t);
        orc_rns_add(q, out, t, out);
        orc_rns_multiply_ntt(q, digits + i * P, k1, t);
        orc_rns_add(q, out + P, t, out + P);
    }
    free(t);
    free(digits);
    return 0;
}

int orc_keyswitch_hybrid(const orc_basis* q, const orc_basis* e, const uint64_t* poly, const uint64_t* ksk,
                         size_t ksk_digits, uint64_t* out) {
    size_t n = q->n, L = q->levels, Le = e->levels, P = L * n, Pe = Le * n;
    size_t active = orc_num_decomp_digits(orc_total_modulus_bits(q->moduli, L), 30);
    if (ksk_digits < active) return -1;
    uint64_t* digits = (uint64_t*)malloc(active * P * sizeof(uint64_t));
    orc_decompose(q, poly, 30, active, digits); /* levels = active digits (:170) */
    uint64_t* lifted = (uint64_t*)malloc(Pe * sizeof(uint64_t));
    uint64_t* t = (uint64_t*)malloc(Pe * sizeof(uint64_t));
    uint64_t* acc0 = (uint64_t*)calloc(Pe, sizeof(uint64_t));
    uint64_t* acc1 = (uint64_t*)calloc(Pe, sizeof(uint64_t));
    for (size_t i = 0; i < active; i++) {
        const uint64_t* k0 = ksk + (i * 2 + 0) * Pe;
        const uint64_t* k1 = ksk + (i * 2 + 1) * Pe;
        orc_lift(q, e, digits + i * P, lifted);
        orc_rns_multiply_ntt(e, lifted, k0, t);
        orc_rns_add(e, acc0, t, acc0);
        orc_rns_multiply_ntt(e, lifted, k1, t);
        orc_rns_add(e, acc1, t, acc1);
    }
    orc_mod_down(q, e, acc0, out);
    orc_mod_down(q, e, acc1, out + P);
    free(digits);
    free(lifted);
    free(t);
    free(acc0);
    free(acc1);
    return 0;
}

/* relinearize  context/ops.rs:111-122 (keyswitch dispatch :97-109) */
int orc_relinearize(const orc_basis* q, const orc_basis* e, const uint64_t* quad, const uint64_t* ksk,
                    size_t ksk_digits, uint64_t* out) {
    size_t P = q->levels * q->n;
    uint64_t* ks = (uint64_t*)malloc(2 * P * sizeof(uint64_t));
    int rc = e ? orc_keyswitch_hybrid(q, e, quad + 2 * P, ksk, ksk_digits, ks)
               : orc_keyswitch_qonly(q, quad + 2 * P, ksk, ksk_digits, ks);
    if (rc == 0) {
        orc_rns_add(q, quad, ks, out);
        orc_rns_add(q, quad + P, ks + P, out + P);
    }
    free(ks);
    return rc;
}

/* rotate  context/rotation.rs:104-127 */
int orc_rotate(const orc_basis* q, const orc_basis* e, const uint64_t* ct, uint64_t galois, const uint64_t* ksk,
               size_t ksk_digits, uint64_t* out) {
    size_t P = q->levels * q->n;
    uint64_t* rot = (uint64_t*)malloc(2 * P * sizeof(uint64_t));
    uint64_t* ks = (uint64_t*)malloc(2 * P * sizeof(uint64_t));
    orc_rns_galois(q, ct, rot, galois);
    orc_rns_galois(q, ct + P, rot + P, galois);
    int rc = e ? orc_keyswitch_hybrid(q, e, rot + P, ksk, ksk_digits, ks)
               : orc_keyswitch_qonly(q, rot + P, ksk, ksk_digits, ks);
    if (rc == 0) {
        orc_rns_add(q, rot, ks, out);
        memcpy(out + P, ks + P, P * sizeof(uint64_t));
    }
    free(rot);
    free(ks);
    return rc;
}

/* galois_element_for_rotation  rotation.rs:14-26 */
static uint64_t pow_mod_small(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    b %= m;
    while (e) {
        if (e & 1) r = (uint64_t)(((u128)r * b) % m);
        b = (uint64_t)(((u128)b * b) % m);
        e >>= 1;
    }
    return r;
}

uint64_t orc_galois_element_for_rotation(uint64_t poly_degree, int64_t steps) {
    uint64_t m = 2 * poly_degree;
    uint64_t g = 5 % m;
    if (steps >= 0) return pow_mod_small(g, (uint64_t)steps, m);
    int ok;
    uint64_t ginv = orc_mod_inv(g, m, &ok);
    return pow_mod_small(ginv, (uint64_t)(-steps), m);
}

/* ------------------------------------------------------------------ */
/* seeded sampling                                                      */
/* ------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void orc_rng_seed(orc_rng* r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
    r->have_spare = 0;
    r->spare = 0;
}

static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* xoshiro256** */
uint64_t orc_rng_next(orc_rng* r) {
    uint64_t* s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

static double rng_unit(orc_rng* r) { return (double)(orc_rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }

/* Normal(0, sigma) by Box-Muller, rounded to nearest (keys.rs:83-86) */
static int64_t sample_gaussian(orc_rng* r, double sigma) {
    double z;
    if (r->have_spare) {
        r->have_spare = 0;
        z = r->spare;
    } else {
        double u1 = rng_unit(r), u2 = rng_unit(r);
        if (u1 < 1e-300) u1 = 1e-300;
        double rad = sqrt(-2.0 * log(u1));
        z = rad * cos(2.0 * M_PI * u2);
        r->spare = rad * sin(2.0 * M_PI * u2);
        r->have_spare = 1;
    }
    return (int64_t)llround(z * sigma);
}

void orc_sample_uniform(const orc_basis* b, orc_rng* r, uint64_t* out) {
    for (size_t l = 0; l < b->levels; l++)
        for (size_t c = 0; c < b->n; c++) out[l * b->n + c] = orc_rng_next(r) % b->moduli[l];
}

static void lift_small(const orc_basis* b, const int64_t* v, uint64_t* out) {
    for (size_t l = 0; l < b->levels; l++) {
        uint64_t q = b->moduli[l];
        for (size_t c = 0; c < b->n; c++)
            out[l * b->n + c] = v[c] >= 0 ? (uint64_t)v[c] % q : (uint64_t)((i128)q + v[c]) % q;
    }
}

void orc_sample_error(const orc_basis* b, orc_rng* r, double sigma, uint64_t* out) {
    int64_t* v = (int64_t*)malloc(b->n * sizeof(int64_t));
    for (size_t c = 0; c < b->n; c++) v[c] = sample_gaussian(r, sigma);
    lift_small(b, v, out);
    free(v);
}

void orc_sample_ternary(const orc_basis* b, orc_rng* r, uint64_t* out) {
    int64_t* v = (int64_t*)malloc(b->n * sizeof(int64_t));
    for (size_t c = 0; c < b->n; c++) {
        uint64_t k = orc_rng_next(r) % 3;
        v[c] = k == 0 ? 0 : (k == 1 ? 1 : -1);
    }
    lift_small(b, v, out);
    free(v);
}

void orc_sample_sparse_ternary(const orc_basis* b, orc_rng* r, size_t hamming, uint64_t* out) {
    int64_t* v = (int64_t*)calloc(b->n, sizeof(int64_t));
    size_t placed = 0;
    while (placed < hamming) {
        size_t idx = (size_t)(orc_rng_next(r) % b->n);
        if (v[idx] == 0) {
            v[idx] = (orc_rng_next(r) & 1) ? 1 : -1;
            placed++;
        }
    }
    lift_small(b, v, out);
    free(v);
}

/* rns_keygen  keys.rs:153-167 */
void orc_keygen_public(const orc_basis* b, orc_rng* r, double sigma, const uint64_t* secret, uint64_t* pk,
                       uint64_t* error) {
    size_t P = b->levels * b->n;
    uint64_t* p1 = pk + P;
    orc_sample_uniform(b, r, p1);
    orc_sample_error(b, r, sigma, error);
    uint64_t* p1s = (uint64_t*)malloc(P * sizeof(uint64_t));
    orc_rns_multiply_ntt(b, p1, secret, p1s);
    orc_rns_sub(b, error, p1s, pk);
    free(p1s);
}

/* generate_keyswitch_key_rns_hybrid  keyswitching.rs:81-128.
 * plaintext_i = centred(source) * 2^{30 i} * P reduced into every eval prime; computed per prime:
 * (centred(source) mod p_j) * (2^{30 i} mod p_j) * (P mod p_j) mod p_j, which equals the reference's
 * BigInt product reduced mod QP then mod p_j (p_j divides QP). */
void orc_keygen_hybrid(const orc_basis* q, const orc_basis* e, orc_rng* r, double sigma, const uint64_t* source,
                       const uint64_t* secret, size_t digits, int noiseless, uint64_t* out) {
    size_t n = q->n, L = q->levels, Le = e->levels, Pe = Le * n;
    uint64_t* cen = (uint64_t*)malloc(Pe * sizeof(uint64_t)); /* centred source mod every eval prime */
    orc_lift(q, e, source, cen);
    uint64_t* tsec = (uint64_t*)malloc(Pe * sizeof(uint64_t));
    orc_lift(q, e, secret, tsec); /* target_secret (:97); secret's first L limbs */
    uint64_t* pt = (uint64_t*)malloc(Pe * sizeof(uint64_t));
    uint64_t* a1s = (uint64_t*)malloc(Pe * sizeof(uint64_t));
    uint64_t* err = (uint64_t*)malloc(Pe * sizeof(uint64_t));
    for (size_t i = 0; i < digits; i++) {
        for (size_t l = 0; l < Le; l++) {
            uint64_t pj = e->moduli[l];
            uint64_t bp = orc_mod_pow(2 % pj, 30 * (uint64_t)i, pj);
            uint64_t pm = 1 % pj;
            for (size_t k = L; k < Le; k++) pm = orc_mod_mul(pm, e->moduli[k] % pj, pj);
            uint64_t f = orc_mod_mul(bp, pm, pj);
            for (size_t c = 0; c < n; c++) pt[l * n + c] = orc_mod_mul(cen[l * n + c], f, pj);
        }
        uint64_t* a0 = out + (i * 2 + 0) * Pe;
        uint64_t* a1 = out + (i * 2 + 1) * Pe;
        if (noiseless) {
            memcpy(a0, pt, Pe * sizeof(uint64_t));
            memset(a1, 0, Pe * sizeof(uint64_t));
            continue;
        }
        orc_sample_uniform(e, r, a1);
        orc_sample_error(e, r, sigma, err);
        orc_rns_multiply_ntt(e, a1, tsec, a1s);
        orc_rns_add(e, pt, err, a0);
        orc_rns_sub(e, a0, a1s, a0);
    }
    free(cen);
    free(tsec);
    free(pt);
    free(a1s);
    free(err);
}

/* generate_keyswitch_key_rns  keyswitching.rs:44-79 (plaintext_i = source * (2^{30 i} mod Q) mod Q, encrypted
 * with encrypt_rns at scale 0, ops.rs:8-23) */
void orc_keygen_qonly(const orc_basis* q, orc_rng* r, double sigma, const uint64_t* source, const uint64_t* pk,
                      size_t digits, int noiseless, uint64_t* out) {
    size_t n = q->n, L = q->levels, P = L * n;
    uint64_t* pt = (uint64_t*)malloc(P * sizeof(uint64_t));
    uint64_t* u = (uint64_t*)malloc(P * sizeof(uint64_t));
    uint64_t* e0 = (uint64_t*)malloc(P * sizeof(uint64_t));
    uint64_t* e1 = (uint64_t*)malloc(P * sizeof(uint64_t));
    for (size_t i = 0; i < digits; i++) {
        for (size_t l = 0; l < L; l++) {
            uint64_t ql = q->moduli[l];
            uint64_t bp = orc_mod_pow(2 % ql, 30 * (uint64_t)i, ql);
            for (size_t c = 0; c < n; c++) pt[l * n + c] = orc_mod_mul(source[l * n + c], bp, ql);
        }
        uint64_t* ct = out + i * 2 * P;
        if (noiseless) {
            memcpy(ct, pt, P * sizeof(uint64_t));
            memset(ct + P, 0, P * sizeof(uint64_t));
            continue;
        }
        orc_sample_ternary(q, r, u);
        orc_sample_error(q, r, sigma, e0);
        orc_sample_error(q, r, sigma, e1);
        orc_encrypt_with_samples(q, pt, pk, u, e0, e1, ct);
    }
    free(pt);
    free(u);
    free(e0);
    free(e1);
}
