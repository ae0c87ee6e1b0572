"""CKKS encode / decode and CRT reconstruction (TEST INFRASTRUCTURE ONLY).

Restates src/rns/encoding.rs:9-87, src/math/canonical_embedding.rs:21-79 and
RnsPolynomial::reconstruct_coefficients / from_bigint_coeffs (src/rns/polynomial.rs:57-85,
162-201, 221-235) with Python integers and the crate's own FFT: bit reversal, radix-2 stages whose
twiddles follow the recurrence w *= w_len from w_len = from_polar(1, +-2 pi / len), complex products
(ac - bd, ad + bc) evaluated one rounded operation at a time (numpy arrays of real and imaginary
parts; no fused multiply-add), the 1/n scaling, the xi^-j twist by num-traits square-and-multiply on
the inverse of xi, and Rust's round-half-away-from-zero.  A constant slot vector therefore encodes
exactly to round(c 2^s) in coefficient 0 and zeros elsewhere, as in the crate.  Bit-exact parity of
encode with the crate rests on libm's cos/sin and num-complex 0.4's operation order (unpinned by any
reference vector); decode is checked against FP64 within the reference's tolerances.
"""
import math

import numpy as np


def slot_exponents(n):
    """slot_exponents (encoding.rs:68-79): 5^i mod 2N for i < N/2."""
    out, cur = [], 1
    for _ in range(n // 2):
        out.append(cur)
        cur = cur * 5 % (2 * n)
    return out


def _bit_reverse_indices(n):
    """bit_reverse_permute (canonical_embedding.rs:81-95) as a permutation: out[i] = in[perm[i]]."""
    bits = n.bit_length() - 1
    idx = np.arange(n)
    rev = np.zeros(n, dtype=np.int64)
    for b in range(bits):
        rev |= ((idx >> b) & 1) << (bits - 1 - b)
    return rev


def _cmul(ar, ai, br, bi):
    """num-complex product (ac - bd, ad + bc), every operation rounded on its own."""
    return ar * br - ai * bi, ar * bi + ai * br


def fft_crate(re, im, inverse):
    """fft_in_place (canonical_embedding.rs:44-79) on arrays of real and imaginary parts."""
    n = re.size
    assert n & (n - 1) == 0, "FFT size must be a power of two"
    perm = _bit_reverse_indices(n)
    re, im = np.array(re, dtype=np.float64)[perm], np.array(im, dtype=np.float64)[perm]
    length = 2
    while length <= n:
        angle = (-2.0 * math.pi / length) if inverse else (2.0 * math.pi / length)
        lr, li = 1.0 * math.cos(angle), 1.0 * math.sin(angle)   # from_polar(1, angle)
        half = length // 2
        wr, wi = np.empty(half), np.empty(half)
        cr, ci = 1.0, 0.0
        for i in range(half):                                   # w *= w_len, sequentially
            wr[i], wi[i] = cr, ci
            cr, ci = cr * lr - ci * li, cr * li + ci * lr
        r, m = re.reshape(-1, length), im.reshape(-1, length)
        ur, ui = r[:, :half].copy(), m[:, :half].copy()
        tr, ti = _cmul(r[:, half:], m[:, half:], wr, wi)
        r[:, :half], m[:, :half] = ur + tr, ui + ti
        r[:, half:], m[:, half:] = ur - tr, ui - ti
        length *= 2
    if inverse:
        sc = 1.0 / n
        re, im = re * sc, im * sc
    return re, im


def _powi(br, bi, exps):
    """num-complex powi(base, e) for e >= 0 by num-traits' square-and-multiply, for many e at once: the
    squarings base^(2^k) are one shared chain; each result starts at its lowest set bit's square and
    multiplies the higher set bits' squares in increasing order."""
    exps = np.asarray(exps, dtype=np.int64)
    top = int(exps.max()).bit_length() if exps.size else 0
    sq = [(br, bi)]
    for _ in range(1, max(top, 1)):
        r, i = sq[-1]
        sq.append((r * r - i * i, r * i + i * r))
    rr, ri = np.ones(exps.size), np.zeros(exps.size)
    started = np.zeros(exps.size, dtype=bool)
    for k in range(top):
        bit = ((exps >> k) & 1).astype(bool)
        first = bit & ~started
        more = bit & started
        rr[first], ri[first] = sq[k][0], sq[k][1]
        if more.any():
            pr, pi_ = _cmul(rr[more], ri[more], sq[k][0], sq[k][1])
            rr[more], ri[more] = pr, pi_
        started |= bit
    return rr, ri


def _round_half_away(x):
    """f64::round: halves go away from zero."""
    t = np.trunc(x)
    frac = x - t
    return t + np.where(np.abs(frac) >= 0.5, np.sign(x), 0.0)


def _xi(n):
    t = math.pi / n                                             # primitive_2n_root (encoding.rs:85-87)
    return 1.0 * math.cos(t), 1.0 * math.sin(t)


def embedding_forward(coeffs):
    """canonical_embedding_forward (canonical_embedding.rs:21-28): y_k = sum_j a_j w^{jk}, w = e^{2 pi i/n}."""
    a = np.asarray(coeffs, dtype=np.complex128)
    r, i = fft_crate(a.real, a.imag, False)
    return r + 1j * i


def embedding_backward(values):
    """canonical_embedding_backward (canonical_embedding.rs:35-42): a_j = (1/n) sum_k y_k w^{-jk}."""
    y = np.asarray(values, dtype=np.complex128)
    r, i = fft_crate(y.real, y.imag, True)
    return r + 1j * i


def encode_integer_coeffs(n, scale, slots):
    """encode_integer_coeffs (encoding.rs:9-31): returns Python ints (rounded, scaled, twisted)."""
    slots = np.asarray(slots, dtype=np.complex128)
    assert slots.size == n // 2, "slot count must be N/2"
    er, ei = np.zeros(n), np.zeros(n)
    for v, e in zip(slots, slot_exponents(n)):
        er[(e - 1) // 2], ei[(e - 1) // 2] = v.real, v.imag
        j = ((2 * n - e) % (2 * n) - 1) // 2
        er[j], ei[j] = v.real, -v.imag
    cr, ci = fft_crate(er, ei, True)
    xr, xi = _xi(n)
    ns = xr * xr + xi * xi                                      # Complex::inv: (re / |z|^2, -im / |z|^2)
    pr, pi_ = _powi(xr / ns, -xi / ns, np.arange(n))
    tr, _ = _cmul(cr, ci, pr, pi_)
    return [int(x) for x in _round_half_away(tr * scale)]       # re.mul_add(scale, 0.0).round()


def decode_integer_coeffs(n, scale, centered):
    """decode_integer_coeffs (encoding.rs:33-66) from centred integer coefficients."""
    inv_delta = 1.0 / scale
    c = np.array([float(x) for x in centered]) * inv_delta
    xr, xi = _xi(n)
    pr, pi_ = _powi(xr, xi, np.arange(n))
    br, bi = _cmul(c, np.zeros(n), pr, pi_)
    er, ei = fft_crate(br, bi, False)
    idx = [(e - 1) // 2 for e in slot_exponents(n)]
    return er[idx] + 1j * ei[idx]


def total_modulus(moduli):
    q = 1
    for m in moduli:
        q *= int(m)
    return q


def poly_from_bigint_coeffs(moduli, n, coeffs):
    """RnsPolynomial::from_bigint_coeffs (polynomial.rs:57-85): centred ints -> residues per prime."""
    coeffs = list(coeffs) + [0] * (n - len(coeffs))
    out = np.zeros((len(moduli), n), dtype=np.uint64)
    for l, q in enumerate(moduli):
        q = int(q)
        out[l] = np.array([c % q for c in coeffs], dtype=np.uint64)
    return out


def reconstruct(moduli, residues):
    """reconstruct_coefficients (polynomial.rs:162-201): CRT to ints in [0, Q)."""
    moduli = [int(m) for m in moduli]
    Q = total_modulus(moduli)
    acc = np.zeros(residues.shape[-1], dtype=object)
    for l, q in enumerate(moduli):
        part = Q // q
        inv = pow(part % q, -1, q)
        acc = acc + residues[l].astype(object) * ((part * inv) % Q)
    return [int(x) % Q for x in acc]


def centered(values, modulus):
    """mod_q_biguint_to_centered (polynomial.rs:227-235): v <= Q/2 stays, else v - Q."""
    half = modulus >> 1
    return [v if v <= half else v - modulus for v in values]


def encode(moduli, n, scale_bits, slots):
    return poly_from_bigint_coeffs(moduli, n, encode_integer_coeffs(n, 2.0 ** scale_bits, slots))


def decode(moduli, n, scale_bits, residues):
    Q = total_modulus(moduli)
    return decode_integer_coeffs(n, 2.0 ** scale_bits, centered(reconstruct(moduli, residues), Q))
