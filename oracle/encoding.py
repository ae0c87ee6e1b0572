"""CKKS encode / decode and CRT reconstruction (TEST INFRASTRUCTURE ONLY).

Restates src/rns/encoding.rs:9-87, src/math/canonical_embedding.rs:21-79 and
RnsPolynomial::reconstruct_coefficients / from_bigint_coeffs (src/rns/polynomial.rs:57-85,
162-201, 221-235) with numpy FFTs and Python integers.  Encoding is used only to build test
plaintexts and decoding only to measure CKKS error against FP64; neither is on the GPU path.
Bit-exact parity of encode with the reference is unpinned (it depends on num-complex 0.4's
from_polar/powi rounding); decode is checked against FP64 within the reference's tolerances.
"""
import numpy as np


def slot_exponents(n):
    """slot_exponents (encoding.rs:68-79): 5^i mod 2N for i < N/2."""
    out, cur = [], 1
    for _ in range(n // 2):
        out.append(cur)
        cur = cur * 5 % (2 * n)
    return out


def embedding_forward(coeffs):
    """canonical_embedding_forward (canonical_embedding.rs:21-28): y_k = sum_j a_j w^{jk}, w = e^{2 pi i/n}."""
    a = np.asarray(coeffs, dtype=np.complex128)
    return np.fft.ifft(a) * a.size


def embedding_backward(values):
    """canonical_embedding_backward (canonical_embedding.rs:35-42): a_j = (1/n) sum_k y_k w^{-jk}."""
    y = np.asarray(values, dtype=np.complex128)
    return np.fft.fft(y) / y.size


def encode_integer_coeffs(n, scale, slots):
    """encode_integer_coeffs (encoding.rs:9-31): returns Python ints (rounded, scaled, twisted)."""
    slots = np.asarray(slots, dtype=np.complex128)
    assert slots.size == n // 2, "slot count must be N/2"
    evals = np.zeros(n, dtype=np.complex128)
    for v, e in zip(slots, slot_exponents(n)):
        evals[(e - 1) // 2] = v
        evals[((2 * n - e) % (2 * n) - 1) // 2] = np.conj(v)
    coeffs = embedding_backward(evals)
    j = np.arange(n)
    twisted = coeffs * np.exp(-1j * np.pi * j / n)
    return [int(x) for x in np.round(twisted.real * scale)]


def decode_integer_coeffs(n, scale, centered):
    """decode_integer_coeffs (encoding.rs:33-66) from centred integer coefficients."""
    c = np.array([float(x) for x in centered]) / scale
    j = np.arange(n)
    b = c * np.exp(1j * np.pi * j / n)
    evals = embedding_forward(b)
    return np.array([evals[(e - 1) // 2] for e in slot_exponents(n)])


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
