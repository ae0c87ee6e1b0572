"""Paterson-Stockmeyer evaluation restated over the oracle (TEST INFRASTRUCTURE ONLY).

Follows src/algorithms/paterson_stockmeyer.rs:18-336 step for step with the oracle's ring
operations: constants are encoded with the general encoder (oracle/encoding.py, the restatement of
encode_at_level_and_scale) and multiplied with multiply_plain (rns/ops.rs:54-66), ct x ct products
are mult -> relinearize -> rescale (context/ops.rs:124-133).  A ciphertext is (array[2][limbs][N],
scale_bits); its level is limbs - 1.

Parity pins: block size / required levels against the crate's unit tests (:353-366); the whole
evaluation against the library bit for bit on the same inputs and keys, and against FP64 polynomial
evaluation after decryption within the crate's tolerance (:368-404).
"""
import numpy as np

from . import encoding as E
from . import oracle as O


class PsError(ValueError):
    pass


def block_size(term_count):
    """paterson_stockmeyer_block_size (:18-26): smallest b with b*b >= term_count."""
    if term_count <= 0:
        raise PsError("term_count must be positive")
    b = 1
    while b * b < term_count:
        b += 1
    return b


def required_levels(ciphertext_level, term_count):
    """paterson_stockmeyer_required_levels (:108-126)."""
    if term_count <= 0:
        raise PsError("term_count must be positive")
    if term_count == 1:
        return []
    b = block_size(term_count)
    blocks = -(-term_count // b)
    top = ciphertext_level - (b - 1)
    if top < 0:
        raise PsError("polynomial degree is too large for the available CKKS level budget")
    final = top - (blocks - 1)
    if final < 0:
        raise PsError("polynomial degree is too large for the available CKKS level budget")
    return list(range(ciphertext_level, final, -1))


class Ring:
    """Q chain + aux chain of one ring degree, with the relinearization keys by level."""

    def __init__(self, n, primes, roots, aux_primes, aux_roots, relin_keys):
        self.n, self.primes, self.roots = n, list(primes), list(roots)
        self.aux, self.aux_roots = list(aux_primes), list(aux_roots)
        self.keys = relin_keys
        self._b = {}

    def q(self, limbs):
        k = ("q", limbs)
        if k not in self._b:
            self._b[k] = O.Basis(self.n, self.primes[:limbs], self.roots[:limbs])
        return self._b[k]

    def e(self, limbs):
        if not self.aux:
            return None
        k = ("e", limbs)
        if k not in self._b:
            self._b[k] = O.Basis(self.n, self.primes[:limbs] + self.aux, self.roots[:limbs] + self.aux_roots)
        return self._b[k]

    def const_plain(self, value, limbs, scale_bits):
        """constant_plaintext (:325-332): encode a constant slot vector at (level, scale)."""
        slots = np.full(self.n // 2, complex(value, 0.0))
        return E.encode(self.primes[:limbs], self.n, scale_bits, slots)


def _limbs(ct):
    return ct[0].shape[1]


def constant_ciphertext(R, level, scale, value):
    z = np.zeros((2, level + 1, R.n), dtype=np.uint64)
    return add_constant_term(R, (z, scale), value)


def add_constant_term(R, ct, value):
    if value == 0.0:
        return ct
    arr, scale = ct
    plain = R.const_plain(value, _limbs(ct), scale)
    b = R.q(_limbs(ct))
    out = arr.copy()
    out[0] = O.rns_add(b, arr[0], plain)
    return out, scale


def align(ct, target_level):
    """align_ciphertext_to_level (:218-225)."""
    if target_level > _limbs(ct) - 1:
        raise PsError("target level cannot exceed ciphertext level")
    return np.ascontiguousarray(ct[0][:, : target_level + 1]), ct[1]


def lift(R, ct, target):
    """lift_ciphertext_scale (:227-251)."""
    arr, scale = ct
    if scale > target:
        raise PsError("cannot lower ciphertext scale with a plaintext-one lift")
    if scale == target:
        return ct
    plain = R.const_plain(1.0, _limbs(ct), target - scale)
    return O.multiply_plain(R.q(_limbs(ct)), arr, plain), target


def mult_plain_const(R, ct, value, plain_scale):
    """mult_plain_slots_real_at_scale (:253-273) with a constant slot vector."""
    arr, scale = ct
    plain = R.const_plain(value, _limbs(ct), plain_scale)
    return O.multiply_plain(R.q(_limbs(ct)), arr, plain), scale + plain_scale


def add(R, a, b):
    if a[1] != b[1]:
        raise PsError("ciphertexts must share the same scale to add")
    if _limbs(a) != _limbs(b):
        raise PsError("ciphertexts must be at the same level to add")
    q = R.q(_limbs(a))
    return np.stack([O.rns_add(q, a[0][0], b[0][0]), O.rns_add(q, a[0][1], b[0][1])]), a[1]


def mult_relinearized(R, a, b):
    """mult (context/ops.rs:63-70) -> relinearize (:111-122) -> rescale (:32-37)."""
    if _limbs(a) != _limbs(b):
        raise PsError("ciphertexts must be at the same level to multiply")
    if a[1] != b[1]:
        raise PsError("ciphertexts must share the same scale to multiply")
    limbs = _limbs(a)
    level = limbs - 1
    if level not in R.keys:
        raise PsError(f"missing Paterson-Stockmeyer relinearization key at level {level}")
    q = R.q(limbs)
    quad = O.multiply_ciphertexts(q, a[0], b[0])
    rel = O.relinearize(q, R.e(limbs), quad, R.keys[level])
    out = O.rescale(q, rel)
    drop = int(R.primes[level]).bit_length() - 1
    return out, a[1] + b[1] - drop


def working_scale_bits(R, level, term_count):
    """paterson_stockmeyer_working_scale_bits (:275-295)."""
    return min(int(R.primes[lv]).bit_length() - 1 for lv in required_levels(level, term_count))


def baby_powers(R, ct, block, scale):
    """precompute_baby_powers (:146-174)."""
    powers = [lift(R, ct, scale)]
    base = powers[0]
    for _ in range(2, block + 1):
        prev = powers[-1]
        nxt = mult_relinearized(R, lift(R, prev, scale), align(base, _limbs(prev) - 1))
        powers.append(lift(R, nxt, scale))
    return powers


def evaluate_block(R, powers, coeffs, bi, block, target, working, coeff_scale):
    """evaluate_block (:176-216)."""
    acc = constant_ciphertext(R, target, working, 0.0)
    for off in range(block):
        idx = bi * block + off
        if idx >= len(coeffs):
            break
        c = coeffs[idx]
        if c == 0.0:
            continue
        if off == 0:
            acc = add_constant_term(R, acc, c)
            continue
        pw = powers[off - 1]
        term = mult_plain_const(R, lift(R, align(pw, target), pw[1]), c, coeff_scale)
        acc = add(R, acc, term)
    return acc


def ps_eval(R, ct, coeffs):
    """paterson_stockmeyer_eval (:28-100).  ct = (array[2][limbs][N], scale_bits)."""
    if len(coeffs) == 0:
        raise PsError("polynomial must contain at least one coefficient")
    level = _limbs(ct) - 1
    block = block_size(len(coeffs))
    blocks = -(-len(coeffs) // block)
    if len(coeffs) == 1:
        return constant_ciphertext(R, level, ct[1], coeffs[0])
    working = working_scale_bits(R, level, len(coeffs))
    if working < ct[1]:
        raise PsError("Paterson-Stockmeyer working scale must dominate the input ciphertext scale")
    coeff_scale = working - ct[1]
    powers = baby_powers(R, ct, block, ct[1])
    giant = powers[-1]
    if _limbs(giant) < blocks:
        raise PsError("polynomial degree is too large for the available CKKS level budget")
    top = _limbs(giant) - 1
    bl = [evaluate_block(R, powers, coeffs, bi, block, top - (blocks - 1 - bi), working, coeff_scale)
          for bi in range(blocks)]
    result = bl[-1]
    for b in reversed(bl[:-1]):
        g = lift(R, align(giant, _limbs(result) - 1), result[1])
        m = mult_relinearized(R, result, g)
        result = add(R, lift(R, m, result[1]), b)
    return result
