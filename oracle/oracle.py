"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a C restatement of the reference crate's CPU path (Akram275/rns-ckks-rust);
see ckks_oracle.h for the per-function citations and how each part is pinned.  Only the
test suite, the smoke check and the benchmark's CPU-baseline leg import this module, and
only as the checker or the timed CPU baseline.

Arrays are numpy uint64 with shape (..., levels, N), C-contiguous, limb-major.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u64p = ctypes.POINTER(ctypes.c_uint64)
i64p = ctypes.POINTER(ctypes.c_int64)


class NttPlan(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("modulus", "size", "root", "root_inv", "psi", "psi_inv", "size_inv")]


class _Basis(ctypes.Structure):
    _fields_ = [("n", ctypes.c_size_t), ("levels", ctypes.c_size_t), ("moduli", u64p), ("roots", u64p),
                ("plans", ctypes.POINTER(NttPlan))]


class _Rng(ctypes.Structure):
    _fields_ = [("s", ctypes.c_uint64 * 4), ("have_spare", ctypes.c_int), ("spare", ctypes.c_double)]


def build():
    """Compile liboracle.so with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sz, B, R = ctypes.c_size_t, ctypes.POINTER(_Basis), ctypes.POINTER(_Rng)
        u64, cp = ctypes.c_uint64, ctypes.c_char_p
        sigs = {
            "orc_mod_mul": (u64, [u64, u64, u64]),
            "orc_mod_pow": (u64, [u64, u64, u64]),
            "orc_ntt_plan_new": (ctypes.c_int, [u64, u64, u64, ctypes.POINTER(NttPlan), cp, sz]),
            "orc_ntt_forward": (None, [ctypes.POINTER(NttPlan), u64p]),
            "orc_ntt_inverse": (None, [ctypes.POINTER(NttPlan), u64p]),
            "orc_poly_reduce": (None, [u64p, u64p, sz, u64]),
            "orc_poly_add": (None, [u64p, u64p, u64p, sz, u64]),
            "orc_poly_sub": (None, [u64p, u64p, u64p, sz, u64]),
            "orc_poly_multiply_naive": (None, [u64p, u64p, u64p, sz, u64]),
            "orc_poly_galois": (None, [u64p, u64p, sz, u64, u64]),
            "orc_poly_multiply_ntt": (None, [u64p, u64p, u64p, ctypes.POINTER(NttPlan)]),
            "orc_basis_new": (ctypes.c_int, [sz, u64p, u64p, sz, B, cp, sz]),
            "orc_basis_free": (None, [B]),
            "orc_total_modulus_bits": (u64, [u64p, sz]),
            "orc_num_decomp_digits": (sz, [u64, sz]),
            "orc_rns_add": (None, [B, u64p, u64p, u64p]),
            "orc_rns_sub": (None, [B, u64p, u64p, u64p]),
            "orc_rns_multiply_ntt": (None, [B, u64p, u64p, u64p]),
            "orc_rns_galois": (None, [B, u64p, u64p, u64]),
            "orc_multiply_ciphertexts": (None, [B, u64p, u64p, u64p]),
            "orc_multiply_plain": (None, [B, u64p, u64p, u64p]),
            "orc_rescale": (None, [B, u64p, sz, u64p]),
            "orc_decrypt": (None, [B, u64p, u64p, u64p]),
            "orc_decrypt_quadratic": (None, [B, u64p, u64p, u64p]),
            "orc_encrypt_with_samples": (None, [B, u64p, u64p, u64p, u64p, u64p, u64p]),
            "orc_decompose": (None, [B, u64p, sz, sz, u64p]),
            "orc_decompose_signed": (None, [B, u64p, sz, sz, i64p]),
            "orc_lift": (None, [B, B, u64p, u64p]),
            "orc_mod_down": (None, [B, B, u64p, u64p]),
            "orc_keyswitch_qonly": (ctypes.c_int, [B, u64p, u64p, sz, u64p]),
            "orc_keyswitch_hybrid": (ctypes.c_int, [B, B, u64p, u64p, sz, u64p]),
            "orc_rng_seed": (None, [R, u64]),
            "orc_rng_next": (u64, [R]),
            "orc_sample_uniform": (None, [B, R, u64p]),
            "orc_sample_error": (None, [B, R, ctypes.c_double, u64p]),
            "orc_sample_ternary": (None, [B, R, u64p]),
            "orc_sample_sparse_ternary": (None, [B, R, sz, u64p]),
            "orc_keygen_public": (None, [B, R, ctypes.c_double, u64p, u64p, u64p]),
            "orc_keygen_hybrid": (None, [B, B, R, ctypes.c_double, u64p, u64p, sz, ctypes.c_int, u64p]),
            "orc_keygen_qonly": (None, [B, R, ctypes.c_double, u64p, u64p, sz, ctypes.c_int, u64p]),
            "orc_galois_element_for_rotation": (u64, [u64, ctypes.c_int64]),
            "orc_relinearize": (ctypes.c_int, [B, B, u64p, u64p, sz, u64p]),
            "orc_rotate": (ctypes.c_int, [B, B, u64p, u64, u64p, sz, u64p]),
            "orc_set_threads": (None, [ctypes.c_int]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(u64p)


def _arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


class OracleError(ValueError):
    pass


# ---------------------------------------------------------------- scalar / plan
def mod_mul(a, b, q):
    return lib().orc_mod_mul(a, b, q)


def ntt_plan(size, modulus, root):
    """NttPlan::new (src/math/ntt.rs:17-60); raises OracleError with the reference's message."""
    plan = NttPlan()
    err = ctypes.create_string_buffer(256)
    if lib().orc_ntt_plan_new(size, modulus, root, ctypes.byref(plan), err, 256) != 0:
        raise OracleError(err.value.decode())
    return plan


def ntt_forward(plan, values):
    v = _arr(values).copy()
    assert v.shape == (plan.size,)
    lib().orc_ntt_forward(ctypes.byref(plan), _p(v))
    return v


def ntt_inverse(plan, values):
    v = _arr(values).copy()
    assert v.shape == (plan.size,)
    lib().orc_ntt_inverse(ctypes.byref(plan), _p(v))
    return v


# ---------------------------------------------------------------- ModularPolynomial
def poly_new(coeffs, q):
    a = _arr(coeffs)
    out = np.empty_like(a)
    lib().orc_poly_reduce(_p(a), _p(out), a.size, q)
    return out


def poly_add(a, b, q):
    a, b = _arr(a), _arr(b)
    out = np.empty_like(a)
    lib().orc_poly_add(_p(a), _p(b), _p(out), a.size, q)
    return out


def poly_sub(a, b, q):
    a, b = _arr(a), _arr(b)
    out = np.empty_like(a)
    lib().orc_poly_sub(_p(a), _p(b), _p(out), a.size, q)
    return out


def poly_multiply_naive(a, b, q):
    a, b = _arr(a), _arr(b)
    out = np.empty_like(a)
    lib().orc_poly_multiply_naive(_p(a), _p(b), _p(out), a.size, q)
    return out


def poly_galois(a, g, q):
    a = _arr(a)
    out = np.empty_like(a)
    lib().orc_poly_galois(_p(a), _p(out), a.size, q, g)
    return out


def poly_multiply_ntt(a, b, plan):
    a, b = _arr(a), _arr(b)
    out = np.empty_like(a)
    lib().orc_poly_multiply_ntt(_p(a), _p(b), _p(out), ctypes.byref(plan))
    return out


# ---------------------------------------------------------------- RNS basis
class Basis:
    """RnsContext (src/rns/basis.rs:24-112): N plus an ordered prime chain with NTT plans."""

    def __init__(self, n, moduli, roots):
        self.n = int(n)
        self.moduli = [int(m) for m in moduli]
        self.roots = [int(r) for r in roots]
        self._m = _arr(self.moduli)
        self._r = _arr(self.roots)
        self._b = _Basis()
        err = ctypes.create_string_buffer(512)
        if lib().orc_basis_new(self.n, _p(self._m), _p(self._r), len(self.moduli), ctypes.byref(self._b), err,
                               512) != 0:
            raise OracleError(err.value.decode())

    def __del__(self):
        try:
            if _lib is not None and self._b.levels:
                _lib.orc_basis_free(ctypes.byref(self._b))
        except Exception:
            pass

    @property
    def levels(self):
        return len(self.moduli)

    @property
    def ref(self):
        return ctypes.byref(self._b)

    def prefix(self, k):
        return Basis(self.n, self.moduli[:k], self.roots[:k])

    def extend(self, other):
        return Basis(self.n, self.moduli + other.moduli, self.roots + other.roots)

    def total_modulus_bits(self):
        return int(lib().orc_total_modulus_bits(_p(self._m), self.levels))

    def plan(self, i):
        return self._b.plans[i]

    def zeros(self, polys=None):
        shape = (self.levels, self.n) if polys is None else (polys, self.levels, self.n)
        return np.zeros(shape, dtype=np.uint64)


def num_decomp_digits(total_bits, decomp_bits=30):
    return int(lib().orc_num_decomp_digits(total_bits, decomp_bits))


def _unop(fn, basis, x, *extra, polys_out=None):
    x = _arr(x)
    out = np.empty_like(x) if polys_out is None else basis.zeros(polys_out)
    fn(basis.ref, _p(x), *extra, _p(out))
    return out


def rns_add(b, x, y):
    x, y = _arr(x), _arr(y)
    out = np.empty_like(x)
    lib().orc_rns_add(b.ref, _p(x), _p(y), _p(out))
    return out


def rns_sub(b, x, y):
    x, y = _arr(x), _arr(y)
    out = np.empty_like(x)
    lib().orc_rns_sub(b.ref, _p(x), _p(y), _p(out))
    return out


def rns_multiply_ntt(b, x, y):
    x, y = _arr(x), _arr(y)
    out = np.empty_like(x)
    lib().orc_rns_multiply_ntt(b.ref, _p(x), _p(y), _p(out))
    return out


def rns_galois(b, x, g):
    x = _arr(x)
    out = np.empty_like(x)
    lib().orc_rns_galois(b.ref, _p(x), _p(out), g)
    return out


def multiply_ciphertexts(b, lhs, rhs):
    lhs, rhs = _arr(lhs), _arr(rhs)
    out = b.zeros(3)
    lib().orc_multiply_ciphertexts(b.ref, _p(lhs), _p(rhs), _p(out))
    return out


def multiply_plain(b, ct, plain):
    ct, plain = _arr(ct), _arr(plain)
    out = b.zeros(2)
    lib().orc_multiply_plain(b.ref, _p(ct), _p(plain), _p(out))
    return out


def rescale(b, x):
    """rescale_poly_rns over each polynomial of x (shape (polys, L, N) or (L, N))."""
    x = _arr(x)
    polys = 1 if x.ndim == 2 else x.shape[0]
    out = np.zeros(x.shape[:-2] + (b.levels - 1, b.n), dtype=np.uint64)
    lib().orc_rescale(b.ref, _p(x), polys, _p(out))
    return out


def decrypt(b, ct, secret):
    ct, secret = _arr(ct), _arr(secret[: b.levels])
    out = b.zeros()
    lib().orc_decrypt(b.ref, _p(ct), _p(secret), _p(out))
    return out


def decrypt_quadratic(b, ct, secret):
    ct, secret = _arr(ct), _arr(secret[: b.levels])
    out = b.zeros()
    lib().orc_decrypt_quadratic(b.ref, _p(ct), _p(secret), _p(out))
    return out


def encrypt_with_samples(b, plain, pk, u, e0, e1):
    args = [_arr(a) for a in (plain, pk, u, e0, e1)]
    out = b.zeros(2)
    lib().orc_encrypt_with_samples(b.ref, *[_p(a) for a in args], _p(out))
    return out


def decompose(b, poly, levels, decomp_bits=30):
    poly = _arr(poly)
    out = np.zeros((levels, b.levels, b.n), dtype=np.uint64)
    lib().orc_decompose(b.ref, _p(poly), decomp_bits, levels, _p(out))
    return out


def decompose_signed(b, poly, levels, decomp_bits=30):
    poly = _arr(poly)
    out = np.zeros((levels, b.n), dtype=np.int64)
    lib().orc_decompose_signed(b.ref, _p(poly), decomp_bits, levels, out.ctypes.data_as(i64p))
    return out


def lift(s, t, poly):
    poly = _arr(poly)
    out = t.zeros()
    lib().orc_lift(s.ref, t.ref, _p(poly), _p(out))
    return out


def mod_down(q, e, poly):
    poly = _arr(poly)
    out = q.zeros()
    lib().orc_mod_down(q.ref, e.ref, _p(poly), _p(out))
    return out


def keyswitch(q, e, poly, ksk):
    """keyswitch dispatch (context/ops.rs:97-109): hybrid when e is given, Q-only otherwise."""
    poly, ksk = _arr(poly), _arr(ksk)
    out = q.zeros(2)
    if e is None:
        rc = lib().orc_keyswitch_qonly(q.ref, _p(poly), _p(ksk), ksk.shape[0], _p(out))
    else:
        rc = lib().orc_keyswitch_hybrid(q.ref, e.ref, _p(poly), _p(ksk), ksk.shape[0], _p(out))
    if rc != 0:
        raise OracleError("keyswitch key must cover the active modulus chain")
    return out


def relinearize(q, e, quad, ksk):
    quad, ksk = _arr(quad), _arr(ksk)
    out = q.zeros(2)
    if lib().orc_relinearize(q.ref, e.ref if e is not None else None, _p(quad), _p(ksk), ksk.shape[0], _p(out)):
        raise OracleError("keyswitch key must cover the active modulus chain")
    return out


def rotate(q, e, ct, galois, ksk):
    ct, ksk = _arr(ct), _arr(ksk)
    out = q.zeros(2)
    if lib().orc_rotate(q.ref, e.ref if e is not None else None, _p(ct), galois, _p(ksk), ksk.shape[0], _p(out)):
        raise OracleError("keyswitch key must cover the active modulus chain")
    return out


def galois_element_for_rotation(n, steps):
    return int(lib().orc_galois_element_for_rotation(n, steps))


def set_threads(n):
    lib().orc_set_threads(n)


# ---------------------------------------------------------------- seeded sampling / keygen
class Rng:
    """xoshiro256** seeded through splitmix64 (the seeded stand-in for the reference's thread_rng)."""

    def __init__(self, seed):
        self._r = _Rng()
        lib().orc_rng_seed(ctypes.byref(self._r), seed)

    @property
    def ref(self):
        return ctypes.byref(self._r)

    def next(self):
        return int(lib().orc_rng_next(self.ref))


def sample_uniform(b, rng):
    out = b.zeros()
    lib().orc_sample_uniform(b.ref, rng.ref, _p(out))
    return out


def sample_error(b, rng, sigma=3.19):
    out = b.zeros()
    lib().orc_sample_error(b.ref, rng.ref, sigma, _p(out))
    return out


def sample_ternary(b, rng):
    out = b.zeros()
    lib().orc_sample_ternary(b.ref, rng.ref, _p(out))
    return out


def sample_sparse_ternary(b, rng, hamming):
    out = b.zeros()
    lib().orc_sample_sparse_ternary(b.ref, rng.ref, hamming, _p(out))
    return out


def keygen_public(b, rng, secret, sigma=3.19):
    secret = _arr(secret[: b.levels])
    pk = b.zeros(2)
    err = b.zeros()
    lib().orc_keygen_public(b.ref, rng.ref, sigma, _p(secret), _p(pk), _p(err))
    return pk, err


def keygen_hybrid(q, e, rng, source, secret, sigma=3.19, noiseless=False, digits=None):
    if digits is None:
        digits = num_decomp_digits(q.total_modulus_bits())
    source, secret = _arr(source[: q.levels]), _arr(secret[: q.levels])
    out = np.zeros((digits, 2, e.levels, q.n), dtype=np.uint64)
    lib().orc_keygen_hybrid(q.ref, e.ref, rng.ref, sigma, _p(source), _p(secret), digits, int(noiseless), _p(out))
    return out


def keygen_qonly(q, rng, source, pk, sigma=3.19, noiseless=False, digits=None):
    if digits is None:
        digits = num_decomp_digits(q.total_modulus_bits())
    source, pk = _arr(source[: q.levels]), _arr(pk)
    out = np.zeros((digits, 2, q.levels, q.n), dtype=np.uint64)
    lib().orc_keygen_qonly(q.ref, rng.ref, sigma, _p(source), _p(pk), digits, int(noiseless), _p(out))
    return out
