"""ctypes binding of librnsckks.so (include/rnsckks.h).

Thin by design: device buffers, the context, keyswitching keys and one Python method per C
entry point.  Every call goes straight to the HIP library; there is no CPU fallback — a
missing or unloadable library raises at import of this module's first use.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librnsckks.so")
_lib = None

vp = ctypes.c_void_p
u32, u64, i64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64
u32p, u64p = ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64)

KS_EXACT_CONVOLUTION = 0  # keyswitch methods (rck_context_set_keyswitch_method)
KS_DIGIT_NTT = 1
KS_AUTO = 2  # both key forms, the faster method per keyswitch
ARITH_FP64 = 0  # kernel arithmetic (rck_context_set_arithmetic)
ARITH_INT64 = 1

EXPORTS = {
    "rck_last_error": (ctypes.c_char_p, []),
    "rck_version": (ctypes.c_int, []),
    "rck_malloc": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_size_t]),
    "rck_free": (ctypes.c_int, [vp]),
    "rck_memcpy_htod": (ctypes.c_int, [vp, vp, ctypes.c_size_t]),
    "rck_memcpy_dtoh": (ctypes.c_int, [vp, vp, ctypes.c_size_t]),
    "rck_memcpy_dtod": (ctypes.c_int, [vp, vp, ctypes.c_size_t, vp]),
    "rck_device_synchronize": (ctypes.c_int, []),
    "rck_stream_create": (ctypes.c_int, [ctypes.POINTER(vp)]),
    "rck_stream_destroy": (ctypes.c_int, [vp]),
    "rck_stream_synchronize": (ctypes.c_int, [vp]),
    "rck_event_create": (ctypes.c_int, [ctypes.POINTER(vp)]),
    "rck_event_destroy": (ctypes.c_int, [vp]),
    "rck_event_record": (ctypes.c_int, [vp, vp]),
    "rck_event_synchronize": (ctypes.c_int, [vp]),
    "rck_event_elapsed_ms": (ctypes.c_int, [vp, vp, ctypes.POINTER(ctypes.c_float)]),
    "rck_context_create": (ctypes.c_int, [u32, u64p, u64p, u32, u64p, u64p, u32, u32, ctypes.POINTER(vp)]),
    "rck_context_destroy": (ctypes.c_int, [vp]),
    "rck_num_decomp_digits": (ctypes.c_int, [vp, u32, u32p]),
    "rck_context_set_keyswitch_method": (ctypes.c_int, [vp, ctypes.c_int]),
    "rck_ksk_method": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int)]),
    "rck_keyswitch_exact_groups": (ctypes.c_int, [vp, u32, u32p, u32p]),
    "rck_context_set_arithmetic": (ctypes.c_int, [vp, ctypes.c_int]),
    "rck_context_arithmetic": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int)]),
    "rck_galois_element_for_rotation": (u64, [u32, i64]),
    "rck_ntt_forward": (ctypes.c_int, [vp, vp, u32, u32, vp]),
    "rck_ntt_inverse": (ctypes.c_int, [vp, vp, u32, u32, vp]),
    "rck_poly_add": (ctypes.c_int, [vp, vp, vp, vp, u32, u32, vp]),
    "rck_poly_sub": (ctypes.c_int, [vp, vp, vp, vp, u32, u32, vp]),
    "rck_poly_multiply_ntt": (ctypes.c_int, [vp, vp, vp, vp, u32, u32, vp]),
    "rck_poly_multiply_ntt_noalloc": (ctypes.c_int, [vp, vp, vp, vp, u32, u32, vp]),
    "rck_poly_galois": (ctypes.c_int, [vp, vp, vp, u32, u32, u64, vp]),
    "rck_poly_mul_add": (ctypes.c_int, [vp, vp, vp, vp, vp, u32, u32, vp]),
    "rck_ct_add": (ctypes.c_int, [vp, vp, vp, u32, u32, u32, vp, u32p, vp]),
    "rck_mult": (ctypes.c_int, [vp, vp, vp, u32, u32, u32, vp, u32p, vp]),
    "rck_mult_plain": (ctypes.c_int, [vp, vp, vp, u32, u32, vp, u32p, vp]),
    "rck_rescale": (ctypes.c_int, [vp, vp, u32, u32, vp, u32p, vp]),
    "rck_rescale_quadratic": (ctypes.c_int, [vp, vp, u32, u32, vp, u32p, vp]),
    "rck_ksk_create": (ctypes.c_int, [vp, vp, u32, u32, u32, u64, ctypes.POINTER(vp)]),
    "rck_ksk_create_device": (ctypes.c_int, [vp, vp, u32, u32, u32, u64, ctypes.POINTER(vp)]),
    "rck_ksk_destroy": (ctypes.c_int, [vp]),
    "rck_keyswitch": (ctypes.c_int, [vp, vp, u32, vp, vp, u32p, vp]),
    "rck_relinearize": (ctypes.c_int, [vp, vp, u32, u32, vp, vp, u32p, vp]),
    "rck_mult_relinearized": (ctypes.c_int, [vp, vp, vp, u32, u32, u32, vp, vp, u32p, vp]),
    "rck_ct_galois": (ctypes.c_int, [vp, vp, u32, u64, vp, vp]),
    "rck_rotate": (ctypes.c_int, [vp, vp, u32, u32, vp, vp, u32p, vp]),
    "rck_tune_keyswitch": (ctypes.c_int, [vp, u32, vp, u32, ctypes.c_char_p, ctypes.c_size_t]),
    "rck_tune_table_export": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "rck_tune_table_import": (ctypes.c_int, [ctypes.c_char_p, u32p]),
    "rck_rng_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(vp)]),
    "rck_rng_destroy": (ctypes.c_int, [vp]),
    "rck_rng_next": (ctypes.c_uint64, [vp]),
    "rck_sample_uniform": (ctypes.c_int, [vp, vp, vp, u32, u32]),
    "rck_sample_error": (ctypes.c_int, [vp, vp, ctypes.c_double, vp, u32, u32]),
    "rck_sample_ternary": (ctypes.c_int, [vp, vp, vp, u32, u32]),
    "rck_sample_sparse_ternary": (ctypes.c_int, [vp, vp, u32, vp, u32, u32]),
    "rck_keygen_public": (ctypes.c_int, [vp, vp, ctypes.c_double, vp, u32, vp]),
    "rck_keygen_relin": (ctypes.c_int, [vp, vp, ctypes.c_double, vp, vp, u32, u32, ctypes.POINTER(vp)]),
    "rck_keygen_rotation": (ctypes.c_int, [vp, vp, ctypes.c_double, vp, vp, u32, ctypes.c_int64, u32,
                                           ctypes.POINTER(vp)]),
    "rck_ksk_export": (ctypes.c_int, [vp, vp, vp]),
    "rck_ksk_info": (ctypes.c_int, [vp, u32p, u32p, u32p, ctypes.POINTER(ctypes.c_uint64)]),
    "rck_encrypt": (ctypes.c_int, [vp, vp, ctypes.c_double, vp, vp, u32, u32, vp]),
    "rck_decrypt": (ctypes.c_int, [vp, vp, u32, vp, vp, vp]),
    "rck_encode": (ctypes.c_int, [vp, vp, vp, u32, u32, u32, vp]),
    "rck_decode": (ctypes.c_int, [vp, vp, u32, u32, vp, vp]),
    "rck_ps_block_size": (ctypes.c_int, [u32, u32p]),
    "rck_ps_required_levels": (ctypes.c_int, [u32, u32, u32p, u32, u32p]),
    "rck_ps_plan": (ctypes.c_int, [vp, u32, u32, vp, u32, u32p, u32p, u32p, ctypes.POINTER(ctypes.c_uint64)]),
    "rck_ps_eval": (ctypes.c_int, [vp, vp, u32, u32, vp, u32, vp, u32, vp, u32p, u32p, vp]),
    "rck_server_create": (ctypes.c_int, [ctypes.c_uint64, u32, u32, ctypes.POINTER(vp)]),
    "rck_server_register_ring": (ctypes.c_int, [vp, vp, vp, u32]),
    "rck_server_submit": (ctypes.c_int, [vp, vp, vp, u32, u32, vp, u32, vp, ctypes.POINTER(ctypes.c_uint64)]),
    "rck_server_wait": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(ctypes.c_int), u32p,
                                       u32p, ctypes.POINTER(ctypes.c_double)]),
    "rck_server_drain": (ctypes.c_int, [vp, ctypes.c_double]),
    "rck_server_stats": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_size_t]),
    "rck_server_destroy": (ctypes.c_int, [vp]),
    "rck_bsgs_rotation_steps": (ctypes.c_int, [u32, u32p, u32, u32p]),
    "rck_diag_pack": (ctypes.c_int, [vp, vp, u32, u32, u32, ctypes.POINTER(vp)]),
    "rck_diag_info": (ctypes.c_int, [vp, u32p, u32p, u32p, u32p, u32p]),
    "rck_diag_export": (ctypes.c_int, [vp, vp]),
    "rck_diag_destroy": (ctypes.c_int, [vp]),
    "rck_bsgs_matvec": (ctypes.c_int, [vp, vp, u32, u32, u32, vp, vp, u32, vp, u32p, vp]),
    "rck_serialize_polys": (ctypes.c_int, [vp, vp, u32, u32, u32, u32, vp, ctypes.c_size_t,
                                           ctypes.POINTER(ctypes.c_size_t)]),
    "rck_deserialize_polys": (ctypes.c_int, [vp, vp, ctypes.c_size_t, vp, ctypes.c_size_t, u32p, u32p, u32p, u32p]),
    "rck_ksk_serialize": (ctypes.c_int, [vp, vp, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "rck_ksk_deserialize": (ctypes.c_int, [vp, vp, ctypes.c_size_t, ctypes.POINTER(vp)]),
    "rck_batch_create": (ctypes.c_int, [vp, ctypes.POINTER(vp)]),
    "rck_batch_add": (ctypes.c_int, [vp, ctypes.c_int, u32, vp, vp, vp, vp]),
    "rck_batch_compile": (ctypes.c_int, [vp, ctypes.c_int]),
    "rck_batch_run": (ctypes.c_int, [vp, vp]),
    "rck_batch_run_eager": (ctypes.c_int, [vp]),
    "rck_batch_synchronize": (ctypes.c_int, [vp]),
    "rck_batch_report": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_size_t]),
    "rck_batch_destroy": (ctypes.c_int, [vp]),
    "rck_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "rck_diag_copy": (ctypes.c_int, [vp, vp, ctypes.c_size_t, vp]),
    "rck_profile_report": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
}


class RckError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code
        self.msg = msg


def lib():
    """Load librnsckks.so (raises if it was not built: there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C rns-ckks-rust_amd`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RckError(rc, lib().rck_last_error().decode())


def _u64arr(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.uint64))
    return a, a.ctypes.data_as(u64p)


# ------------------------------------------------------------------ device memory
class DeviceBuffer:
    """A device allocation with a numpy shape/dtype for transfers."""

    def __init__(self, shape, dtype=np.uint64):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = vp()
        _check(lib().rck_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.shape, arr.dtype)
        _check(lib().rck_memcpy_htod(b.ptr, arr.ctypes.data, b.nbytes))
        return b

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes
        _check(lib().rck_memcpy_htod(self.ptr, arr.ctypes.data, self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        synchronize()
        _check(lib().rck_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def offset(self, elems):
        return self.ptr + int(elems) * self.dtype.itemsize

    def free(self):
        if self.ptr:
            _check(lib().rck_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.rck_free(self.ptr)
        except Exception:
            pass


def synchronize():
    _check(lib().rck_device_synchronize())


class Stream:
    """A non-blocking HIP stream (rck_stream_create)."""

    def __init__(self):
        p = vp()
        _check(lib().rck_stream_create(ctypes.byref(p)))
        self.ptr = p.value

    def synchronize(self):
        _check(lib().rck_stream_synchronize(self.ptr))

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.rck_stream_destroy(self.ptr)
        except Exception:
            pass


class Event:
    def __init__(self):
        p = vp()
        _check(lib().rck_event_create(ctypes.byref(p)))
        self.ptr = p.value

    def record(self, stream=None):
        _check(lib().rck_event_record(self.ptr, stream))

    def synchronize(self):
        _check(lib().rck_event_synchronize(self.ptr))

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        _check(lib().rck_event_elapsed_ms(self.ptr, end.ptr, ctypes.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            if _lib is not None:
                _lib.rck_event_destroy(self.ptr)
        except Exception:
            pass


def diag_copy(dst, src, nbytes, stream=None):
    _check(lib().rck_diag_copy(_p(dst), _p(src), nbytes, stream))


def tune_table_export():
    """The process-wide keyswitch tuning table as text (one entry per line)."""
    need = ctypes.c_size_t(0)
    _check(lib().rck_tune_table_export(None, 0, ctypes.byref(need)))
    buf = ctypes.create_string_buffer(need.value)
    _check(lib().rck_tune_table_export(buf, len(buf), ctypes.byref(need)))
    return buf.value.decode()


def tune_table_import(text):
    """Merges a table written by tune_table_export; returns the number of entries."""
    k = ctypes.c_uint32(0)
    _check(lib().rck_tune_table_import(text.encode(), ctypes.byref(k)))
    return k.value


def ps_block_size(term_count):
    out = ctypes.c_uint32()
    _check(lib().rck_ps_block_size(term_count, ctypes.byref(out)))
    return out.value


def ps_required_levels(ciphertext_level, term_count):
    cnt = ctypes.c_uint32()
    buf = (ctypes.c_uint32 * 128)()
    _check(lib().rck_ps_required_levels(ciphertext_level, term_count, buf, 128, ctypes.byref(cnt)))
    return [buf[i] for i in range(cnt.value)]


def bsgs_rotation_steps(dim):
    cnt = ctypes.c_uint32()
    buf = (ctypes.c_uint32 * 4096)()
    _check(lib().rck_bsgs_rotation_steps(dim, buf, 4096, ctypes.byref(cnt)))
    return [buf[i] for i in range(cnt.value)]


class Diag:
    """rck_diag: packed BSGS diagonals of a matrix, NTT-resident."""

    def __init__(self, ctx, matrix, level, scale_bits):
        m = np.ascontiguousarray(matrix, dtype=np.float64)
        assert m.ndim == 2 and m.shape[0] == m.shape[1]
        out = vp()
        _check(lib().rck_diag_pack(ctx.ptr, m.ctypes.data, m.shape[0], level, scale_bits, ctypes.byref(out)))
        self.ptr, self.ctx = out.value, ctx
        v = [ctypes.c_uint32() for _ in range(5)]
        _check(lib().rck_diag_info(self.ptr, *[ctypes.byref(x) for x in v]))
        self.dim, self.baby_step, self.giant_steps, self.level, self.scale_bits = [x.value for x in v]

    def export(self):
        out = np.zeros((self.giant_steps, self.baby_step, self.level + 1, self.ctx.n), dtype=np.uint64)
        _check(lib().rck_diag_export(self.ptr, out.ctypes.data))
        return out

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.rck_diag_destroy(self.ptr)
        except Exception:
            pass


def profile_enable(on=True):
    _check(lib().rck_profile_enable(1 if on else 0))


def profile_report():
    """{kernel: (launches, total_ms)} since the previous report (synchronizes the device)."""
    import json
    buf = ctypes.create_string_buffer(1 << 16)
    _check(lib().rck_profile_report(buf, len(buf)))
    return {k: (int(v[0]), float(v[1])) for k, v in json.loads(buf.value.decode()).items()}


def galois_element_for_rotation(poly_degree, steps):
    return int(lib().rck_galois_element_for_rotation(poly_degree, steps))


# ------------------------------------------------------------------ context
def _p(x):
    """device pointer of a DeviceBuffer / int / None"""
    if x is None:
        return None
    return x.ptr if isinstance(x, DeviceBuffer) else int(x)


class Ksk:
    def __init__(self, ptr, digits, level, aux, galois):
        self.ptr, self.digits, self.level, self.aux, self.galois = ptr, digits, level, aux, galois

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.rck_ksk_destroy(self.ptr)
        except Exception:
            pass


class Rng:
    """rck_rng: the seeded sampling stream (xoshiro256** seeded through splitmix64)."""

    def __init__(self, seed):
        out = vp()
        _check(lib().rck_rng_create(seed, ctypes.byref(out)))
        self.ptr = out.value

    def next(self):
        return int(lib().rck_rng_next(self.ptr))

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.rck_rng_destroy(self.ptr)
        except Exception:
            pass


class Context:
    """rck_context: RnsCkksContext's hot-path entry points on device buffers."""

    def __init__(self, params=None, *, poly_degree=None, primes=None, roots=None, aux_primes=(), aux_roots=(),
                 scale_bits=40):
        if params is not None:
            poly_degree, primes, roots = params.poly_degree, params.primes, params.roots
            aux_primes, aux_roots, scale_bits = params.aux_primes, params.aux_roots, params.scale_bits
        self.n = int(poly_degree)
        self.primes = [int(x) for x in primes]
        self.aux_primes = [int(x) for x in aux_primes]
        qa, qp = _u64arr(self.primes)
        ga, gp = _u64arr(roots)
        pa, pp = _u64arr(self.aux_primes if self.aux_primes else [0])
        ha, hp = _u64arr(list(aux_roots) if aux_primes else [0])
        out = vp()
        _check(lib().rck_context_create(self.n, qp, gp, len(self.primes), pp, hp, len(self.aux_primes), scale_bits,
                                        ctypes.byref(out)))
        self.ptr = out.value
        self.scale_bits = scale_bits

    def __del__(self):
        try:
            if getattr(self, "ptr", None) and _lib is not None:
                _lib.rck_context_destroy(self.ptr)
        except Exception:
            pass

    @property
    def L(self):
        return len(self.primes)

    @property
    def K(self):
        return len(self.aux_primes)

    def num_decomp_digits(self, limbs):
        d = ctypes.c_uint32()
        _check(lib().rck_num_decomp_digits(self.ptr, limbs, ctypes.byref(d)))
        return d.value

    # L0 / L1 / L2
    def ntt_forward(self, data, nq, np_=0, stream=None):
        _check(lib().rck_ntt_forward(self.ptr, _p(data), nq, np_, stream))

    def ntt_inverse(self, data, nq, np_=0, stream=None):
        _check(lib().rck_ntt_inverse(self.ptr, _p(data), nq, np_, stream))

    def poly_add(self, a, b, out, nq, np_=0, stream=None):
        _check(lib().rck_poly_add(self.ptr, _p(a), _p(b), _p(out), nq, np_, stream))

    def poly_sub(self, a, b, out, nq, np_=0, stream=None):
        _check(lib().rck_poly_sub(self.ptr, _p(a), _p(b), _p(out), nq, np_, stream))

    def poly_multiply_ntt(self, a, b, out, nq, np_=0, stream=None):
        _check(lib().rck_poly_multiply_ntt(self.ptr, _p(a), _p(b), _p(out), nq, np_, stream))

    def poly_multiply_ntt_noalloc(self, a, b, out, nq, np_=0, stream=None):
        """The same product through the library's fixed device scratch (no device allocation)."""
        _check(lib().rck_poly_multiply_ntt_noalloc(self.ptr, _p(a), _p(b), _p(out), nq, np_, stream))

    def poly_galois(self, a, out, nq, galois, np_=0, stream=None):
        _check(lib().rck_poly_galois(self.ptr, _p(a), _p(out), nq, np_, galois, stream))

    def poly_mul_add(self, a, b, c, out, nq, np_=0, stream=None):
        _check(lib().rck_poly_mul_add(self.ptr, _p(a), _p(b), _p(c), _p(out), nq, np_, stream))

    # L3
    def _scaled(self, fn, *args):
        s = ctypes.c_uint32()
        _check(fn(*args, ctypes.byref(s), None))
        return s.value

    def ct_add(self, lhs, rhs, limbs, ls, rs, out):
        return self._scaled(lib().rck_ct_add, self.ptr, _p(lhs), _p(rhs), limbs, ls, rs, _p(out))

    def mult(self, lhs, rhs, limbs, ls, rs, out):
        return self._scaled(lib().rck_mult, self.ptr, _p(lhs), _p(rhs), limbs, ls, rs, _p(out))

    def mult_plain(self, ct, pt, limbs, scale, out):
        return self._scaled(lib().rck_mult_plain, self.ptr, _p(ct), _p(pt), limbs, scale, _p(out))

    def rescale(self, ct, limbs, scale, out):
        return self._scaled(lib().rck_rescale, self.ptr, _p(ct), limbs, scale, _p(out))

    def rescale_quadratic(self, quad, limbs, scale, out):
        return self._scaled(lib().rck_rescale_quadratic, self.ptr, _p(quad), limbs, scale, _p(out))

    def set_keyswitch_method(self, method):
        """Keyswitch method of the keys created from now on (KS_AUTO, KS_EXACT_CONVOLUTION or KS_DIGIT_NTT)."""
        _check(lib().rck_context_set_keyswitch_method(self.ptr, method))

    def set_arithmetic(self, arithmetic):
        """ARITH_FP64 (moduli below 2^50 + 2^46 only) or ARITH_INT64 (any modulus below 2^64)."""
        _check(lib().rck_context_set_arithmetic(self.ptr, arithmetic))

    @property
    def arithmetic(self):
        m = ctypes.c_int()
        _check(lib().rck_context_arithmetic(self.ptr, ctypes.byref(m)))
        return m.value

    def exact_group_size(self, limbs):
        """(digits per group, groups) of the exact-convolution keyswitch at `limbs` Q limbs."""
        g, k = ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().rck_keyswitch_exact_groups(self.ptr, limbs, ctypes.byref(g), ctypes.byref(k)))
        return g.value, k.value

    @staticmethod
    def ksk_method(ksk):
        m = ctypes.c_int()
        _check(lib().rck_ksk_method(ksk.ptr, ctypes.byref(m)))
        return m.value

    def ksk_create(self, key, level, aux, galois=0):
        """key: numpy uint64 [digits][2][level+1+aux][N] (host) or a DeviceBuffer of that shape."""
        out = vp()
        if isinstance(key, DeviceBuffer):
            digits = key.shape[0]
            _check(lib().rck_ksk_create_device(self.ptr, key.ptr, digits, level, aux, galois, ctypes.byref(out)))
        else:
            key = np.ascontiguousarray(key, dtype=np.uint64)
            digits = key.shape[0]
            _check(lib().rck_ksk_create(self.ptr, key.ctypes.data, digits, level, aux, galois, ctypes.byref(out)))
        return Ksk(out.value, digits, level, aux, galois)

    # ---- seeded sampling, keys, encryption
    def sample_uniform(self, rng, nq, np_=0):
        out = DeviceBuffer((nq + np_, self.n))
        _check(lib().rck_sample_uniform(self.ptr, rng.ptr, out.ptr, nq, np_))
        return out

    def sample_error(self, rng, nq, np_=0, sigma=3.19):
        out = DeviceBuffer((nq + np_, self.n))
        _check(lib().rck_sample_error(self.ptr, rng.ptr, sigma, out.ptr, nq, np_))
        return out

    def sample_ternary(self, rng, nq, np_=0):
        out = DeviceBuffer((nq + np_, self.n))
        _check(lib().rck_sample_ternary(self.ptr, rng.ptr, out.ptr, nq, np_))
        return out

    def sample_sparse_ternary(self, rng, hamming, nq, np_=0):
        out = DeviceBuffer((nq + np_, self.n))
        _check(lib().rck_sample_sparse_ternary(self.ptr, rng.ptr, hamming, out.ptr, nq, np_))
        return out

    def keygen_public(self, rng, secret, limbs, sigma=3.19):
        pk = DeviceBuffer((2, limbs, self.n))
        _check(lib().rck_keygen_public(self.ptr, rng.ptr, sigma, _p(secret), limbs, pk.ptr))
        return pk

    def _key_info(self, ptr):
        d, l, a, g = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        _check(lib().rck_ksk_info(ptr, ctypes.byref(d), ctypes.byref(l), ctypes.byref(a), ctypes.byref(g)))
        return Ksk(ptr, d.value, l.value, a.value, g.value)

    def keygen_relin(self, rng, secret, level, pk=None, pk_limbs=0, sigma=3.19):
        out = vp()
        _check(lib().rck_keygen_relin(self.ptr, rng.ptr, sigma, _p(secret), _p(pk) if pk is not None else None,
                                      pk_limbs, level, ctypes.byref(out)))
        return self._key_info(out.value)

    def keygen_rotation(self, rng, secret, steps, level, pk=None, pk_limbs=0, sigma=3.19):
        out = vp()
        _check(lib().rck_keygen_rotation(self.ptr, rng.ptr, sigma, _p(secret), _p(pk) if pk is not None else None,
                                         pk_limbs, steps, level, ctypes.byref(out)))
        return self._key_info(out.value)

    def encode(self, slots, limbs, scale_bits, out=None):
        """slots: N/2 complex (or real) values -> DeviceBuffer [limbs][N]"""
        z = np.asarray(slots)
        re = np.ascontiguousarray(z.real, dtype=np.float64)
        im = np.ascontiguousarray(z.imag if np.iscomplexobj(z) else np.zeros_like(re), dtype=np.float64)
        out = out if out is not None else DeviceBuffer((limbs, self.n))
        _check(lib().rck_encode(self.ptr, re.ctypes.data, im.ctypes.data, re.size, limbs, scale_bits, _p(out)))
        return out

    def decode(self, poly, limbs, scale_bits):
        """DeviceBuffer [limbs][N] -> N/2 complex slots"""
        re = np.zeros(self.n // 2)
        im = np.zeros(self.n // 2)
        _check(lib().rck_decode(self.ptr, _p(poly), limbs, scale_bits, re.ctypes.data, im.ctypes.data))
        return re + 1j * im

    def ps_plan(self, limbs, scale_bits, coeffs):
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        ol, os_, m, fp = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        _check(lib().rck_ps_plan(self.ptr, limbs, scale_bits, c.ctypes.data, c.size, ctypes.byref(ol),
                                 ctypes.byref(os_), ctypes.byref(m), ctypes.byref(fp)))
        return {"limbs": ol.value, "scale_bits": os_.value, "mults": m.value, "footprint": fp.value}

    def ps_eval(self, ct, limbs, scale_bits, coeffs, relin_keys, out, stream=None):
        """relin_keys: {level: Ksk}.  Returns (out_limbs, out_scale)."""
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        nk = (max(relin_keys) + 1) if relin_keys else 0
        arr = (vp * max(nk, 1))()
        for lv, k in relin_keys.items():
            arr[lv] = k.ptr
        ol, os_ = ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().rck_ps_eval(self.ptr, _p(ct), limbs, scale_bits, c.ctypes.data, c.size, arr, nk, _p(out),
                                 ctypes.byref(ol), ctypes.byref(os_), stream))
        return ol.value, os_.value

    def bsgs_matvec(self, cts, batch, limbs, scale_bits, diag, keys_by_step, out, stream=None):
        """keys_by_step: {step: Ksk}.  Returns the output scale."""
        nk = (max(keys_by_step) + 1) if keys_by_step else 0
        arr = (vp * max(nk, 1))()
        for st, k in keys_by_step.items():
            arr[st] = k.ptr
        sc = ctypes.c_uint32()
        _check(lib().rck_bsgs_matvec(self.ptr, _p(cts), batch, limbs, scale_bits, diag.ptr, arr, nk, _p(out),
                                     ctypes.byref(sc), stream))
        return sc.value

    def serialize_polys(self, polys, count, nq, np_=0, scale_bits=0):
        need = ctypes.c_size_t()
        _check(lib().rck_serialize_polys(self.ptr, _p(polys), count, nq, np_, scale_bits, None, 0, ctypes.byref(need)))
        buf = (ctypes.c_uint8 * need.value)()
        _check(lib().rck_serialize_polys(self.ptr, _p(polys), count, nq, np_, scale_bits, buf, need.value,
                                         ctypes.byref(need)))
        return bytes(buf)

    def deserialize_polys(self, data):
        raw = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
        v = [ctypes.c_uint32() for _ in range(4)]
        _check(lib().rck_deserialize_polys(self.ptr, raw, len(data), None, 0, *[ctypes.byref(x) for x in v]))
        count, nq, np_, scale = [x.value for x in v]
        out = DeviceBuffer((count, nq + np_, self.n))
        _check(lib().rck_deserialize_polys(self.ptr, raw, len(data), out.ptr, out.nbytes, *[ctypes.byref(x) for x in v]))
        return out, {"count": count, "nq": nq, "np": np_, "scale_bits": scale}

    def ksk_serialize(self, ksk):
        need = ctypes.c_size_t()
        _check(lib().rck_ksk_serialize(self.ptr, ksk.ptr, None, 0, ctypes.byref(need)))
        buf = (ctypes.c_uint8 * need.value)()
        _check(lib().rck_ksk_serialize(self.ptr, ksk.ptr, buf, need.value, ctypes.byref(need)))
        return bytes(buf)

    def ksk_deserialize(self, data):
        raw = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
        out = vp()
        _check(lib().rck_ksk_deserialize(self.ptr, raw, len(data), ctypes.byref(out)))
        return self._key_info(out.value)

    def ksk_export(self, ksk):
        out = np.zeros((ksk.digits, 2, ksk.level + 1 + ksk.aux, self.n), dtype=np.uint64)
        _check(lib().rck_ksk_export(self.ptr, ksk.ptr, out.ctypes.data))
        return out

    def encrypt(self, rng, plain, pk, pk_limbs, limbs, sigma=3.19):
        out = DeviceBuffer((2, limbs, self.n))
        _check(lib().rck_encrypt(self.ptr, rng.ptr, sigma, _p(plain), _p(pk), pk_limbs, limbs, out.ptr))
        return out

    def decrypt(self, ct, limbs, secret, out=None, stream=None):
        out = out if out is not None else DeviceBuffer((limbs, self.n))
        _check(lib().rck_decrypt(self.ptr, _p(ct), limbs, _p(secret), _p(out), stream))
        return out

    def tune_keyswitch(self, limbs, ksk, batch=1):
        import json
        buf = ctypes.create_string_buffer(1 << 20)
        _check(lib().rck_tune_keyswitch(self.ptr, limbs, ksk.ptr, batch, buf, len(buf)))
        return json.loads(buf.value.decode())

    def keyswitch(self, poly, limbs, ksk, out):
        return self._scaled(lib().rck_keyswitch, self.ptr, _p(poly), limbs, ksk.ptr, _p(out))

    def relinearize(self, quad, limbs, scale, ksk, out):
        return self._scaled(lib().rck_relinearize, self.ptr, _p(quad), limbs, scale, ksk.ptr, _p(out))

    def mult_relinearized(self, lhs, rhs, limbs, ls, rs, ksk, out, stream=None):
        s = ctypes.c_uint32()
        _check(lib().rck_mult_relinearized(self.ptr, _p(lhs), _p(rhs), limbs, ls, rs, ksk.ptr, _p(out),
                                           ctypes.byref(s), stream))
        return s.value

    def ct_galois(self, ct, limbs, galois, out):
        _check(lib().rck_ct_galois(self.ptr, _p(ct), limbs, galois, _p(out), None))

    def rotate(self, ct, limbs, scale, ksk, out):
        return self._scaled(lib().rck_rotate, self.ptr, _p(ct), limbs, scale, ksk.ptr, _p(out))


OP_MULT_RELIN, OP_ROTATE, OP_MULT_PLAIN, OP_MULT_RELIN_RESCALE = 1, 2, 3, 4


class Batch:
    """rck_batch: independent ops planned into grouped launches and captured in one hipGraph."""

    def __init__(self, ctx):
        self.ctx = ctx
        p = vp()
        _check(lib().rck_batch_create(ctx.ptr, ctypes.byref(p)))
        self.ptr = p.value
        self._keep = []

    def add(self, op, limbs, in0, in1, key, out):
        self._keep.append((in0, in1, key, out))
        _check(lib().rck_batch_add(self.ptr, op, limbs, _p(in0), _p(in1), key.ptr if key is not None else None,
                                   _p(out)))

    def compile(self, tune=True):
        _check(lib().rck_batch_compile(self.ptr, 1 if tune else 0))

    def run(self, stream=None):
        """Launch the captured graph on `stream` (a Stream; None = the batch's own stream)."""
        _check(lib().rck_batch_run(self.ptr, stream.ptr if isinstance(stream, Stream) else stream))

    def run_eager(self):
        _check(lib().rck_batch_run_eager(self.ptr))

    def synchronize(self):
        _check(lib().rck_batch_synchronize(self.ptr))

    def report(self):
        import json
        buf = ctypes.create_string_buffer(1 << 20)
        _check(lib().rck_batch_report(self.ptr, buf, len(buf)))
        return json.loads(buf.value.decode())

    def __del__(self):
        try:
            if getattr(self, "ptr", None) and _lib is not None:
                _lib.rck_batch_destroy(self.ptr)
        except Exception:
            pass


class Server:
    """rck_server: continuous-batching request loop for Paterson-Stockmeyer evaluations."""

    def __init__(self, budget_bytes, max_active=0, lanes=4):
        out = vp()
        _check(lib().rck_server_create(int(budget_bytes), max_active, lanes, ctypes.byref(out)))
        self.ptr = out.value
        self._keep = []

    def register_ring(self, ctx, relin_keys):
        nk = (max(relin_keys) + 1) if relin_keys else 0
        arr = (vp * max(nk, 1))()
        for lv, k in relin_keys.items():
            arr[lv] = k.ptr
        self._keep.append((ctx, dict(relin_keys)))
        _check(lib().rck_server_register_ring(self.ptr, ctx.ptr, arr, nk))

    def submit(self, ctx, ct, limbs, scale_bits, coeffs, out):
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        rid = ctypes.c_uint64()
        _check(lib().rck_server_submit(self.ptr, ctx.ptr, _p(ct), limbs, scale_bits, c.ctypes.data, c.size, _p(out),
                                       ctypes.byref(rid)))
        return rid.value

    def wait(self, rid, timeout=60.0):
        done, ol, osc, lat = ctypes.c_int(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_double()
        _check(lib().rck_server_wait(self.ptr, rid, timeout, ctypes.byref(done), ctypes.byref(ol), ctypes.byref(osc),
                                     ctypes.byref(lat)))
        if not done.value:
            return None
        return {"limbs": ol.value, "scale_bits": osc.value, "latency_ms": lat.value}

    def drain(self, timeout=600.0):
        _check(lib().rck_server_drain(self.ptr, timeout))

    def stats(self):
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        _check(lib().rck_server_stats(self.ptr, buf, len(buf)))
        return json.loads(buf.value.decode())

    def close(self):
        if self.ptr and _lib is not None:
            _check(_lib.rck_server_destroy(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
