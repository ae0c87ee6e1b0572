"""Allocation-free negacyclic product (rck_poly_multiply_ntt_noalloc): bit-identical to
rck_poly_multiply_ntt on two calls in a row through the shared fixed scratch (the second smaller),
and inputs whose NTT forms exceed the fixed scratch are rejected with an error before any launch."""
import numpy as np
import pytest

from rnsckks import capi as G
from rnsckks import params as PR

pytestmark = pytest.mark.gpu


def uniform(moduli, n, rs, polys):
    out = np.empty((polys, len(moduli), n), dtype=np.uint64)
    for l, q in enumerate(moduli):
        out[:, l] = rs.integers(0, q, size=(polys, n), dtype=np.uint64)
    return out


def test_noalloc_product_matches_the_allocating_entry_point():
    p = PR.config_per_op_gpu()  # N=2^16, 24 Q primes + 2 aux primes
    n = p.poly_degree
    ctx = G.Context(p)
    rs = np.random.default_rng(71)
    for nq, np_ in ((24, 2), (5, 0)):  # the whole eval basis, then a smaller input on the same scratch
        moduli = p.primes[:nq] + (p.aux_primes if np_ else [])
        a = G.DeviceBuffer.from_numpy(uniform(moduli, n, rs, 1)[0])
        b = G.DeviceBuffer.from_numpy(uniform(moduli, n, rs, 1)[0])
        want, got = G.DeviceBuffer((nq + np_, n)), G.DeviceBuffer((nq + np_, n))
        print(f"[no-alloc] N={n} nq={nq} np={np_}", flush=True)
        ctx.poly_multiply_ntt(a, b, want, nq, np_)
        ctx.poly_multiply_ntt_noalloc(a, b, got, nq, np_)
        assert np.array_equal(got.numpy(), want.numpy()), (nq, np_)
    # a smaller ring on another context shares the same device scratch
    small = PR.config_plain_cpu()  # N=2^12, 3 primes
    c2 = G.Context(small)
    a = G.DeviceBuffer.from_numpy(uniform(small.primes, small.poly_degree, rs, 1)[0])
    b = G.DeviceBuffer.from_numpy(uniform(small.primes, small.poly_degree, rs, 1)[0])
    want, got = G.DeviceBuffer((3, small.poly_degree)), G.DeviceBuffer((3, small.poly_degree))
    c2.poly_multiply_ntt(a, b, want, 3)
    c2.poly_multiply_ntt_noalloc(a, b, got, 3)
    assert np.array_equal(got.numpy(), want.numpy())


def test_noalloc_rejects_inputs_beyond_the_fixed_scratch():
    n = 1 << 16
    ps = PR.primes_below(31, n, 66)  # 64 Q primes of 31 bits (1984 bits, within the decomposition's range)
    p = PR.RnsCkksParams.with_primes(n, ps[:64], ps[64:], 40)
    ctx = G.Context(p)
    a = G.DeviceBuffer((66, n))
    out = G.DeviceBuffer((66, n))
    with pytest.raises(G.RckError, match="above the fixed"):
        ctx.poly_multiply_ntt_noalloc(a, a, out, 64, 2)  # 2 x 66 x 2^16 x 8 bytes > 64 MiB
    with pytest.raises(G.RckError, match="limb count must be in"):
        ctx.poly_multiply_ntt_noalloc(a, a, out, 65, 0)
