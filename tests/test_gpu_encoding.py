"""Encode / decode (host FFT in the library, residues in device memory) against the oracle's
restatement and the closed forms.  Both restate the crate's FFT operation by operation, so encode is
checked bit for bit (bit-exact parity with the crate itself is unpinned: no reference vector exists),
decode to 1e-9, and the constant encoding exactly."""
import numpy as np
import pytest

from oracle import encoding as E
from rnsckks import capi as G
from rnsckks import params as PR

pytestmark = pytest.mark.gpu


def ctx_and_params(n=1 << 12, limbs=4):
    p = PR.config_batched(n=1 << 15, limbs=limbs, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    return ctx, p


def ints_of(moduli, residues):
    return np.array(E.centered(E.reconstruct(moduli, residues), E.total_modulus(moduli)), dtype=object)


@pytest.mark.parametrize("n,scale", [(1 << 12, 40), (1 << 13, 50), (1 << 12, 90)])
def test_encode_matches_oracle(n, scale):
    ctx, p = ctx_and_params(n)
    rs = np.random.default_rng(n + scale)
    z = rs.uniform(-1, 1, n // 2) + 1j * rs.uniform(-1, 1, n // 2)
    limbs = 3
    got = ints_of(p.primes[:limbs], ctx.encode(z, limbs, scale).numpy())
    want = np.array(E.encode_integer_coeffs(n, 2.0 ** scale, z), dtype=object)
    diff = np.abs((got - want).astype(float))
    # both round an FP64 value computed through two different FFTs: they agree to a few ulps
    tol = max(1.0, 2.0 ** (scale - 40))
    assert diff.max() <= tol
    if scale == 40:
        assert np.count_nonzero(diff) <= n // 100


@pytest.mark.parametrize("n,scale", [(1 << 12, 40), (1 << 13, 50), (1 << 11, 90)])
def test_encode_bit_exact_vs_oracle(n, scale):
    """The library's host encoder and the oracle's numpy restatement of the crate's FFT (same operations
    in the same order, no fused multiply-add) give the same integers."""
    ctx, p = ctx_and_params(n)
    rs = np.random.default_rng(7 * n + scale)
    z = rs.uniform(-1, 1, n // 2) + 1j * rs.uniform(-1, 1, n // 2)
    limbs = 3
    got = ints_of(p.primes[:limbs], ctx.encode(z, limbs, scale).numpy())
    want = np.array(E.encode_integer_coeffs(n, 2.0 ** scale, z), dtype=object)
    assert np.array_equal(got, want)


def test_constant_encoding_closed_form():
    """A constant c in every slot encodes to round(c * 2^s) in coefficient 0 and zeros elsewhere
    (the inverse embedding of a constant vector is the constant polynomial)."""
    n = 1 << 12
    ctx, p = ctx_and_params(n)
    for c, s in ((0.75, 40), (-1.25, 45), (1.0, 10), (0.3, 30)):
        got = ints_of(p.primes[:2], ctx.encode(np.full(n // 2, c), 2, s).numpy())
        assert got[0] == round(c * 2 ** s)
        assert all(v == 0 for v in got[1:])


def test_decode_matches_oracle_and_round_trip():
    n = 1 << 12
    ctx, p = ctx_and_params(n, limbs=6)
    rs = np.random.default_rng(5)
    z = rs.uniform(-1, 1, n // 2) + 1j * rs.uniform(-1, 1, n // 2)
    for limbs, scale in ((6, 40), (2, 40), (1, 30)):
        enc = ctx.encode(z, limbs, scale)
        got = ctx.decode(enc, limbs, scale)
        want = E.decode(p.primes[:limbs], n, scale, enc.numpy())
        assert np.max(np.abs(got - want)) <= 1e-9
        assert np.max(np.abs(got - z)) <= n * 2.0 ** -scale * 4
    # negative (centred) values near +-Q/2 reconstruct through the big-integer path
    limbs = 3
    Q = E.total_modulus(p.primes[:limbs])
    ints = [(-1) ** k * (Q // 2 - k * 12345) for k in range(8)] + [0] * (n - 8)
    res = E.poly_from_bigint_coeffs(p.primes[:limbs], n, ints)
    buf = G.DeviceBuffer.from_numpy(res)
    got = ctx.decode(buf, limbs, 100)
    want = E.decode(p.primes[:limbs], n, 100, res)
    assert np.max(np.abs(got - want)) <= 1e-9 * np.max(np.abs(want))


def test_encode_errors():
    ctx, p = ctx_and_params()
    with pytest.raises(G.RckError, match="slot count must be N/2"):
        ctx.encode(np.zeros(10), 2, 40)
