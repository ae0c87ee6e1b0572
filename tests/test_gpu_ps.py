"""Paterson-Stockmeyer evaluation on the device equals the oracle's restatement of the crate's
algorithm bit for bit, and decrypts to the FP64 polynomial within the crate's tolerance."""
import numpy as np
import pytest

from oracle import encoding as E
from oracle import oracle as O
from oracle import ps as OPS
from rnsckks import capi as G
from rnsckks import params as PR

pytestmark = pytest.mark.gpu


def uniform(moduli, n, rs, polys):
    out = np.empty((polys, len(moduli), n), dtype=np.uint64)
    for l, q in enumerate(moduli):
        out[:, l] = rs.integers(0, q, size=(polys, n), dtype=np.uint64)
    return out


@pytest.mark.parametrize("coeffs", [
    [0.5, -0.25],                                   # degree 1: one block of 2
    [0.75, -0.5, 1.25, -0.25, 0.125],               # the crate's test polynomial (block 3, 2 blocks)
    [0.0, 1.0, 0.0, -0.5, 0.0, 0.25, 0.0, -0.125],  # zero coefficients skipped, block 3, 3 blocks
    [0.3],                                          # constant
])
def test_ps_matches_oracle_bit_exact(coeffs):
    n, L = 1 << 12, 7
    p = PR.config_batched(n=1 << 15, limbs=L, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    rs = np.random.default_rng(len(coeffs))
    keys_np, keys = {}, {}
    for lv in OPS.required_levels(L - 1, len(coeffs)):
        D = ctx.num_decomp_digits(lv + 1)
        keys_np[lv] = np.stack([uniform(p.primes[: lv + 1] + p.aux_primes, n, rs, 2) for _ in range(D)])
        keys[lv] = ctx.ksk_create(keys_np[lv], lv, 2)
    ct = uniform(p.primes, n, rs, 2)
    out = G.DeviceBuffer((2, L, n))
    ol, osc = ctx.ps_eval(G.DeviceBuffer.from_numpy(ct), L, 40, coeffs, keys, out)
    R = OPS.Ring(n, p.primes, p.roots, p.aux_primes, p.aux_roots, keys_np)
    want, wscale = OPS.ps_eval(R, (ct, 40), coeffs)
    assert (ol, osc) == (want.shape[1], wscale)
    got = out.numpy().reshape(-1)[: 2 * ol * n].reshape(2, ol, n)
    assert np.array_equal(got, want)
    plan = ctx.ps_plan(L, 40, coeffs)
    assert (plan["limbs"], plan["scale_bits"]) == (ol, osc)


def test_ps_reference_polynomial_decrypts_within_tolerance():
    """paterson_stockmeyer_eval_matches_slotwise_polynomial_on_realistic_params (:368-404) with device
    keys: |decoded - FP64| <= 2e-2 on the crate's inputs."""
    p = PR.realistic_8_level()
    n, L = p.poly_degree, len(p.primes)
    ctx = G.Context(p)
    g = G.Rng(64)
    s = ctx.sample_sparse_ternary(g, 64, L)
    pk = ctx.keygen_public(g, s, L)
    coeffs = [0.75, -0.5, 1.25, -0.25, 0.125]
    keys = {lv: ctx.keygen_relin(g, s, lv) for lv in G.ps_required_levels(L - 1, len(coeffs))}
    x = np.zeros(n // 2)
    x[:4] = [0.125, -0.2, 0.05, 0.3]
    ct = ctx.encrypt(g, G.DeviceBuffer.from_numpy(E.encode(p.primes, n, 45, x.astype(complex))), pk, L, L)
    out = G.DeviceBuffer((2, L, n))
    ol, osc = ctx.ps_eval(ct, L, 45, coeffs, keys, out)
    dec = ctx.decrypt(out, ol, s)
    got = ctx.decode(dec, ol, osc).real[:4]
    want = np.polyval(coeffs[::-1], x[:4])
    assert np.max(np.abs(got - want)) <= 2e-2


def test_ps_degree_31_decode_error():
    """A degree-31 polynomial with coefficients in [-1, 1] on inputs in [-1, 1] at scale 2^40 with
    primes just above 2^50: the result matches FP64 evaluation of the coefficients as the algorithm
    encodes them (round(c * 2^(working - 40)) / 2^(working - 40)) to 1e-3, and the exact
    coefficients to the quantisation bound."""
    n, L = 1 << 13, 12
    p = PR.config_batched(n=1 << 15, limbs=L, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    g = G.Rng(5)
    s = ctx.sample_ternary(g, L)
    pk = ctx.keygen_public(g, s, L)
    rs = np.random.default_rng(9)
    coeffs = rs.uniform(-1, 1, 32)
    levels = G.ps_required_levels(L - 1, 32)
    keys = {lv: ctx.keygen_relin(g, s, lv) for lv in levels}
    x = rs.uniform(-1, 1, n // 2)
    ct = ctx.encrypt(g, ctx.encode(x, L, 40), pk, L, L)
    out = G.DeviceBuffer((2, L, n))
    ol, osc = ctx.ps_eval(ct, L, 40, coeffs, keys, out)
    got = ctx.decode(ctx.decrypt(out, ol, s), ol, osc).real
    working = min(int(p.primes[lv]).bit_length() - 1 for lv in levels)
    cs = working - 40
    block = G.ps_block_size(32)
    quant = np.array([c if i % block == 0 else np.round(c * 2.0 ** cs) / 2.0 ** cs for i, c in enumerate(coeffs)])
    assert np.max(np.abs(got - np.polyval(quant[::-1], x))) <= 1e-3
    bound = sum(2.0 ** -(cs + 1) for i in range(32) if i % block) + 1e-3
    assert np.max(np.abs(got - np.polyval(coeffs[::-1], x))) <= bound


def test_ps_errors():
    n, L = 1 << 12, 4
    p = PR.config_batched(n=1 << 15, limbs=L, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    ct = G.DeviceBuffer((2, L, n))
    out = G.DeviceBuffer((2, L, n))
    with pytest.raises(G.RckError, match="polynomial degree is too large"):
        ctx.ps_eval(ct, L, 40, np.ones(20), {}, out)
    with pytest.raises(G.RckError, match="missing Paterson-Stockmeyer relinearization key at level 3"):
        ctx.ps_eval(ct, L, 40, [1.0, 2.0, 3.0], {}, out)
    with pytest.raises(G.RckError, match="at least one coefficient"):
        ctx.ps_eval(ct, L, 40, [], {}, out)
