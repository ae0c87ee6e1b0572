"""Every digit-MAC launch shape the autotuner may pick gives the same keyswitch bit for bit:
row-pass width 16 (rows 4 / 16, every load schedule) and width 8 (rows 8 / 16), eval-limb groups
and digit splits, for every row length (ring degrees 2^12 .. 2^16), with digit-NTT keys; and every
multiply-accumulate tile of the exact-convolution method.  The default shapes are pinned to the
oracle by the parity tests; here each shape is forced through the tuning table."""
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


SHAPES = [  # group, dsplit, rows, prefetch, width
    (None, 1, 16, 0, 16), (None, 2, 4, 1, 16), (None, 1, 4, 2, 16), (None, 1, 16, 3, 16), (None, 1, 8, 0, 8),
    (None, 2, 16, 0, 8), (4, 1, 8, 0, 8), (1, 1, 16, 3, 16), (None, 1, 4, 4, 16), (None, 2, 16, 4, 16),
    (None, 1, 4, 5, 16), (None, 2, 4, 5, 16), (3, 1, 4, 5, 16),
    (None, 1, 4, 6, 16), (None, 2, 4, 6, 16), (2, 1, 4, 6, 16),
]


@pytest.mark.parametrize("logn,limbs", [(12, 4), (13, 6), (14, 5), (15, 4), (16, 8)])
def test_mac_shapes_bit_identical(logn, limbs):
    n = 1 << logn
    p = PR.config_batched(n=1 << 16, limbs=limbs, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    rs = np.random.default_rng(logn)
    D = ctx.num_decomp_digits(limbs)
    E = limbs + 2
    key = np.stack([uniform(p.primes + p.aux_primes, n, rs, 2) for _ in range(D)])
    ctx.set_keyswitch_method(G.KS_DIGIT_NTT)
    k = ctx.ksk_create(key, limbs - 1, 2)
    assert ctx.ksk_method(k) == G.KS_DIGIT_NTT
    quad = G.DeviceBuffer.from_numpy(uniform(p.primes, n, rs, 3))
    out = G.DeviceBuffer((2, limbs, n))
    ctx.relinearize(quad, limbs, 80, k, out)
    want = out.numpy()
    for grp, ds, rows, pf, width in SHAPES:
        g = E if grp is None else min(grp, E)
        G.tune_table_import(f"{logn} {limbs} {E} {D} 1 {g} {ds} {rows} {pf} {width} 1.0\n")
        out.upload(np.zeros_like(want))
        ctx.relinearize(quad, limbs, 80, k, out)
        assert np.array_equal(out.numpy(), want), (grp, ds, rows, pf, width)


@pytest.mark.parametrize("logn,limbs", [(12, 4), (13, 5), (16, 6)])
def test_mac_streamed_batch_bit_identical(logn, limbs):
    """The streamed MACs (rows copied into LDS while the previous digit is transformed): four
    ciphertexts per workgroup (load schedule 5; a batch of 5 leaves one workgroup with a single
    active wave) and four digit ranges of one ciphertext per workgroup (6), with one or two digit
    splits, on a batch of 5 mult+relin give the per-op results."""
    n, B = 1 << logn, 5
    p = PR.config_batched(n=1 << 16, limbs=limbs, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    rs = np.random.default_rng(100 + logn)
    D = ctx.num_decomp_digits(limbs)
    E = limbs + 2
    key = np.stack([uniform(p.primes + p.aux_primes, n, rs, 2) for _ in range(D)])
    ctx.set_keyswitch_method(G.KS_DIGIT_NTT)
    k = ctx.ksk_create(key, limbs - 1, 2)
    P = limbs * n
    lhs = G.DeviceBuffer.from_numpy(uniform(p.primes, n, rs, 2 * B).reshape(B, 2, limbs, n))
    rhs = G.DeviceBuffer.from_numpy(uniform(p.primes, n, rs, 2 * B).reshape(B, 2, limbs, n))
    quad = G.DeviceBuffer((3, limbs, n))
    one = G.DeviceBuffer((2, limbs, n))
    want = []
    for b in range(B):
        ctx.mult(lhs.offset(b * 2 * P), rhs.offset(b * 2 * P), limbs, 40, 40, quad)
        ctx.relinearize(quad, limbs, 80, k, one)
        want.append(one.numpy())
    for ds, pf in ((1, 5), (2, 5), (1, 6), (2, 6)):
        G.tune_table_import(f"{logn} {limbs} {E} {D} {B} {E} {ds} 4 {pf} 16 1.0\n")
        out = G.DeviceBuffer((B, 2, limbs, n))
        bt = G.Batch(ctx)
        for b in range(B):
            bt.add(G.OP_MULT_RELIN, limbs, lhs.offset(b * 2 * P), rhs.offset(b * 2 * P), k, out.offset(b * 2 * P))
        bt.compile(tune=True)
        assert bt.report()["groups"][0]["load_schedule"] == pf
        got = out.numpy()
        for b in range(B):
            assert np.array_equal(got[b], want[b]), (ds, pf, b)


@pytest.mark.parametrize("logn,limbs,B", [(12, 4, 1), (13, 6, 5), (16, 5, 3), (14, 5, 9), (13, 9, 17)])
def test_exact_mac_shapes_bit_identical(logn, limbs, B):
    """Every launch shape of the exact-convolution keyswitch's matrix-core multiply-accumulate (1 / 2 / 4
    slots per wave), forced through the tuning table, gives the digit-NTT result: one or two tiles of 16
    ciphertexts, key columns filling 12 / 14 / 16 / 22 of the 16-column tiles, 1 to 2 digit blocks."""
    n = 1 << logn
    p = PR.config_batched(n=1 << 16, limbs=limbs, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    rs = np.random.default_rng(300 + logn)
    D = ctx.num_decomp_digits(limbs)
    E = limbs + 2
    key = np.stack([uniform(p.primes + p.aux_primes, n, rs, 2) for _ in range(D)])
    ctx.set_keyswitch_method(G.KS_DIGIT_NTT)
    kc = ctx.ksk_create(key, limbs - 1, 2)
    ctx.set_keyswitch_method(G.KS_EXACT_CONVOLUTION)
    ke = ctx.ksk_create(key, limbs - 1, 2)
    assert ctx.ksk_method(ke) == G.KS_EXACT_CONVOLUTION
    P = limbs * n
    lhs = G.DeviceBuffer.from_numpy(uniform(p.primes, n, rs, 2 * B).reshape(B, 2, limbs, n))
    rhs = G.DeviceBuffer.from_numpy(uniform(p.primes, n, rs, 2 * B).reshape(B, 2, limbs, n))
    quad = G.DeviceBuffer((3, limbs, n))
    one = G.DeviceBuffer((2, limbs, n))
    want = []
    for b in range(B):
        ctx.mult(lhs.offset(b * 2 * P), rhs.offset(b * 2 * P), limbs, 40, 40, quad)
        ctx.relinearize(quad, limbs, 80, kc, one)
        want.append(one.numpy())
    for spw in (1, 2, 4):
        G.tune_table_import(f"{logn} {limbs} {E} {D} {B} {E} 1 {spw} 7 16 1.0\n")
        out = G.DeviceBuffer((B, 2, limbs, n))
        batch = G.Batch(ctx)
        for b in range(B):
            batch.add(G.OP_MULT_RELIN, limbs, lhs.offset(b * 2 * P), rhs.offset(b * 2 * P), ke,
                      out.offset(b * 2 * P))
        batch.compile(tune=True)
        grp = batch.report()["groups"][0]
        assert (grp["mac_rows"], grp["load_schedule"]) == (spw, 7)
        got = out.numpy()
        for b in range(B):
            assert np.array_equal(got[b], want[b]), (spw, b)
