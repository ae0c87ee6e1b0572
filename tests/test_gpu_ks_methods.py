"""The two keyswitch methods produce the crate's keyswitch bit for bit.

* digit NTT: every digit transformed under every eval prime, key kept in NTT form mod each prime;
* exact convolution (hybrid keys at N >= 2^12): digits and centred key limbs
  transformed under two auxiliary primes s1, s2, the products summed as exact integers modulo s1 s2
  in groups of digits small enough that no sum can wrap, then recombined and reduced mod each prime.
  When typical digits allow, that is one group, with the digits' L1 norms checked on the device and
  a batch whose norms could wrap a sum recomputed with the worst-case groups.

Keys created with the automatic method (the default) hold both forms and every keyswitch takes one of
them by batch size; both choices give the same bits.

Both are checked against each other and the oracle (keyswitch_rns_hybrid, keyswitching.rs:159-194),
including inputs built to drive the exact sums to their largest possible magnitude: every digit at
-2^29 and every key coefficient at (p_j - 1)/2 on the largest primes the context accepts, with the
group size at its limit (one group), split in two, and beyond the one-group bound inside a batch of
uniform ciphertexts."""
import numpy as np
import pytest

from oracle import oracle as O
from rnsckks import capi as G
from rnsckks import params as PR

pytestmark = pytest.mark.gpu


def uniform(moduli, n, rs, polys):
    out = np.empty((polys, len(moduli), n), dtype=np.uint64)
    for l, q in enumerate(moduli):
        out[:, l] = rs.integers(0, q, size=(polys, n), dtype=np.uint64)
    return out


def dev(a):
    return G.DeviceBuffer.from_numpy(np.ascontiguousarray(a, dtype=np.uint64))


def both_keys(ctx, key, level, aux, galois=0):
    ctx.set_keyswitch_method(G.KS_DIGIT_NTT)
    kc = ctx.ksk_create(key, level, aux, galois)
    ctx.set_keyswitch_method(G.KS_EXACT_CONVOLUTION)
    ke = ctx.ksk_create(key, level, aux, galois)
    assert (ctx.ksk_method(kc), ctx.ksk_method(ke)) == (G.KS_DIGIT_NTT, G.KS_EXACT_CONVOLUTION)
    return kc, ke


@pytest.mark.parametrize("logn,limbs,K", [(12, 4, 2), (13, 8, 1), (15, 6, 3), (14, 10, 4)])
def test_methods_agree_with_each_other_and_the_oracle(logn, limbs, K):
    n = 1 << logn
    p = PR.config_batched(n=1 << 16, limbs=limbs, aux=K)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    qb = O.Basis(n, p.primes, p.roots)
    eb = O.Basis(n, p.primes + p.aux_primes, p.roots + p.aux_roots)
    rs = np.random.default_rng(70 + logn)
    D = ctx.num_decomp_digits(limbs)
    key = np.stack([uniform(eb.moduli, n, rs, 2) for _ in range(D)])
    gel = G.galois_element_for_rotation(n, 5)
    kc, ke = both_keys(ctx, key, limbs - 1, K)
    rc, re = both_keys(ctx, key, limbs - 1, K, gel)
    # the exact form converts back to the crate's coefficient form
    assert np.array_equal(ctx.ksk_export(ke), key)
    assert np.array_equal(ctx.ksk_export(re), key)
    quad = uniform(qb.moduli, n, rs, 3)
    ct = uniform(qb.moduli, n, rs, 2)
    lhs, rhs = uniform(qb.moduli, n, rs, 2), uniform(qb.moduli, n, rs, 2)
    outs = {}
    for name, k, r in (("digit", kc, rc), ("exact", ke, re)):
        o1, o2, o3, o4 = (G.DeviceBuffer((2, limbs, n)) for _ in range(4))
        ctx.relinearize(dev(quad), limbs, 80, k, o1)
        ctx.keyswitch(dev(quad[2]), limbs, k, o2)
        ctx.rotate(dev(ct), limbs, 40, r, o3)
        ctx.mult_relinearized(dev(lhs), dev(rhs), limbs, 40, 40, k, o4)
        outs[name] = [o.numpy() for o in (o1, o2, o3)] + [o4.numpy().reshape(-1)[: 2 * (limbs - 1) * n]]
    for a, b in zip(outs["digit"], outs["exact"]):
        assert np.array_equal(a, b)
    assert np.array_equal(outs["exact"][0], O.relinearize(qb, eb, quad, key))
    assert np.array_equal(outs["exact"][2], O.rotate(qb, eb, ct, gel, key))


@pytest.mark.parametrize("limbs", [20, 21])
def test_exact_sums_at_their_largest_magnitude(limbs):
    """N = 2^16, the largest primes below 2^50 + 2^46, every digit of every coefficient at -2^29 (the top
    digit 0) and every key coefficient at (p_j - 1)/2: each exact sum reaches D - 1 = 33 (one group of
    the largest size the auxiliary primes allow) or 35 times N 2^29 (p_j - 1)/2.  At 35 the keyswitch
    plans one group (typical digits fit), the decomposition finds these digits beyond the one-group
    bound and the batch is recomputed with two groups of 18; a one-group result would be wrong."""
    n, K = 1 << 16, 2
    step, top = 2 * n, (1 << 50) + (1 << 46)
    q, cand = [], (top - 1) // step * step + 1
    while len(q) < limbs + K + 2:  # the context takes the next two as its auxiliary NTT primes
        if PR.is_prime(cand):
            q.append(cand)
        cand -= step
    q = q[2:]  # leave the two largest to the auxiliary primes, so the eval primes are the next ones
    g = [PR.smallest_primitive_root(x) for x in q]
    ctx = G.Context(poly_degree=n, primes=q[:limbs], roots=g[:limbs], aux_primes=q[limbs:limbs + K],
                    aux_roots=g[limbs:limbs + K])
    qb = O.Basis(n, q[:limbs], g[:limbs])
    eb = O.Basis(n, q[:limbs + K], g[:limbs + K])
    D = ctx.num_decomp_digits(limbs)
    assert D == (34 if limbs == 20 else 36)
    assert ctx.exact_group_size(limbs) == ((34, 1) if limbs == 20 else (36, 1))
    Q = 1
    for x in q[:limbs]:
        Q *= x
    X = -(2 ** 29) * sum(2 ** (30 * i) for i in range(D - 1))
    assert abs(X) < Q // 2
    c2 = np.array([[(X % m)] * n for m in q[:limbs]], dtype=np.uint64)
    sd = O.decompose_signed(qb, c2, D)
    assert (sd[: D - 1] == -(2 ** 29)).all() and (sd[D - 1] == 0).all()
    key = np.empty((D, 2, limbs + K, n), dtype=np.uint64)
    for j, m in enumerate(q[:limbs + K]):
        key[:, :, j, :] = (m - 1) // 2
    rs = np.random.default_rng(9)
    quad = np.concatenate([uniform(qb.moduli, n, rs, 2), c2[None]], axis=0)
    kc, ke = both_keys(ctx, key, limbs - 1, K)
    oc, oe = G.DeviceBuffer((2, limbs, n)), G.DeviceBuffer((2, limbs, n))
    ctx.relinearize(dev(quad), limbs, 80, kc, oc)
    ctx.relinearize(dev(quad), limbs, 80, ke, oe)
    got = oe.numpy()
    assert np.array_equal(got, oc.numpy())
    O.set_threads(16)
    assert np.array_equal(got, O.relinearize(qb, eb, quad, key))


def test_auto_keys_hold_both_forms_and_take_either_method():
    """N = 2^13, 5 limbs: a key created with the automatic method keeps both forms; a single
    mult+relin (digit NTT by default) and an untuned batch of 5 (exact convolution from four up) and a
    tuned batch (both families timed) all equal the oracle's relinearization."""
    n, L, K, B = 1 << 13, 5, 2, 5
    p = PR.config_batched(n=1 << 16, limbs=L, aux=K)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    qb = O.Basis(n, p.primes, p.roots)
    eb = O.Basis(n, p.primes + p.aux_primes, p.roots + p.aux_roots)
    rs = np.random.default_rng(77)
    D = ctx.num_decomp_digits(L)
    key = np.stack([uniform(eb.moduli, n, rs, 2) for _ in range(D)])
    k = ctx.ksk_create(key, L - 1, K)
    assert ctx.ksk_method(k) == G.KS_AUTO
    assert np.array_equal(ctx.ksk_export(k), key)
    P = L * n
    lhs = uniform(qb.moduli, n, rs, 2 * B).reshape(B, 2, L, n)
    rhs = uniform(qb.moduli, n, rs, 2 * B).reshape(B, 2, L, n)
    want = [O.relinearize(qb, eb, O.multiply_ciphertexts(qb, lhs[b], rhs[b]), key) for b in range(B)]
    quad, one = G.DeviceBuffer((3, L, n)), G.DeviceBuffer((2, L, n))
    dl, dr = dev(lhs), dev(rhs)
    for b in range(B):
        ctx.mult(dl.offset(b * 2 * P), dr.offset(b * 2 * P), L, 40, 40, quad)
        ctx.relinearize(quad, L, 80, k, one)
        assert np.array_equal(one.numpy(), want[b]), b
    for tune in (False, True):
        out = G.DeviceBuffer((B, 2, L, n))
        bt = G.Batch(ctx)
        for b in range(B):
            bt.add(G.OP_MULT_RELIN, L, dl.offset(b * 2 * P), dr.offset(b * 2 * P), k, out.offset(b * 2 * P))
        bt.compile(tune=tune)
        got = out.numpy()
        for b in range(B):
            assert np.array_equal(got[b], want[b]), (tune, b)


def test_one_group_with_a_device_checked_recomputation():
    """N = 2^16, 24 primes below 2^50, 40 digits: two groups in the worst case, and the exact
    keyswitch runs one (the digits of uniform ciphertexts leave a margin of about 1.8), recomputing a
    batch with two groups when some ciphertext's digit L1 norm could carry its sums out of range.
    Op 2 of a batch of four has a1 = 1 and b1 = X (every digit 2^29 - 1, the top digit 0) and the key
    every coefficient at (p_j - 1)/2: its one-group sum at coefficient N - 1 is
    39 (2^29 - 1) N (p_j - 1)/2, beyond (s1 s2 - 1)/2 for any auxiliary primes below 2^50 + 2^46,
    so only the recomputation gives the crate's result.  Batched, alone, and mixed with uniform
    ciphertexts, every output equals the digit-NTT method's; op 2 also the oracle's."""
    p = PR.config_per_op_gpu()
    n, L, K = p.poly_degree, len(p.primes), len(p.aux_primes)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    D = ctx.num_decomp_digits(L)
    assert D == 40 and ctx.exact_group_size(L) == (40, 1)
    mods = p.primes + p.aux_primes
    assert 39 * (2 ** 29 - 1) * n * ((min(mods) - 1) // 2) > ((1 << 50) + (1 << 46)) ** 2 // 2
    key = np.empty((D, 2, L + K, n), dtype=np.uint64)
    for j, m in enumerate(mods):
        key[:, :, j, :] = (m - 1) // 2
    Q = 1
    for x in p.primes:
        Q *= x
    X = (2 ** 29 - 1) * sum(2 ** (30 * i) for i in range(D - 1))
    assert X < Q // 2
    qb = O.Basis(n, p.primes, p.roots)
    eb = O.Basis(n, mods, p.roots + p.aux_roots)
    rs = np.random.default_rng(41)
    B, P = 4, L * n
    lhs = uniform(qb.moduli, n, rs, 2 * B).reshape(B, 2, L, n)
    rhs = uniform(qb.moduli, n, rs, 2 * B).reshape(B, 2, L, n)
    lhs[2, 1] = 0
    lhs[2, 1, :, 0] = 1
    rhs[2, 1] = np.array([[X % m] * n for m in p.primes], dtype=np.uint64)
    sd = O.decompose_signed(qb, rhs[2, 1], D)
    assert (sd[: D - 1] == 2 ** 29 - 1).all() and (sd[D - 1] == 0).all()
    kc, ke = both_keys(ctx, key, L - 1, K)
    dl, dr = dev(lhs), dev(rhs)
    want, quad, one = [], G.DeviceBuffer((3, L, n)), G.DeviceBuffer((2, L, n))
    for b in range(B):
        ctx.mult(dl.offset(b * 2 * P), dr.offset(b * 2 * P), L, 40, 40, quad)
        ctx.relinearize(quad, L, 80, kc, one)
        want.append(one.numpy())
        ctx.relinearize(quad, L, 80, ke, one)
        assert np.array_equal(one.numpy(), want[b]), ("alone", b)
    for ops in ([0, 1, 3], [0, 1, 2, 3]):
        out = G.DeviceBuffer((B, 2, L, n))
        bt = G.Batch(ctx)
        for b in ops:
            bt.add(G.OP_MULT_RELIN, L, dl.offset(b * 2 * P), dr.offset(b * 2 * P), ke, out.offset(b * 2 * P))
        bt.compile(tune=False)
        for rep in range(2):
            bt.run()
            bt.synchronize()
            got = out.numpy()
            for b in ops:
                assert np.array_equal(got[b], want[b]), (ops, rep, b)
    O.set_threads(16)
    assert np.array_equal(want[2], O.relinearize(qb, eb, O.multiply_ciphertexts(qb, lhs[2], rhs[2]), key))
