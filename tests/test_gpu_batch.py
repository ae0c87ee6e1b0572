"""Batched executor: a captured hipGraph over many ciphertexts at uneven levels with mixed
operations gives exactly the per-op results, on every replay (with eager work in between)."""
import os

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


def test_batched_mult_relin_matches_per_op_and_oracle():
    """B=4 mult+relin at one level sharing one hybrid key, contiguous (strided) operands."""
    n, L, B = 1 << 13, 8, 4
    p = PR.realistic_8_level()
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots,
                    scale_bits=45)
    qb = O.Basis(n, p.primes, p.roots)
    eb = O.Basis(n, p.primes + p.aux_primes, p.roots + p.aux_roots)
    rs = np.random.default_rng(5)
    D = ctx.num_decomp_digits(L)
    key = np.stack([uniform(eb.moduli, n, rs, 2) for _ in range(D)])
    k = ctx.ksk_create(key, L - 1, 2)
    lhs = uniform(p.primes, n, rs, 2 * B).reshape(B, 2, L, n)
    rhs = uniform(p.primes, n, rs, 2 * B).reshape(B, 2, L, n)
    dl, dr = dev(lhs), dev(rhs)
    out = G.DeviceBuffer((B, 2, L, n))
    bt = G.Batch(ctx)
    P = L * n
    for b in range(B):
        bt.add(G.OP_MULT_RELIN, L, dl.offset(b * 2 * P), dr.offset(b * 2 * P), k, out.offset(b * 2 * P))
    bt.compile(tune=True)
    got = out.numpy()
    one = G.DeviceBuffer((2, L, n))
    quad = G.DeviceBuffer((3, L, n))
    for b in range(B):
        ctx.mult(dl.offset(b * 2 * P), dr.offset(b * 2 * P), L, 45, 45, quad)
        ctx.relinearize(quad, L, 90, k, one)
        assert np.array_equal(one.numpy(), got[b]), f"ciphertext {b}"
    want0 = O.relinearize(qb, eb, O.multiply_ciphertexts(qb, lhs[0], rhs[0]), key)
    assert np.array_equal(got[0], want0)
    # replays: clear the outputs between launches with eager work, every replay must match
    for _ in range(3):
        out.upload(np.zeros((B, 2, L, n), dtype=np.uint64))
        ctx.mult(dl, dr, L, 45, 45, quad)  # eager work on the context between replays
        bt.run()
        bt.synchronize()
        assert np.array_equal(out.numpy(), got)
    rep = bt.report()
    assert rep["ops"] == B and len(rep["groups"]) == 1 and rep["groups"][0]["ops"] == B


def test_batched_mixed_levels_and_ops_match_per_op():
    """Config-3 shape in miniature: ciphertexts at levels 4..8 limbs (primes just above 2^50),
    40% mult+relin, 40% rotation by +-2^k, 20% ct x pt, scattered (staged) operands."""
    n = 1 << 12
    p = PR.config_batched(n=1 << 15, limbs=8, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    rs = np.random.default_rng(17)
    eprimes = p.primes + p.aux_primes
    keys = {}

    def key_for(limbs, galois):
        kk = (limbs, galois)
        if kk not in keys:
            D = ctx.num_decomp_digits(limbs)
            moduli = p.primes[:limbs] + p.aux_primes
            kd = np.stack([uniform(moduli, n, rs, 2) for _ in range(D)])
            keys[kk] = ctx.ksk_create(kd, limbs - 1, 2, galois)
        return keys[kk]

    bt = G.Batch(ctx)
    ops = []
    mix = rs.permutation(["mr"] * 8 + ["rot"] * 8 + ["pt"] * 4)
    bufs = []
    for t, kind in enumerate(mix):
        limbs = int(rs.integers(4, 9))
        a = dev(uniform(p.primes[:limbs], n, rs, 2))
        out = G.DeviceBuffer((2, limbs, n))
        if kind == "mr":
            b = dev(uniform(p.primes[:limbs], n, rs, 2))
            k = key_for(limbs, 0)
            bt.add(G.OP_MULT_RELIN, limbs, a, b, k, out)
        elif kind == "rot":
            step = int(rs.choice([1, 2, 4, 8, 1024])) * int(rs.choice([-1, 1]))
            gel = G.galois_element_for_rotation(n, step)
            b, k = None, key_for(limbs, gel)
            bt.add(G.OP_ROTATE, limbs, a, None, k, out)
        else:
            b, k = dev(uniform(p.primes[:limbs], n, rs, 1)[0]), None
            bt.add(G.OP_MULT_PLAIN, limbs, a, b, None, out)
        ops.append((kind, limbs, a, b, k, out))
        bufs.append((a, b, out))
    bt.compile(tune=True)
    first = [o[5].numpy() for o in ops]
    for rep in range(2):
        for o in ops:
            o[5].upload(np.zeros(o[5].shape, dtype=np.uint64))
        bt.run()
        bt.synchronize()
        for i, o in enumerate(ops):
            assert np.array_equal(o[5].numpy(), first[i]), f"replay {rep} op {i}"
    for i, (kind, limbs, a, b, k, out) in enumerate(ops):
        ref = G.DeviceBuffer((2, limbs, n))
        if kind == "mr":
            quad = G.DeviceBuffer((3, limbs, n))
            ctx.mult(a, b, limbs, 40, 40, quad)
            ctx.relinearize(quad, limbs, 80, k, ref)
        elif kind == "rot":
            ctx.rotate(a, limbs, 40, k, ref)
        else:
            ctx.mult_plain(a, b, limbs, 40, ref)
        assert np.array_equal(ref.numpy(), first[i]), f"op {i} ({kind}, {limbs} limbs)"


def test_batch_rejects_bad_ops():
    n = 1 << 12
    p = PR.config_batched(n=1 << 15, limbs=4, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    bt = G.Batch(ctx)
    with pytest.raises(G.RckError, match="batch has no operations"):
        bt.compile()
    a = G.DeviceBuffer((2, 4, n))
    with pytest.raises(G.RckError, match="unknown batch op kind"):
        bt.add(9, 4, a, a, None, a)
    rs = np.random.default_rng(1)
    D = ctx.num_decomp_digits(3)
    k = ctx.ksk_create(np.stack([uniform(p.primes[:3] + p.aux_primes, n, rs, 2) for _ in range(D)]), 2, 2, 5)
    bt.add(G.OP_ROTATE, 4, a, None, k, G.DeviceBuffer((2, 4, n)))
    with pytest.raises(G.RckError, match="rotation key level must match the ciphertext level"):
        bt.compile()
    # a failed compile leaves no groups behind, and a retry fails the same way with no duplicates
    assert bt.report()["groups"] == []
    with pytest.raises(G.RckError, match="rotation key level must match the ciphertext level"):
        bt.compile()
    assert bt.report()["groups"] == []


@pytest.mark.parametrize("tune", [False, True])
def test_batch_compile_failing_in_the_run_leaves_no_groups(tune):
    """A valid group and a mult+relin whose key has another level: the compile fails inside the eager
    run (or the tuner), and every retry starts from no groups (each op in exactly one group)."""
    n = 1 << 12
    p = PR.config_batched(n=1 << 15, limbs=4, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    rs = np.random.default_rng(3)
    D = ctx.num_decomp_digits(3)
    k3 = ctx.ksk_create(np.stack([uniform(p.primes[:3] + p.aux_primes, n, rs, 2) for _ in range(D)]), 2, 2)
    bt = G.Batch(ctx)
    a = dev(uniform(p.primes[:3], n, rs, 2))
    b = dev(uniform(p.primes[:3], n, rs, 2))
    bt.add(G.OP_MULT_RELIN, 3, a, b, k3, G.DeviceBuffer((2, 3, n)))
    a4 = dev(uniform(p.primes[:4], n, rs, 2))
    bt.add(G.OP_MULT_RELIN, 4, a4, a4, k3, G.DeviceBuffer((2, 4, n)))
    for _ in range(2):
        with pytest.raises(G.RckError, match="keyswitch key level must match"):
            bt.compile(tune=tune)
        rep = bt.report()
        assert rep["ops"] == 2 and rep["groups"] == []


def test_headline_workload_batched_matches_per_op_and_oracle():
    """The headline workload itself: 16 mult+relinearize at N=2^16, L=24 primes below 2^50, K=2,
    D=40, sharing one hybrid key, through the batch executor with both streamed MAC shapes (the
    relinearization terms folded into the accumulators, inverse rows in the MAC epilogue): every
    ciphertext bit-exact against the per-op calls, the first against the oracle."""
    p = PR.config_per_op_gpu()
    n, L, K, B = p.poly_degree, len(p.primes), len(p.aux_primes), 16
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    rs = np.random.default_rng(2024)
    D = ctx.num_decomp_digits(L)
    assert D == 40
    key = np.stack([uniform(p.primes + p.aux_primes, n, rs, 2) for _ in range(D)])
    k = ctx.ksk_create(key, L - 1, K)  # both forms (the default)
    ctx.set_keyswitch_method(G.KS_DIGIT_NTT)
    kc = ctx.ksk_create(key, L - 1, K)
    ctx.set_keyswitch_method(G.KS_EXACT_CONVOLUTION)
    ke = ctx.ksk_create(key, L - 1, K)
    P = L * n
    lhs_h = uniform(p.primes, n, rs, 2 * B).reshape(B, 2, L, n)
    rhs_h = uniform(p.primes, n, rs, 2 * B).reshape(B, 2, L, n)
    lhs, rhs = dev(lhs_h), dev(rhs_h)
    quad = G.DeviceBuffer((3, L, n))
    one = G.DeviceBuffer((2, L, n))
    want = []
    for b in range(B):
        ctx.mult(lhs.offset(b * 2 * P), rhs.offset(b * 2 * P), L, 40, 40, quad)
        ctx.relinearize(quad, L, 80, k, one)
        want.append(one.numpy())
    qb = O.Basis(n, p.primes, p.roots)
    eb = O.Basis(n, p.primes + p.aux_primes, p.roots + p.aux_roots)
    O.set_threads(16)
    assert np.array_equal(want[0], O.relinearize(qb, eb, O.multiply_ciphertexts(qb, lhs_h[0], rhs_h[0]), key))
    for pf in (5, 6, 7):  # the streamed digit-NTT MACs, then the exact-convolution method
        key_used = ke if pf == 7 else kc
        G.tune_table_import(f"16 {L} {L + K} {D} {B} {L + K} 1 {2 if pf == 7 else 4} {pf} 16 1.0\n")
        out = G.DeviceBuffer((B, 2, L, n))
        bt = G.Batch(ctx)
        for b in range(B):
            bt.add(G.OP_MULT_RELIN, L, lhs.offset(b * 2 * P), rhs.offset(b * 2 * P), key_used, out.offset(b * 2 * P))
        bt.compile(tune=True)
        assert bt.report()["groups"][0]["load_schedule"] == pf
        bt.run()
        bt.synchronize()
        got = out.numpy()
        for b in range(B):
            assert np.array_equal(got[b], want[b]), (pf, b)


def test_groups_larger_than_32_match_per_op():
    """Groups of 40 mult+relin, 36 rotations (one Galois element) and 34 mult+relin+rescale at one
    level, scattered operands: the mod-down epilogue takes ciphertexts in chunks of 32, so every
    group spans two chunks."""
    n, limbs = 1 << 12, 3
    p = PR.config_batched(n=1 << 15, limbs=8, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    rs = np.random.default_rng(23)
    moduli = p.primes[:limbs] + p.aux_primes
    D = ctx.num_decomp_digits(limbs)
    gel = G.galois_element_for_rotation(n, -5)
    rkey = np.stack([uniform(moduli, n, rs, 2) for _ in range(D)])
    kr = ctx.ksk_create(np.stack([uniform(moduli, n, rs, 2) for _ in range(D)]), limbs - 1, 2)
    kg = ctx.ksk_create(rkey, limbs - 1, 2, gel)
    bt = G.Batch(ctx)
    ops = []
    for kind, count in (("mr", 40), ("rot", 36), ("mrr", 34)):
        for _ in range(count):
            a = dev(uniform(p.primes[:limbs], n, rs, 2))
            b = dev(uniform(p.primes[:limbs], n, rs, 2)) if kind != "rot" else None
            out = G.DeviceBuffer((2, limbs - 1 if kind == "mrr" else limbs, n))
            op = {"mr": G.OP_MULT_RELIN, "rot": G.OP_ROTATE, "mrr": G.OP_MULT_RELIN_RESCALE}[kind]
            bt.add(op, limbs, a, b, kg if kind == "rot" else kr, out)
            ops.append((kind, a, b, out))
    bt.compile(tune=False)
    rep = bt.report()
    assert sorted(g["ops"] for g in rep["groups"]) == [34, 36, 40]
    got = [o[3].numpy() for o in ops]
    quad = G.DeviceBuffer((3, limbs, n))
    for i, (kind, a, b, out) in enumerate(ops):
        ref = G.DeviceBuffer(out.shape)
        if kind == "rot":
            ctx.rotate(a, limbs, 40, kg, ref)
        elif kind == "mr":
            ctx.mult(a, b, limbs, 40, 40, quad)
            ctx.relinearize(quad, limbs, 80, kr, ref)
        else:
            ctx.mult_relinearized(a, b, limbs, 40, 40, kr, ref)
        assert np.array_equal(ref.numpy(), got[i]), f"op {i} ({kind})"
    # the last rotation of the second chunk against the oracle
    qb = O.Basis(n, p.primes[:limbs], p.roots[:limbs])
    eb = O.Basis(n, moduli, p.roots[:limbs] + p.aux_roots)
    kind, a, _, out = ops[40 + 35]
    assert np.array_equal(got[40 + 35], O.rotate(qb, eb, a.numpy(), gel, rkey))


BATCH_SEEDS = int(os.environ.get("RCK_BATCH_SWEEP_SEEDS", "8"))  # longer: RCK_BATCH_SWEEP_SEEDS=600


@pytest.mark.parametrize("seed", range(BATCH_SEEDS))
def test_seeded_batch_mixes_match_per_op(seed):
    """Drawn batches: ring degree 2^12 or 2^13, 8-30 ops of all four kinds at 2-8 limbs, rotation
    steps from +-1 .. +-N/4, tuned or untuned launch shapes; every output equals the per-op call,
    on the eager run and on two graph replays."""
    rs = np.random.default_rng(500 + seed)
    n = 1 << int(rs.integers(12, 14))
    p = PR.config_batched(n=1 << 15, limbs=8, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    keys = {}

    def key_for(limbs, galois):
        if (limbs, galois) not in keys:
            D = ctx.num_decomp_digits(limbs)
            kd = np.stack([uniform(p.primes[:limbs] + p.aux_primes, n, rs, 2) for _ in range(D)])
            keys[(limbs, galois)] = ctx.ksk_create(kd, limbs - 1, 2, galois)
        return keys[(limbs, galois)]

    bt = G.Batch(ctx)
    ops = []
    for _ in range(int(rs.integers(8, 31))):
        kind = int(rs.integers(1, 5))
        limbs = int(rs.integers(2, 9))
        a = dev(uniform(p.primes[:limbs], n, rs, 2))
        b, k = None, None
        if kind in (G.OP_MULT_RELIN, G.OP_MULT_RELIN_RESCALE):
            b, k = dev(uniform(p.primes[:limbs], n, rs, 2)), key_for(limbs, 0)
        elif kind == G.OP_ROTATE:
            step = int(rs.integers(1, n // 4 + 1)) * (1 if rs.integers(0, 2) else -1)
            k = key_for(limbs, G.galois_element_for_rotation(n, step))
        else:
            b = dev(uniform(p.primes[:limbs], n, rs, 1)[0])
        out = G.DeviceBuffer((2, limbs - 1 if kind == G.OP_MULT_RELIN_RESCALE else limbs, n))
        bt.add(kind, limbs, a, b, k, out)
        ops.append((kind, limbs, a, b, k, out))
    print(f"[batch sweep] seed={seed} n={n} ops={len(ops)}", flush=True)
    bt.compile(tune=bool(seed % 2))
    first = [o[5].numpy() for o in ops]
    for _ in range(2):
        for o in ops:
            o[5].upload(np.zeros(o[5].shape, dtype=np.uint64))
        bt.run()
        bt.synchronize()
        for i, o in enumerate(ops):
            assert np.array_equal(o[5].numpy(), first[i]), f"replay, op {i}"
    quad = {}
    for i, (kind, limbs, a, b, k, out) in enumerate(ops):
        ref = G.DeviceBuffer(out.shape)
        if kind == G.OP_MULT_RELIN:
            q3 = quad.setdefault(limbs, G.DeviceBuffer((3, limbs, n)))
            ctx.mult(a, b, limbs, 40, 40, q3)
            ctx.relinearize(q3, limbs, 80, k, ref)
        elif kind == G.OP_MULT_RELIN_RESCALE:
            ctx.mult_relinearized(a, b, limbs, 40, 40, k, ref)
        elif kind == G.OP_ROTATE:
            ctx.rotate(a, limbs, 40, k, ref)
        else:
            ctx.mult_plain(a, b, limbs, 40, ref)
        assert np.array_equal(ref.numpy(), first[i]), f"op {i} (kind {kind}, {limbs} limbs)"
