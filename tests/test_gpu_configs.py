"""The three serving-style workloads at their full sizes, on the GPU:

* mixed batch: 256 ciphertexts at N=2^15, each at 4..20 limbs, 40% mult+relin / 40% rotation by
  +-2^k (k <= 13) / 20% ct x pt, one autotuned hipGraph -- every op equals its per-op call on the eager
  run and on two replays, and one op of each kind at 20 limbs equals the oracle;
* batched BSGS matvec: a 4096 x 4096 matrix at N=2^14, 12 limbs, 32 vectors -- the batch equals the
  vectors run one by one bit for bit, and every vector decodes to the FP64 product within 1e-4;
* request server: a Poisson stream of Paterson-Stockmeyer requests (degree 7..63, N in {2^15, 2^16},
  12..24 starting limbs) -- every result equals rck_ps_eval alone, decodes within the bound set by the
  crate's coefficient quantisation, and a degree-7 request at N=2^15 equals the oracle's restatement.

Reference algorithms: context/ops.rs:20-133, context/rotation.rs:104-127 (ops),
ckks/bootstrapping/linear_transform.rs:111-150 (BSGS), algorithms/paterson_stockmeyer.rs:28-108 (PS).
"""
import time

import numpy as np
import pytest

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


def say(msg):
    print(f"[configs] {msg}", flush=True)


def test_mixed_batch_256_ops_n32768_matches_per_op_and_oracle():
    n, top = 1 << 15, 20
    p = PR.config_batched(n=n, limbs=top, aux=2)
    ctx = G.Context(p)
    rs = np.random.default_rng(2026)
    rng = G.Rng(31)
    s = ctx.sample_ternary(rng, top)
    kinds = list(rs.permutation(["mr"] * 102 + ["rot"] * 102 + ["pt"] * 52))
    ops = []
    for i, kind in enumerate(kinds):
        limbs = int(rs.integers(4, top + 1))
        step = int(rs.choice([1 << k for k in range(14)])) * int(rs.choice([-1, 1])) if kind == "rot" else 0
        ops.append([kind, limbs, step])
    # one op of each kind at the top level (the oracle checks those)
    for kind in ("mr", "rot", "pt"):
        ops[kinds.index(kind)][1] = top
    keys = {}
    t0 = time.perf_counter()
    for kind, limbs, step in ops:
        if kind != "pt" and (limbs, step) not in keys:
            keys[(limbs, step)] = (ctx.keygen_rotation(rng, s, step, limbs - 1, sigma=3.2) if step
                                   else ctx.keygen_relin(rng, s, limbs - 1, sigma=3.2))
    say(f"mixed: {len(ops)} ops, {len(keys)} device keys in {time.perf_counter() - t0:.1f} s")
    bt = G.Batch(ctx)
    bufs = []
    for kind, limbs, step in ops:
        a = G.DeviceBuffer.from_numpy(uniform(p.primes[:limbs], n, rs, 2))
        out = G.DeviceBuffer((2, limbs, n))
        if kind == "mr":
            b = G.DeviceBuffer.from_numpy(uniform(p.primes[:limbs], n, rs, 2))
            bt.add(G.OP_MULT_RELIN, limbs, a, b, keys[(limbs, 0)], out)
        elif kind == "rot":
            b = None
            bt.add(G.OP_ROTATE, limbs, a, None, keys[(limbs, step)], out)
        else:
            b = G.DeviceBuffer.from_numpy(uniform(p.primes[:limbs], n, rs, 1)[0])
            bt.add(G.OP_MULT_PLAIN, limbs, a, b, None, out)
        bufs.append((kind, limbs, step, a, b, out))
    t0 = time.perf_counter()
    bt.compile(tune=True)
    rep = bt.report()
    say(f"mixed: compiled and tuned {len(rep['groups'])} groups in {time.perf_counter() - t0:.1f} s")
    assert rep["ops"] == len(ops) and sum(g["ops"] for g in rep["groups"]) == len(ops)
    first = [o[5].numpy() for o in bufs]
    quad = {}

    def per_op(i):
        kind, limbs, step, a, b, _ = bufs[i]
        ref = G.DeviceBuffer((2, limbs, n))
        if kind == "mr":
            q3 = quad.setdefault(limbs, G.DeviceBuffer((3, limbs, n)))
            ctx.mult(a, b, limbs, 40, 40, q3)
            ctx.relinearize(q3, limbs, 80, keys[(limbs, 0)], ref)
        elif kind == "rot":
            ctx.rotate(a, limbs, 40, keys[(limbs, step)], ref)
        else:
            ctx.mult_plain(a, b, limbs, 40, ref)
        return ref.numpy()

    for rep_i in range(2):
        for o in bufs:
            o[5].upload(np.zeros(o[5].shape, dtype=np.uint64))
        per_op(rep_i)  # eager work on the context between replays
        bt.run()
        bt.synchronize()
        for i, o in enumerate(bufs):
            assert np.array_equal(o[5].numpy(), first[i]), f"replay {rep_i}, op {i} ({o[0]}, {o[1]} limbs)"
    say("mixed: two replays equal the eager run")
    for i in range(len(bufs)):
        assert np.array_equal(per_op(i), first[i]), f"op {i} ({bufs[i][0]}, {bufs[i][1]} limbs)"
    say("mixed: every op equals its per-op call")
    O.set_threads(16)
    qb = O.Basis(n, p.primes, p.roots)
    eb = O.Basis(n, p.primes + p.aux_primes, p.roots + p.aux_roots)
    for kind in ("mr", "rot", "pt"):
        i = kinds.index(kind)
        _, limbs, step, a, b, _ = bufs[i]
        assert limbs == top
        x = a.numpy().reshape(2, top, n)
        if kind == "mr":
            key = ctx.ksk_export(keys[(top, 0)])
            want = O.relinearize(qb, eb, O.multiply_ciphertexts(qb, x, b.numpy().reshape(2, top, n)), key)
        elif kind == "rot":
            key = ctx.ksk_export(keys[(top, step)])
            want = O.rotate(qb, eb, x, G.galois_element_for_rotation(n, step), key)
        else:
            want = O.multiply_plain(qb, x, b.numpy().reshape(top, n))
        assert np.array_equal(first[i], want), f"oracle, {kind}"
        say(f"mixed: {kind} at {top} limbs equals the oracle")


def test_bsgs_4096_n16384_l12_batch_of_32():
    n, L, dim, B = 1 << 14, 12, 4096, 32
    p = PR.config_batched(n=n, limbs=L, aux=2)
    ctx = G.Context(p)
    rs = np.random.default_rng(4096)
    rng = G.Rng(41)
    s = ctx.sample_sparse_ternary(rng, 64, L)
    pk = ctx.keygen_public(rng, s, L, sigma=3.2)
    steps = G.bsgs_rotation_steps(dim)
    t0 = time.perf_counter()
    keys = {st: ctx.keygen_rotation(rng, s, st, L - 1, sigma=3.2) for st in steps}
    matrix = rs.uniform(-1, 1, (dim, dim))
    diag = G.Diag(ctx, matrix, L - 1, 40)
    assert (diag.baby_step, diag.giant_steps) == (64, 64)
    say(f"bsgs: {len(keys)} rotation keys and {dim} diagonals in {time.perf_counter() - t0:.1f} s")
    vecs = rs.uniform(-1, 1, (B, dim))
    CT = 2 * L * n * 8
    cts = G.DeviceBuffer((B, 2, L, n))
    for v in range(B):
        ct = ctx.encrypt(rng, ctx.encode(np.tile(vecs[v], n // 2 // dim), L, 40), pk, L, L, sigma=3.2)
        G.lib().rck_memcpy_dtod(cts.ptr + v * CT, ct.ptr, CT, None)
    out = G.DeviceBuffer((B, 2, L - 1, n))
    sc = ctx.bsgs_matvec(cts, B, L, 40, diag, keys, out)
    got = out.numpy()
    say("bsgs: batch of 32 done")
    one = G.DeviceBuffer((1, 2, L - 1, n))
    for v in range(B):
        sc1 = ctx.bsgs_matvec(cts.offset(v * 2 * L * n), 1, L, 40, diag, keys, one)
        assert sc1 == sc
        assert np.array_equal(one.numpy()[0], got[v]), f"vector {v}: batched != one by one"
    say("bsgs: batched == one by one for all 32 vectors")
    P = 2 * (L - 1) * n * 8
    worst = 0.0
    for v in range(B):
        dec = ctx.decrypt(out.ptr + v * P, L - 1, s)
        vals = ctx.decode(dec, L - 1, sc).real
        want = np.tile(matrix @ vecs[v], n // 2 // dim)
        err = float(np.max(np.abs(vals - want)))
        worst = max(worst, err)
        assert err <= 1e-4, f"vector {v}: decode error {err}"
    say(f"bsgs: decode error max {worst:.2e}")


def _quantised(coeffs, block, coeff_bits):
    """The polynomial the crate's PS evaluates: coefficients at a block offset > 0 are encoded at
    scale 2^coeff_bits (round to nearest), offset-0 coefficients at the 2^50 working scale."""
    q = np.array(coeffs, dtype=float)
    for i in range(len(q)):
        if i % block:
            q[i] = np.round(q[i] * 2.0**coeff_bits) / 2.0**coeff_bits
    return q


def test_ps_serving_poisson_stream_full_rings():
    rs = np.random.default_rng(5151)
    count, qps = 24, 200.0
    plans = []
    for _ in range(count):
        n = int(rs.choice([1 << 15, 1 << 16]))
        limbs = int(rs.integers(12, 25))
        deg = int(rs.integers(7, 64))
        while True:  # cap the degree by the level budget, keeping 2 limbs for the result
            try:
                G.ps_required_levels(limbs - 2, deg + 1)
                break
            except G.RckError:
                deg -= 1
        plans.append((n, limbs, rs.uniform(-1, 1, deg + 1)))
    # one degree-7 request at N=2^15 and 12 limbs (checked against the oracle)
    plans[0] = (1 << 15, 12, rs.uniform(-1, 1, 8))
    need = {}
    for n, limbs, coeffs in plans:
        need.setdefault(n, set()).update(G.ps_required_levels(limbs - 1, len(coeffs)))
    rings = {}
    t0 = time.perf_counter()
    for n in sorted(need):
        p = PR.config_batched(n=n, limbs=24, aux=2)
        ctx = G.Context(p)
        rng = G.Rng(n + 7)
        s = ctx.sample_sparse_ternary(rng, 64, 24)
        pk = ctx.keygen_public(rng, s, 24, sigma=3.2)
        keys = {lv: ctx.keygen_relin(rng, s, lv, sigma=3.2) for lv in sorted(need[n])}
        rings[n] = (p, ctx, rng, s, pk, keys)
    say(f"serve: rings {sorted(rings)} with {sum(len(r[5]) for r in rings.values())} relin keys in "
        f"{time.perf_counter() - t0:.1f} s")
    reqs = []
    for n, limbs, coeffs in plans:
        p, ctx, rng, s, pk, keys = rings[n]
        x = rs.uniform(-1, 1, n // 2)
        ct = ctx.encrypt(rng, ctx.encode(x, limbs, 40), pk, 24, limbs, sigma=3.2)
        reqs.append({"n": n, "limbs": limbs, "coeffs": coeffs, "x": x, "ct": ct, "out": G.DeviceBuffer((2, limbs, n))})
    srv = G.Server(16 << 30, lanes=4)
    for n, r in rings.items():
        srv.register_ring(r[1], r[5])
    gaps = rs.exponential(1.0 / qps, count)
    due = time.perf_counter()
    ids = []
    for r, gap in zip(reqs, gaps):
        due += gap
        delay = due - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(rings[r["n"]][1], r["ct"], r["limbs"], 40, r["coeffs"], r["out"]))
    srv.drain(300)
    st = srv.stats()
    srv.close()
    say(f"serve: {st['completed']} requests, p50 {st['p50_ms']:.1f} ms, p99 {st['p99_ms']:.1f} ms, "
        f"{st['mult_launches']} launches for {st['mults']} mults")
    assert st["completed"] == count and st["failed"] == 0
    assert st["p99_ms"] >= st["p50_ms"] > 0
    worst = worst_q = 0.0
    for i, (r, rid) in enumerate(zip(reqs, ids)):
        p, ctx, rng, s, pk, keys = rings[r["n"]]
        n = r["n"]
        one = G.DeviceBuffer((2, r["limbs"], n))
        ol, osc = ctx.ps_eval(r["ct"], r["limbs"], 40, r["coeffs"], keys, one)
        plan = ctx.ps_plan(r["limbs"], 40, r["coeffs"])
        assert (plan["limbs"], plan["scale_bits"]) == (ol, osc)
        got = r["out"].numpy().reshape(-1)[: 2 * ol * n]
        assert np.array_equal(got, one.numpy().reshape(-1)[: 2 * ol * n]), f"request {i}: mix != alone"
        dec = ctx.decode(ctx.decrypt(r["out"], ol, s), ol, osc).real
        block = G.ps_block_size(len(r["coeffs"]))
        quant = _quantised(r["coeffs"], block, 10)
        err_q = float(np.max(np.abs(dec - np.polyval(quant[::-1], r["x"]))))
        err = float(np.max(np.abs(dec - np.polyval(np.array(r["coeffs"])[::-1], r["x"]))))
        worst, worst_q = max(worst, err), max(worst_q, err_q)
        # against the quantised polynomial only the CKKS noise remains: the baby powers are rescaled down
        # to scale 2^30 before their lift, where the rescale rounding (about 2^11 with h = 64) is ~2e-6
        # of the value, summed over up to 63 terms (measured up to 1.2e-4); against the exact polynomial
        # the quantisation of at most deg coefficients by 2^-11 each (|x| <= 1) comes on top
        assert err_q <= 1e-3, f"request {i}: {err_q} vs the quantised polynomial"
        assert err <= 1e-3 + (len(r["coeffs"]) - 1) * 2.0**-11, f"request {i}: {err} vs FP64"
    say(f"serve: every result equals ps_eval alone; decode error vs FP64 max {worst:.2e}, vs the quantised "
        f"polynomial max {worst_q:.2e}")
    # the degree-7 request against the oracle's restatement of the crate's algorithm
    r = reqs[0]
    p, ctx, rng, s, pk, keys = rings[r["n"]]
    lv_needed = OPS.required_levels(r["limbs"] - 1, len(r["coeffs"]))
    O.set_threads(16)
    R = OPS.Ring(r["n"], p.primes, p.roots, p.aux_primes, p.aux_roots,
                 {lv: ctx.ksk_export(keys[lv]) for lv in lv_needed})
    want, wscale = OPS.ps_eval(R, (r["ct"].numpy().reshape(2, r["limbs"], r["n"]), 40), list(r["coeffs"]))
    ol = want.shape[1]
    got = r["out"].numpy().reshape(-1)[: 2 * ol * r["n"]].reshape(2, ol, r["n"])
    assert np.array_equal(got, want)
    say("serve: degree-7 request at N=2^15 equals the oracle")
