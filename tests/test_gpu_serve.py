"""Request server: Paterson-Stockmeyer requests on two ring degrees at uneven levels, submitted
alone and in a mix (all at once and staggered, under a tight working-set budget): every result is
bit-identical to the request evaluated alone with rck_ps_eval, multiplications of different
requests share launches, and admission keeps the working set within the budget."""
import time

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


@pytest.fixture(scope="module")
def rings():
    rs = np.random.default_rng(77)
    out = []
    for n in (1 << 12, 1 << 13):
        L = 8
        p = PR.config_batched(n=1 << 15, limbs=L, aux=2)
        ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes,
                        aux_roots=p.aux_roots)
        keys = {}
        for lv in range(1, L):
            D = ctx.num_decomp_digits(lv + 1)
            keys[lv] = ctx.ksk_create(np.stack([uniform(p.primes[: lv + 1] + p.aux_primes, n, rs, 2)
                                                for _ in range(D)]), lv, 2)
        out.append((ctx, p, keys))
    return out


def make_requests(rings, count, seed):
    rs = np.random.default_rng(seed)
    reqs = []
    for _ in range(count):
        ctx, p, keys = rings[int(rs.integers(0, len(rings)))]
        limbs = int(rs.integers(5, 9))
        while True:
            deg = int(rs.integers(1, 16))
            try:
                G.ps_required_levels(limbs - 1, deg + 1)
                break
            except G.RckError:
                continue
        coeffs = rs.uniform(-1, 1, deg + 1)
        ct = G.DeviceBuffer.from_numpy(uniform(p.primes[:limbs], ctx.n, rs, 2))
        reqs.append({"ctx": ctx, "keys": keys, "limbs": limbs, "coeffs": coeffs, "ct": ct})
    return reqs


def alone(r):
    out = G.DeviceBuffer((2, r["limbs"], r["ctx"].n))
    ol, osc = r["ctx"].ps_eval(r["ct"], r["limbs"], 40, r["coeffs"], r["keys"], out)
    return out.numpy().reshape(-1)[: 2 * ol * r["ctx"].n], ol, osc


def serve(rings, reqs, budget, stagger=0.0, lanes=4):
    srv = G.Server(budget, lanes=lanes)
    for ctx, _, keys in rings:
        srv.register_ring(ctx, keys)
    ids, outs = [], []
    for r in reqs:
        out = G.DeviceBuffer((2, r["limbs"], r["ctx"].n))
        outs.append(out)
        ids.append(srv.submit(r["ctx"], r["ct"], r["limbs"], 40, r["coeffs"], out))
        if stagger:
            time.sleep(stagger)
    srv.drain(120)
    res = []
    for rid, out, r in zip(ids, outs, reqs):
        w = srv.wait(rid, 1.0)
        assert w is not None
        res.append((out.numpy().reshape(-1)[: 2 * w["limbs"] * r["ctx"].n], w["limbs"], w["scale_bits"]))
    st = srv.stats()
    srv.close()
    return res, st


def test_alone_equals_in_mix(rings):
    reqs = make_requests(rings, 12, 1)
    want = [alone(r) for r in reqs]
    # one request alone in the server
    got, st = serve(rings, reqs[:1], 1 << 34)
    assert np.array_equal(got[0][0], want[0][0]) and got[0][1:] == want[0][1:]
    # all at once: multiplications of different requests share launches (3 concurrent lanes)
    got, st = serve(rings, reqs, 1 << 34, lanes=3)
    for i, (g, w) in enumerate(zip(got, want)):
        assert g[1:] == w[1:], f"request {i} metadata"
        assert np.array_equal(g[0], w[0]), f"request {i}"
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert max(int(k) for k in st["batch_hist"]) >= 2
    assert st["mults"] == sum(r["ctx"].ps_plan(r["limbs"], 40, r["coeffs"])["mults"] for r in reqs)
    assert st["p99_ms"] >= st["p50_ms"] > 0


def test_staggered_arrivals_under_tight_budget(rings):
    reqs = make_requests(rings, 10, 2)
    want = [alone(r) for r in reqs]
    fps = [r["ctx"].ps_plan(r["limbs"], 40, r["coeffs"])["footprint"] for r in reqs]
    budget = 2 * max(fps)  # at most a few requests resident at a time
    got, st = serve(rings, reqs, budget, stagger=0.002, lanes=1)
    for i, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g[0], w[0]) and g[1:] == w[1:], f"request {i}"
    assert st["peak_bytes"] <= budget
    assert st["completed"] == len(reqs)


def test_server_errors(rings):
    ctx, p, keys = rings[0]
    other = G.Context(poly_degree=1 << 12, primes=p.primes, roots=p.roots)
    srv = G.Server(1 << 30)
    srv.register_ring(ctx, keys)
    ct = G.DeviceBuffer((2, 8, ctx.n))
    out = G.DeviceBuffer((2, 8, ctx.n))
    with pytest.raises(G.RckError, match="not registered"):
        srv.submit(other, ct, 8, 40, [1.0, 2.0], out)
    with pytest.raises(G.RckError, match="polynomial degree is too large"):
        srv.submit(ctx, ct, 8, 40, np.ones(40), out)
    small = G.Server(1024)
    small.register_ring(ctx, keys)
    with pytest.raises(G.RckError, match="exceeds the server budget"):
        small.submit(ctx, ct, 8, 40, [1.0, 2.0, 3.0], out)
    srv.close()
    small.close()
