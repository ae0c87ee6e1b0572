"""GPU parity: every C-ABI entry point against the CPU oracle, bit for bit, on seeded inputs.

Small and medium rings run the oracle directly; the full per-op workload (N=2^16, 24 limbs,
2 aux primes, 40 digits) is checked limb by limb for mult and relinearization too.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle import encoding as E
from rnsckks import capi as G
from rnsckks import params as PR

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
PRIMES = json.load(open(os.path.join(HERE, "golden", "primes.json")))


def chain(kind, n, count):
    lst = PRIMES[kind][str(n)][:count]
    return [p for p, _ in lst], [g for _, g in lst]


def rand_poly(moduli, n, rs, polys=None):
    shape = (len(moduli), n) if polys is None else (polys, len(moduli), n)
    out = np.empty(shape, dtype=np.uint64)
    for l, q in enumerate(moduli):
        v = rs.integers(0, q, size=shape[:-2] + (n,), dtype=np.uint64)
        if polys is None:
            out[l] = v
        else:
            out[:, l] = v
    return out


def dev(a):
    return G.DeviceBuffer.from_numpy(np.ascontiguousarray(a, dtype=np.uint64))


def empty(shape):
    return G.DeviceBuffer(shape, np.uint64)


def make(n, q, g, p=(), gp=(), scale=40):
    ctx = G.Context(poly_degree=n, primes=q, roots=g, aux_primes=p, aux_roots=gp, scale_bits=scale)
    return ctx, O.Basis(n, list(q), list(g))


SMALL = ([998244353, 1004535809, 469762049], [3, 3, 3])  # the crate's sample_context (rns/tests.rs:14-24)


# ------------------------------------------------------------------ NTT (src/math/ntt.rs)
def test_ntt_reference_kats():
    ctx, b = make(8, [998244353], [3])
    v = dev(np.arange(1, 9, dtype=np.uint64).reshape(1, 8))
    ctx.ntt_forward(v, 1)
    assert list(v.numpy()[0]) == list(O.ntt_forward(b.plan(0), np.arange(1, 9)))
    ctx.ntt_inverse(v, 1)
    assert list(v.numpy()[0]) == list(range(1, 9))
    ctx, b = make(16, [998244353], [3])
    fib = [0, 1, 2, 3, 5, 8, 13, 21, 34, 55, 89, 144, 233, 377, 610, 987]
    v = dev(np.array([fib], dtype=np.uint64))
    ctx.ntt_forward(v, 1)
    ctx.ntt_inverse(v, 1)
    assert list(v.numpy()[0]) == fib


@pytest.mark.parametrize("logn", [3, 5, 10, 11, 12, 13, 14, 15, 16])
def test_ntt_cyclic_matches_oracle(logn):
    n = 1 << logn
    q, g = chain("below_2^50", 1 << 16, 4) if logn <= 16 else None
    p, gp = chain("above_2^50", 1 << 16, 2)
    ctx, _ = make(n, q, g, p, gp)
    rs = np.random.default_rng(logn)
    x = rand_poly(q + p, n, rs)
    v = dev(x)
    ctx.ntt_forward(v, 4, 2)
    got = v.numpy()
    for l, (m, r) in enumerate(zip(q + p, g + gp)):
        plan = O.ntt_plan(n, m, r)
        assert np.array_equal(got[l], O.ntt_forward(plan, x[l])), f"limb {l}"
    ctx.ntt_inverse(v, 4, 2)
    assert np.array_equal(v.numpy(), x)


# ------------------------------------------------------------------ poly ops
def test_multiply_ntt_reference_kat():
    ctx, b = make(8, [998244353], [3])
    a = dev(np.arange(1, 9, dtype=np.uint64).reshape(1, 8))
    c = dev(np.arange(8, 0, -1, dtype=np.uint64).reshape(1, 8))
    out = empty((1, 8))
    ctx.poly_multiply_ntt(a, c, out, 1)
    assert list(out.numpy()[0]) == [998244193, 998244243, 998244297, 0, 56, 110, 160, 204]


@pytest.mark.parametrize("n,L", [(8, 3), (64, 3), (2048, 3), (4096, 3), (1 << 14, 5), (1 << 15, 4), (1 << 16, 3)])
def test_poly_ops_match_oracle(n, L):
    q, g = chain("below_2^50", 1 << 16, L) if n > 8 else SMALL
    q, g = q[:L], g[:L]
    ctx, b = make(n, q, g)
    rs = np.random.default_rng(n + L)
    x, y, z = rand_poly(q, n, rs), rand_poly(q, n, rs), rand_poly(q, n, rs)
    dx, dy, dz, out = dev(x), dev(y), dev(z), empty(x.shape)
    ctx.poly_multiply_ntt(dx, dy, out, L)
    assert np.array_equal(out.numpy(), O.rns_multiply_ntt(b, x, y))
    ctx.poly_add(dx, dy, out, L)
    assert np.array_equal(out.numpy(), O.rns_add(b, x, y))
    ctx.poly_sub(dx, dy, out, L)
    assert np.array_equal(out.numpy(), O.rns_sub(b, x, y))
    for gel in (3, 5, 2 * n - 1):
        ctx.poly_galois(dx, out, L, gel)
        assert np.array_equal(out.numpy(), O.rns_galois(b, x, gel))
    ctx.poly_mul_add(dx, dy, dz, out, L)
    want = np.stack([(x[l].astype(object) * y[l].astype(object) + z[l].astype(object)) % q[l] for l in range(L)])
    assert np.array_equal(out.numpy(), want.astype(np.uint64))


def test_poly_ops_on_eval_basis_aux_limbs():
    n = 1 << 12
    q, g = chain("below_2^50", 1 << 16, 6)
    ctx, _ = make(n, q[:4], g[:4], q[4:], g[4:])
    e = O.Basis(n, q[:2] + q[4:], g[:2] + g[4:])  # Q-prefix of 2 limbs || P
    rs = np.random.default_rng(3)
    x, y = rand_poly(e.moduli, n, rs), rand_poly(e.moduli, n, rs)
    out = empty(x.shape)
    ctx.poly_multiply_ntt(dev(x), dev(y), out, 2, 2)
    assert np.array_equal(out.numpy(), O.rns_multiply_ntt(e, x, y))


# ------------------------------------------------------------------ scheme ops
@pytest.mark.parametrize("n,L", [(8, 3), (4096, 3), (1 << 15, 6)])
def test_mult_mult_plain_rescale_match_oracle(n, L):
    q, g = chain("above_2^50", 1 << 15, L) if n > 8 else SMALL
    ctx, b = make(n, q, g)
    rs = np.random.default_rng(n)
    lhs, rhs, pt = rand_poly(q, n, rs, 2), rand_poly(q, n, rs, 2), rand_poly(q, n, rs)
    out = empty((3, L, n))
    assert ctx.mult(dev(lhs), dev(rhs), L, 40, 40, out) == 80
    quad = O.multiply_ciphertexts(b, lhs, rhs)
    assert np.array_equal(out.numpy(), quad)
    o2 = empty((2, L, n))
    assert ctx.mult_plain(dev(lhs), dev(pt), L, 40, o2) == 80
    assert np.array_equal(o2.numpy(), O.multiply_plain(b, lhs, pt))
    o3 = empty((2, L - 1, n))
    s = ctx.rescale(dev(lhs), L, 90, o3)
    assert s == 90 - (int(q[-1]).bit_length() - 1)
    assert np.array_equal(o3.numpy(), O.rescale(b, lhs))
    o4 = empty((3, L - 1, n))
    ctx.rescale_quadratic(dev(quad), L, 80, o4)
    assert np.array_equal(o4.numpy(), O.rescale(b, quad))
    with pytest.raises(G.RckError, match="ciphertexts must share the same scale to multiply"):
        ctx.mult(dev(lhs), dev(rhs), L, 40, 41, out)


def _qonly_setup(n, L, seed, kind="below_2^50"):
    q, g = chain(kind, 1 << 16, L)
    ctx, b = make(n, q, g)
    rng = O.Rng(seed)
    s = O.sample_ternary(b, rng)
    pk, _ = O.keygen_public(b, rng, s)
    return ctx, b, rng, s, pk


@pytest.mark.parametrize("n,L", [(8, 3), (1024, 3), (4096, 3), (1 << 14, 4)])
def test_qonly_keyswitch_relin_rotate_match_oracle(n, L):
    if n == 8:
        ctx, b = make(8, *SMALL)
        rng = O.Rng(5)
        s = O.sample_sparse_ternary(b, rng, 3)
        pk, _ = O.keygen_public(b, rng, s)
    else:
        ctx, b, rng, s, pk = _qonly_setup(n, L, 5)
    L = b.levels
    rlk = O.keygen_qonly(b, rng, O.rns_multiply_ntt(b, s, s), pk)
    dk = ctx.ksk_create(rlk, L - 1, 0)
    rs = np.random.default_rng(7)
    quad = rand_poly(b.moduli, n, rs, 3)
    out = empty((2, L, n))
    assert ctx.relinearize(dev(quad), L, 80, dk, out) == 80
    assert np.array_equal(out.numpy(), O.relinearize(b, None, quad, rlk))
    assert ctx.keyswitch(dev(quad[2]), L, dk, out) == L  # keyswitching.rs:156 scale quirk
    assert np.array_equal(out.numpy(), O.keyswitch(b, None, quad[2], rlk))
    # a key with more digits than the active chain: all ksk.len() digits are used (keyswitching.rs:138)
    extra = np.concatenate([rlk, O.keygen_qonly(b, rng, s, pk, digits=2)])
    dke = ctx.ksk_create(extra, L - 1, 0)
    ctx.keyswitch(dev(quad[2]), L, dke, out)
    assert np.array_equal(out.numpy(), O.keyswitch(b, None, quad[2], extra))
    # rotation by one slot and conjugation
    for gel in (O.galois_element_for_rotation(n, 1), 2 * n - 1):
        rk = O.keygen_qonly(b, rng, O.rns_galois(b, s, gel), pk)
        drk = ctx.ksk_create(rk, L - 1, 0, gel)
        ct = rand_poly(b.moduli, n, rs, 2)
        assert ctx.rotate(dev(ct), L, 40, drk, out) == 40
        assert np.array_equal(out.numpy(), O.rotate(b, None, ct, gel, rk))


def _hybrid_setup(n, params, L, seed):
    q, g, p, gp = params.primes[:L], params.roots[:L], params.aux_primes, params.aux_roots
    ctx = G.Context(poly_degree=n, primes=params.primes, roots=params.roots, aux_primes=p, aux_roots=gp,
                    scale_bits=params.scale_bits)
    qb = O.Basis(n, q, g)
    eb = O.Basis(n, q + p, g + gp)
    return ctx, qb, eb


@pytest.mark.parametrize("n,L", [(4096, 8), (4096, 3), (1 << 15, 8), (1 << 13, 1)])
def test_hybrid_keyswitch_relin_rotate_match_oracle(n, L):
    params = PR.realistic_8_level()
    ctx, qb, eb = _hybrid_setup(n, params, L, 9)
    rng = O.Rng(11)
    s = O.sample_sparse_ternary(qb, rng, 64)
    rlk = O.keygen_hybrid(qb, eb, rng, O.rns_multiply_ntt(qb, s, s), s)
    dk = ctx.ksk_create(rlk, L - 1, 2)
    rs = np.random.default_rng(n + L)
    quad = rand_poly(qb.moduli, n, rs, 3)
    out = empty((2, L, n))
    ctx.relinearize(dev(quad), L, 90, dk, out)
    assert np.array_equal(out.numpy(), O.relinearize(qb, eb, quad, rlk))
    gel = O.galois_element_for_rotation(n, -3)
    rk = O.keygen_hybrid(qb, eb, rng, O.rns_galois(qb, s, gel), s)
    drk = ctx.ksk_create(rk, L - 1, 2, gel)
    ct = rand_poly(qb.moduli, n, rs, 2)
    ctx.rotate(dev(ct), L, 45, drk, out)
    assert np.array_equal(out.numpy(), O.rotate(qb, eb, ct, gel, rk))
    with pytest.raises(G.RckError, match="rotation key level must match the ciphertext level"):
        ctx.rotate(dev(ct[:, :1].copy()) if L > 1 else dev(ct), 1 if L > 1 else 2, 45, drk, out)


def test_mult_relinearized_matches_oracle_and_decodes():
    """mult_relinearized (context/ops.rs:124-133) at the crate's realistic preset shape, N=2^13."""
    n, params = 1 << 13, PR.realistic_8_level()
    L = 8
    ctx, qb, eb = _hybrid_setup(n, params, L, 1)
    rng = O.Rng(2)
    s = O.sample_sparse_ternary(qb, rng, 64)
    pk, _ = O.keygen_public(qb, rng, s)
    rlk = O.keygen_hybrid(qb, eb, rng, O.rns_multiply_ntt(qb, s, s), s)
    dk = ctx.ksk_create(rlk, L - 1, 2)
    rs = np.random.default_rng(0)
    za = rs.uniform(-1, 1, n // 2) + 1j * rs.uniform(-1, 1, n // 2)
    zb = rs.uniform(-1, 1, n // 2) + 1j * rs.uniform(-1, 1, n // 2)
    cts = []
    for z in (za, zb):
        u, e0, e1 = O.sample_ternary(qb, rng), O.sample_error(qb, rng), O.sample_error(qb, rng)
        cts.append(O.encrypt_with_samples(qb, E.encode(qb.moduli, n, 45, z), pk, u, e0, e1))
    out = empty((2, L - 1, n))
    scale = ctx.mult_relinearized(dev(cts[0]), dev(cts[1]), L, 45, 45, dk, out)
    ob = qb.prefix(L - 1)
    want = O.rescale(qb, O.relinearize(qb, eb, O.multiply_ciphertexts(qb, cts[0], cts[1]), rlk))
    got = out.numpy()
    assert np.array_equal(got, want)
    assert scale == 90 - 50
    dec = E.decode(ob.moduli, n, scale, O.decrypt(ob, got, s))
    assert np.abs(dec - za * zb).max() < 1.8e-2  # rns/tests.rs:874-900 tolerance


def test_config1_hadamard_product_flow_decodes():
    """N=2^12, 3 primes < 2^50, scale 2^40, uniform ternary secret, sigma 3.2, seed 0, Q-only keys:
    encrypt two slot vectors, GPU mult + relinearize, decrypt + decode vs FP64 elementwise product."""
    n = 1 << 12
    params = PR.config_plain_cpu()
    ctx = G.Context(params)
    qb = O.Basis(n, params.primes, params.roots)
    rng = O.Rng(0)
    s = O.sample_ternary(qb, rng)
    pk, _ = O.keygen_public(qb, rng, s, 3.2)
    rlk = O.keygen_qonly(qb, rng, O.rns_multiply_ntt(qb, s, s), pk, 3.2)
    dk = ctx.ksk_create(rlk, 2, 0)
    rs = np.random.default_rng(0)
    za = rs.uniform(-1, 1, n // 2) + 1j * rs.uniform(-1, 1, n // 2)
    zb = rs.uniform(-1, 1, n // 2) + 1j * rs.uniform(-1, 1, n // 2)
    cts = []
    for z in (za, zb):
        u, e0, e1 = O.sample_ternary(qb, rng), O.sample_error(qb, rng, 3.2), O.sample_error(qb, rng, 3.2)
        cts.append(O.encrypt_with_samples(qb, E.encode(qb.moduli, n, 40, z), pk, u, e0, e1))
    quad = empty((3, 3, n))
    assert ctx.mult(dev(cts[0]), dev(cts[1]), 3, 40, 40, quad) == 80
    out = empty((2, 3, n))
    ctx.relinearize(quad, 3, 80, dk, out)
    got = out.numpy()
    assert np.array_equal(got, O.relinearize(qb, None, O.multiply_ciphertexts(qb, cts[0], cts[1]), rlk))
    dec = E.decode(qb.moduli, n, 80, O.decrypt(qb, got, s))
    assert np.abs(dec - za * zb).max() < 1.8e-2


def test_config2_full_size_mult_relin_bit_exact():
    """N=2^16, the 24 largest primes < 2^50 (q = 1 mod 2N) plus the next 2 as P, D=40 digits:
    uniform seeded ciphertexts and a uniform seeded key; GPU mult + relinearize == oracle, limb by limb."""
    params = PR.config_per_op_gpu()
    n, L = params.poly_degree, 24
    ctx = G.Context(params)
    qb = O.Basis(n, params.primes, params.roots)
    eb = O.Basis(n, params.primes + params.aux_primes, params.roots + params.aux_roots)
    D = ctx.num_decomp_digits(L)
    assert D == 40 == O.num_decomp_digits(qb.total_modulus_bits())
    rs = np.random.default_rng(2)
    lhs, rhs = rand_poly(qb.moduli, n, rs, 2), rand_poly(qb.moduli, n, rs, 2)
    key = np.stack([rand_poly(eb.moduli, n, rs, 2) for _ in range(D)])
    dk = ctx.ksk_create(key, L - 1, 2)
    quad_d = empty((3, L, n))
    ctx.mult(dev(lhs), dev(rhs), L, 40, 40, quad_d)
    quad = O.multiply_ciphertexts(qb, lhs, rhs)
    assert np.array_equal(quad_d.numpy(), quad)
    out = empty((2, L, n))
    ctx.relinearize(quad_d, L, 80, dk, out)
    want = O.relinearize(qb, eb, quad, key)
    got = out.numpy()
    for comp in range(2):
        for l in range(L):
            assert np.array_equal(got[comp, l], want[comp, l]), f"c{comp} limb {l}"


def test_config2_full_size_ntt_and_rotate_bit_exact():
    """The rest of the config-2 run at full size (N=2^16, 24 Q + 2 P primes): the crate's cyclic
    forward NTT of all 26 limbs (NttPlan::forward per limb) and its inverse, and a rotation by one
    slot with a seeded uniform key, limb by limb against the oracle."""
    params = PR.config_per_op_gpu()
    n, L = params.poly_degree, 24
    ctx = G.Context(params)
    qb = O.Basis(n, params.primes, params.roots)
    eb = O.Basis(n, params.primes + params.aux_primes, params.roots + params.aux_roots)
    rs = np.random.default_rng(24)
    x = rand_poly(eb.moduli, n, rs)
    v = dev(x)
    ctx.ntt_forward(v, L, 2)
    got = v.numpy()
    for l in range(L + 2):
        assert np.array_equal(got[l], O.ntt_forward(eb.plan(l), x[l])), f"limb {l}"
    ctx.ntt_inverse(v, L, 2)
    assert np.array_equal(v.numpy(), x)
    D = ctx.num_decomp_digits(L)
    key = np.stack([rand_poly(eb.moduli, n, rs, 2) for _ in range(D)])
    gel = O.galois_element_for_rotation(n, 1)
    drk = ctx.ksk_create(key, L - 1, 2, gel)
    ct = rand_poly(qb.moduli, n, rs, 2)
    out = empty((2, L, n))
    ctx.rotate(dev(ct), L, 40, drk, out)
    want = O.rotate(qb, eb, ct, gel, key)
    got = out.numpy()
    for comp in range(2):
        for l in range(L):
            assert np.array_equal(got[comp, l], want[comp, l]), f"c{comp} limb {l}"


@pytest.mark.parametrize("n,L,K", [(4096, 5, 1), (4096, 4, 3), (1 << 14, 5, 4), (1 << 13, 6, 1)])
def test_hybrid_aux_prime_counts_match_oracle(n, L, K):
    """Hybrid relinearize, rotate and mult_relinearized with K = 1, 3, 4 aux primes (the mod-down's
    Garner over P, its centring and the per-K specialised epilogue), uniform seeded key rows."""
    p = PR.config_batched(n=n, limbs=L, aux=K)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    qb = O.Basis(n, p.primes, p.roots)
    eb = O.Basis(n, p.primes + p.aux_primes, p.roots + p.aux_roots)
    D = ctx.num_decomp_digits(L)
    rs = np.random.default_rng(100 * K + L)
    key = np.stack([rand_poly(eb.moduli, n, rs, 2) for _ in range(D)])
    dk = ctx.ksk_create(key, L - 1, K)
    quad = rand_poly(qb.moduli, n, rs, 3)
    out = empty((2, L, n))
    ctx.relinearize(dev(quad), L, 80, dk, out)
    assert np.array_equal(out.numpy(), O.relinearize(qb, eb, quad, key))
    gel = O.galois_element_for_rotation(n, 5)
    drk = ctx.ksk_create(key, L - 1, K, gel)
    ct = rand_poly(qb.moduli, n, rs, 2)
    ctx.rotate(dev(ct), L, 40, drk, out)
    assert np.array_equal(out.numpy(), O.rotate(qb, eb, ct, gel, key))
    lhs, rhs = rand_poly(qb.moduli, n, rs, 2), rand_poly(qb.moduli, n, rs, 2)
    prod = empty((2, L - 1, n))
    ctx.mult_relinearized(dev(lhs), dev(rhs), L, 40, 40, dk, prod)
    want = O.rescale(qb, O.relinearize(qb, eb, O.multiply_ciphertexts(qb, lhs, rhs), key))
    assert np.array_equal(prod.numpy(), want)


@pytest.mark.parametrize("kind", ["below_2^50", "above_2^50"])
def test_relinearize_26_limbs_matches_oracle(kind):
    """26 ciphertext primes (more than the decomposition keeps in registers: its Garner digits go
    through LDS) and 2 aux primes at N=4096, 44 digits: relinearize and rotate == oracle."""
    n = 4096
    q, g = chain(kind, n, 28)
    ctx = G.Context(poly_degree=n, primes=q[:26], roots=g[:26], aux_primes=q[26:], aux_roots=g[26:])
    qb = O.Basis(n, q[:26], g[:26])
    eb = O.Basis(n, q, g)
    D = ctx.num_decomp_digits(26)
    assert D == O.num_decomp_digits(qb.total_modulus_bits())
    rs = np.random.default_rng(26)
    key = np.stack([rand_poly(eb.moduli, n, rs, 2) for _ in range(D)])
    dk = ctx.ksk_create(key, 25, 2)
    quad = rand_poly(qb.moduli, n, rs, 3)
    out = empty((2, 26, n))
    ctx.relinearize(dev(quad), 26, 80, dk, out)
    assert np.array_equal(out.numpy(), O.relinearize(qb, eb, quad, key))
    gel = O.galois_element_for_rotation(n, -3)
    drk = ctx.ksk_create(key, 25, 2, gel)
    ct = rand_poly(qb.moduli, n, rs, 2)
    ctx.rotate(dev(ct), 26, 40, drk, out)
    assert np.array_equal(out.numpy(), O.rotate(qb, eb, ct, gel, key))


def test_relinearize_mixed_size_primes_matches_oracle():
    """A chain mixing ~2^50 and ~2^30 primes (the decomposition's Garner then keeps its reductions
    and its digits in LDS): relinearize == oracle at N=4096."""
    n = 4096
    big, small = PR.primes_below(50, n, 4), PR.primes_below(30, n, 2)
    q = [big[0], small[0], big[1], small[1]]
    aux = big[2:]
    roots = [PR.smallest_primitive_root(x) for x in q]
    aroots = [PR.smallest_primitive_root(x) for x in aux]
    ctx = G.Context(poly_degree=n, primes=q, roots=roots, aux_primes=aux, aux_roots=aroots)
    qb = O.Basis(n, q, roots)
    eb = O.Basis(n, q + aux, roots + aroots)
    D = ctx.num_decomp_digits(4)
    assert D == O.num_decomp_digits(qb.total_modulus_bits())
    rs = np.random.default_rng(7)
    key = np.stack([rand_poly(eb.moduli, n, rs, 2) for _ in range(D)])
    dk = ctx.ksk_create(key, 3, 2)
    quad = rand_poly(qb.moduli, n, rs, 3)
    out = empty((2, 4, n))
    ctx.relinearize(dev(quad), 4, 80, dk, out)
    assert np.array_equal(out.numpy(), O.relinearize(qb, eb, quad, key))


def test_largest_allowed_primes_match_oracle():
    """Primes just below the FP64-exact bound 2^50 + 2^46 (the largest the context accepts, where the
    lazily reduced butterflies and products run closest to their limits): NTT round trip, negacyclic
    product, mult + relinearize + rescale and rotate == oracle at N=4096."""
    n, step, top = 4096, 8192, (1 << 50) + (1 << 46)
    q, cand = [], (top - 1) // step * step + 1
    while len(q) < 6:
        if PR.is_prime(cand):
            q.append(cand)
        cand -= step
    g = [PR.smallest_primitive_root(x) for x in q]
    ctx = G.Context(poly_degree=n, primes=q[:4], roots=g[:4], aux_primes=q[4:], aux_roots=g[4:])
    qb = O.Basis(n, q[:4], g[:4])
    eb = O.Basis(n, q, g)
    rs = np.random.default_rng(46)
    a, b = rand_poly(qb.moduli, n, rs), rand_poly(qb.moduli, n, rs)
    prod = empty((4, n))
    ctx.poly_multiply_ntt(dev(a), dev(b), prod, 4)
    assert np.array_equal(prod.numpy(), O.rns_multiply_ntt(qb, a, b))
    D = ctx.num_decomp_digits(4)
    key = np.stack([rand_poly(eb.moduli, n, rs, 2) for _ in range(D)])
    dk = ctx.ksk_create(key, 3, 2)
    lhs, rhs = rand_poly(qb.moduli, n, rs, 2), rand_poly(qb.moduli, n, rs, 2)
    out = empty((2, 3, n))
    ctx.mult_relinearized(dev(lhs), dev(rhs), 4, 40, 40, dk, out)
    want = O.rescale(qb, O.relinearize(qb, eb, O.multiply_ciphertexts(qb, lhs, rhs), key))
    assert np.array_equal(out.numpy(), want)
    gel = O.galois_element_for_rotation(n, 7)
    drk = ctx.ksk_create(key, 3, 2, gel)
    ct = rand_poly(qb.moduli, n, rs, 2)
    rot = empty((2, 4, n))
    ctx.rotate(dev(ct), 4, 40, drk, rot)
    assert np.array_equal(rot.numpy(), O.rotate(qb, eb, ct, gel, key))


@pytest.mark.parametrize("logn", [13, 15])
def test_largest_allowed_primes_long_digit_runs(logn):
    """Primes just below 2^50 + 2^46 with 10 limbs (D = 17 digits): the streamed MAC's accumulators
    take 7 (N = 2^13, three-stage row pass) or 8 (N = 2^15) products between reductions at the largest
    values the bound allows.  A batch of 4 (one ciphertext per wave, every digit on one wave) and the
    per-op calls (four digit ranges per workgroup) for mult+relin and rotate; op 0 against the oracle,
    the rest batch == per-op."""
    n, L, K, B = 1 << logn, 10, 2, 4
    step, top = 2 * n, (1 << 50) + (1 << 46)
    q, cand = [], (top - 1) // step * step + 1
    while len(q) < L + K:
        if PR.is_prime(cand):
            q.append(cand)
        cand -= step
    g = [PR.smallest_primitive_root(x) for x in q]
    ctx = G.Context(poly_degree=n, primes=q[:L], roots=g[:L], aux_primes=q[L:], aux_roots=g[L:])
    qb = O.Basis(n, q[:L], g[:L])
    eb = O.Basis(n, q, g)
    D = ctx.num_decomp_digits(L)
    assert D >= 15
    rs = np.random.default_rng(4600 + logn)
    key = np.stack([rand_poly(eb.moduli, n, rs, 2) for _ in range(D)])
    gel = O.galois_element_for_rotation(n, -3)
    dk = ctx.ksk_create(key, L - 1, K)
    drk = ctx.ksk_create(key, L - 1, K, gel)
    P = L * n
    lhs = rand_poly(qb.moduli, n, rs, 2 * B).reshape(B, 2, L, n)
    rhs = rand_poly(qb.moduli, n, rs, 2 * B).reshape(B, 2, L, n)
    dl, dr = dev(lhs), dev(rhs)
    out_m, out_r = empty((B, 2, L, n)), empty((B, 2, L, n))
    bt = G.Batch(ctx)
    for b in range(B):
        bt.add(G.OP_MULT_RELIN, L, dl.offset(b * 2 * P), dr.offset(b * 2 * P), dk, out_m.offset(b * 2 * P))
        bt.add(G.OP_ROTATE, L, dl.offset(b * 2 * P), None, drk, out_r.offset(b * 2 * P))
    bt.compile(tune=False)
    got_m, got_r = out_m.numpy(), out_r.numpy()
    O.set_threads(16)
    want_m = O.relinearize(qb, eb, O.multiply_ciphertexts(qb, lhs[0], rhs[0]), key)
    want_r = O.rotate(qb, eb, lhs[0], gel, key)
    assert np.array_equal(got_m[0], want_m)
    assert np.array_equal(got_r[0], want_r)
    quad, one = empty((3, L, n)), empty((2, L, n))
    for b in range(B):
        ctx.mult(dl.offset(b * 2 * P), dr.offset(b * 2 * P), L, 40, 40, quad)
        ctx.relinearize(quad, L, 80, dk, one)
        assert np.array_equal(one.numpy(), got_m[b]), f"mult+relin {b}"
        ctx.rotate(dl.offset(b * 2 * P), L, 40, drk, one)
        assert np.array_equal(one.numpy(), got_r[b]), f"rotate {b}"


def test_context_rejects_invalid_chains():
    """RnsContext / NttPlan::new validation (math/ntt.rs:17-60) with the crate's error texts, at every
    modulus size; the arithmetic follows the largest modulus (FP64-exact below 2^50 + 2^46)."""
    with pytest.raises(G.RckError, match="invalid RNS prime 41"):
        G.Context(poly_degree=8, primes=[41], roots=[6])
    # 2^62 - 2^16 + 1 = 4611686018427322369 is prime with q = 1 mod 16; the root 1 generates no
    # primitive 8th root of unity
    with pytest.raises(G.RckError, match="invalid RNS prime 4611686018427322369"):
        G.Context(poly_degree=8, primes=[4611686018427322369], roots=[1])
    q, g = chain("above_2^50", 1 << 16, 1)
    ctx = G.Context(poly_degree=8, primes=q, roots=g)  # primes just above 2^50: FP64-exact
    assert ctx.arithmetic == G.ARITH_FP64
    big = PR.primes_below(64, 8, 1)[0]
    ctx = G.Context(poly_degree=8, primes=[big], roots=[PR.smallest_primitive_root(big)])
    assert ctx.arithmetic == G.ARITH_INT64


def test_limb_counts_out_of_range_rejected():
    """Empty (0 limbs) and oversized (more limbs than the chain, or a partial aux basis) operands are
    rejected with an error before any launch, for the polynomial ops, the scheme ops and the
    keyswitching entry points."""
    n = 4096
    params = PR.realistic_8_level()
    ctx = G.Context(poly_degree=n, primes=params.primes[:3], roots=params.roots[:3],
                    aux_primes=params.aux_primes, aux_roots=params.aux_roots, scale_bits=45)
    buf = empty((3, 6, n))
    for nq in (0, 4):
        with pytest.raises(G.RckError, match="limb count must be in 1..=3"):
            ctx.ntt_forward(buf, nq)
        with pytest.raises(G.RckError, match="limb count must be in 1..=3"):
            ctx.poly_add(buf, buf, buf, nq)
    with pytest.raises(G.RckError, match="aux limb count must be 0 or the aux basis size"):
        ctx.poly_multiply_ntt(buf, buf, buf, 3, 1)
    with pytest.raises(G.RckError, match="limb count must be in 1..=3, got 0"):
        ctx.mult(buf, buf, 0, 45, 45, buf)
    with pytest.raises(G.RckError, match="limb count must be in 1..=3, got 4"):
        ctx.rescale(buf, 4, 90, buf)
    with pytest.raises(G.RckError, match="rescale requires at least two levels"):
        ctx.rescale(buf, 1, 90, empty((2, 1, n)))
    # the device is still usable afterwards
    rs = np.random.default_rng(3)
    a, b = rand_poly(params.primes[:3], n, rs), rand_poly(params.primes[:3], n, rs)
    out = empty((3, n))
    ctx.poly_add(dev(a), dev(b), out, 3)
    qb = O.Basis(n, params.primes[:3], params.roots[:3])
    assert np.array_equal(out.numpy(), O.rns_add(qb, a, b))
