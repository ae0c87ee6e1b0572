"""64-bit integer arithmetic: moduli up to the crate's full u64 range (src/math/ntt.rs:132-150 reduces
every product with u128), bit for bit.

* A chain of primes below 2^50 run through the 64-bit integer kernels (rck_context_set_arithmetic)
  gives exactly what the FP64-exact kernels give, op by op: both paths on the same inputs.
* Chains with moduli of 56 to 64 bits (the context switches to the integer kernels by itself) match
  the oracle's u128 restatement: cyclic NTT, poly ops, ct x ct, ct x pt, rescale, hybrid relinearize
  and rotate, Q-only keyswitch, key export, keygen draw for draw, encryption and decryption.
* Entry points without an integer path reject such a context."""
import numpy as np
import pytest

from oracle import oracle as O
from rnsckks import capi as G
from rnsckks import params as PR

pytestmark = pytest.mark.gpu


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


_CACHE = {}


def wide_primes(bits, count):
    """The largest primes below 2^bits with q = 1 mod 2^17 (NTT-friendly for every N <= 2^16)."""
    key = (bits, count)
    if key not in _CACHE:
        qs = PR.primes_below(bits, 1 << 16, count)
        _CACHE[key] = (qs, [PR.smallest_primitive_root(q) for q in qs])
    return _CACHE[key]


def keyswitch_key(moduli, n, D, rs):
    return np.stack([rand_poly(moduli, n, rs, 2) for _ in range(D)])


# ------------------------------------------------------------------ integer kernels == FP64 kernels
@pytest.mark.parametrize("logn,L", [(12, 4), (14, 5)])
def test_int64_kernels_equal_fp64_kernels(logn, L):
    n, K = 1 << logn, 2
    p = PR.config_batched(n=1 << 16, limbs=L, aux=K)
    ctxs = []
    for arith in (G.ARITH_FP64, G.ARITH_INT64):
        c = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
        c.set_keyswitch_method(G.KS_DIGIT_NTT)
        c.set_arithmetic(arith)
        assert c.arithmetic == arith
        ctxs.append(c)
    rs = np.random.default_rng(40 + logn)
    q = p.primes
    ev = p.primes + p.aux_primes
    x, y, z = rand_poly(ev, n, rs), rand_poly(ev, n, rs), rand_poly(ev, n, rs)
    lhs, rhs, pt = rand_poly(q, n, rs, 2), rand_poly(q, n, rs, 2), rand_poly(q, n, rs)
    quad = rand_poly(q, n, rs, 3)
    D = ctxs[0].num_decomp_digits(L)
    key = keyswitch_key(ev, n, D, rs)
    gel = G.galois_element_for_rotation(n, 3)
    results = []
    for c in ctxs:
        r = {}
        v = dev(x)
        c.ntt_forward(v, L, K)
        r["ntt"] = v.numpy()
        c.ntt_inverse(v, L, K)
        r["intt"] = v.numpy()
        o = empty(x.shape)
        for name, fn in (("add", c.poly_add), ("sub", c.poly_sub), ("mul", c.poly_multiply_ntt)):
            fn(dev(x), dev(y), o, L, K)
            r[name] = o.numpy()
        c.poly_galois(dev(x), o, L, gel, K)
        r["galois"] = o.numpy()
        c.poly_mul_add(dev(x), dev(y), dev(z), o, L, K)
        r["mul_add"] = o.numpy()
        o3 = empty((3, L, n))
        c.mult(dev(lhs), dev(rhs), L, 40, 40, o3)
        r["mult"] = o3.numpy()
        o2 = empty((2, L, n))
        c.mult_plain(dev(lhs), dev(pt), L, 40, o2)
        r["mult_plain"] = o2.numpy()
        o1 = empty((2, L - 1, n))
        c.rescale(dev(lhs), L, 90, o1)
        r["rescale"] = o1.numpy()
        oq = empty((3, L - 1, n))
        c.rescale_quadratic(dev(quad), L, 90, oq)
        r["rescale_quadratic"] = oq.numpy()
        k = c.ksk_create(key, L - 1, K)
        rk = c.ksk_create(key, L - 1, K, gel)
        r["export"] = c.ksk_export(k)
        for name, fn in (("relin", lambda o: c.relinearize(dev(quad), L, 80, k, o)),
                         ("keyswitch", lambda o: c.keyswitch(dev(quad[2]), L, k, o)),
                         ("rotate", lambda o: c.rotate(dev(lhs), L, 40, rk, o))):
            o = empty((2, L, n))
            fn(o)
            r[name] = o.numpy()
        o = empty((2, L, n))
        c.mult_relinearized(dev(lhs), dev(rhs), L, 40, 40, k, o)
        r["mult_relin"] = o.numpy().reshape(-1)[: 2 * (L - 1) * n]
        results.append(r)
    for name in results[0]:
        assert np.array_equal(results[0][name], results[1][name]), name
    assert np.array_equal(results[1]["export"], key)
    # a key belongs to the arithmetic it was created under
    kf = ctxs[0].ksk_create(key, L - 1, K)
    with pytest.raises(G.RckError, match="another arithmetic"):
        ctxs[1].relinearize(dev(quad), L, 80, kf, empty((2, L, n)))


# ------------------------------------------------------------------ wide moduli against the oracle
@pytest.mark.parametrize("logn,qbits,L", [(10, 56, 3), (13, 60, 4), (12, 62, 3), (11, 64, 3)])
def test_wide_moduli_match_oracle(logn, qbits, L):
    n, K = 1 << logn, 2
    qs, gs = wide_primes(qbits, L + K)
    q, g, p, gp = qs[:L], gs[:L], qs[L:], gs[L:]
    ctx = G.Context(poly_degree=n, primes=q, roots=g, aux_primes=p, aux_roots=gp, scale_bits=40)
    assert ctx.arithmetic == G.ARITH_INT64
    qb, eb = O.Basis(n, q, g), O.Basis(n, q + p, g + gp)
    rs = np.random.default_rng(qbits + logn)
    # cyclic NTT (natural order) of every limb of Q || P, and back
    x = rand_poly(q + p, n, rs)
    v = dev(x)
    ctx.ntt_forward(v, L, K)
    got = v.numpy()
    for l, (m, r) in enumerate(zip(q + p, g + gp)):
        assert np.array_equal(got[l], O.ntt_forward(O.ntt_plan(n, m, r), x[l])), f"limb {l}"
    ctx.ntt_inverse(v, L, K)
    assert np.array_equal(v.numpy(), x)
    # poly ops over the eval basis
    y, z = rand_poly(q + p, n, rs), rand_poly(q + p, n, rs)
    out = empty(x.shape)
    ctx.poly_multiply_ntt(dev(x), dev(y), out, L, K)
    assert np.array_equal(out.numpy(), O.rns_multiply_ntt(eb, x, y))
    ctx.poly_add(dev(x), dev(y), out, L, K)
    assert np.array_equal(out.numpy(), O.rns_add(eb, x, y))
    ctx.poly_sub(dev(x), dev(y), out, L, K)
    assert np.array_equal(out.numpy(), O.rns_sub(eb, x, y))
    for gel in (3, 2 * n - 1):
        ctx.poly_galois(dev(x), out, L, gel, K)
        assert np.array_equal(out.numpy(), O.rns_galois(eb, x, gel))
    ctx.poly_mul_add(dev(x), dev(y), dev(z), out, L, K)
    want = np.stack([(x[l].astype(object) * y[l].astype(object) + z[l].astype(object)) % m
                     for l, m in enumerate(q + p)])
    assert np.array_equal(out.numpy(), want.astype(np.uint64))
    # scheme ops over Q
    lhs, rhs, pt = rand_poly(q, n, rs, 2), rand_poly(q, n, rs, 2), rand_poly(q, n, rs)
    o3 = empty((3, L, n))
    ctx.mult(dev(lhs), dev(rhs), L, 40, 40, o3)
    quad = O.multiply_ciphertexts(qb, lhs, rhs)
    assert np.array_equal(o3.numpy(), quad)
    o2 = empty((2, L, n))
    ctx.mult_plain(dev(lhs), dev(pt), L, 40, o2)
    assert np.array_equal(o2.numpy(), O.multiply_plain(qb, lhs, pt))
    o1 = empty((2, L - 1, n))
    assert ctx.rescale(dev(lhs), L, 120, o1) == 120 - (int(q[-1]).bit_length() - 1)
    assert np.array_equal(o1.numpy(), O.rescale(qb, lhs))
    oq = empty((3, L - 1, n))
    ctx.rescale_quadratic(dev(quad), L, 120, oq)
    assert np.array_equal(oq.numpy(), O.rescale(qb, quad))
    # hybrid keyswitch: relinearize, keyswitch, rotate, mult_relinearized
    D = ctx.num_decomp_digits(L)
    assert D == O.num_decomp_digits(qb.total_modulus_bits())
    key = keyswitch_key(q + p, n, D, rs)
    k = ctx.ksk_create(key, L - 1, K)
    assert np.array_equal(ctx.ksk_export(k), key)
    o = empty((2, L, n))
    ctx.relinearize(dev(quad), L, 80, k, o)
    assert np.array_equal(o.numpy(), O.relinearize(qb, eb, quad, key))
    ctx.keyswitch(dev(quad[2]), L, k, o)
    assert np.array_equal(o.numpy(), O.keyswitch(qb, eb, quad[2], key))
    gel = G.galois_element_for_rotation(n, -2)
    rk = ctx.ksk_create(key, L - 1, K, gel)
    ctx.rotate(dev(lhs), L, 40, rk, o)
    assert np.array_equal(o.numpy(), O.rotate(qb, eb, lhs, gel, key))
    om = empty((2, L, n))
    ctx.mult_relinearized(dev(lhs), dev(rhs), L, 40, 40, k, om)
    want = O.rescale(qb, O.relinearize(qb, eb, quad, key))
    assert np.array_equal(om.numpy().reshape(-1)[: 2 * (L - 1) * n], want.reshape(-1))


@pytest.mark.parametrize("qbits", [60, 64])
def test_wide_qonly_keyswitch_matches_oracle(qbits):
    n, L = 1 << 12, 4
    q, g = wide_primes(qbits, L)
    ctx = G.Context(poly_degree=n, primes=q, roots=g)
    qb = O.Basis(n, q, g)
    rs = np.random.default_rng(qbits)
    D = ctx.num_decomp_digits(L) + 1  # Q-only keys may carry more digits than the chain needs
    key = keyswitch_key(q, n, D, rs)
    k = ctx.ksk_create(key, L - 1, 0)
    quad = rand_poly(q, n, rs, 3)
    o = empty((2, L, n))
    ctx.relinearize(dev(quad), L, 80, k, o)
    assert np.array_equal(o.numpy(), O.relinearize(qb, None, quad, key))
    gel = G.galois_element_for_rotation(n, 7)
    rk = ctx.ksk_create(key, L - 1, 0, gel)
    ct = rand_poly(q, n, rs, 2)
    ctx.rotate(dev(ct), L, 40, rk, o)
    assert np.array_equal(o.numpy(), O.rotate(qb, None, ct, gel, key))


def test_wide_full_ring_relinearize_and_rotate_match_oracle():
    """N = 2^16, 12 primes below 2^60 and 2 aux primes: 24 digits over 14 eval limbs."""
    n, L, K = 1 << 16, 12, 2
    qs, gs = wide_primes(60, L + K)
    ctx = G.Context(poly_degree=n, primes=qs[:L], roots=gs[:L], aux_primes=qs[L:], aux_roots=gs[L:])
    qb, eb = O.Basis(n, qs[:L], gs[:L]), O.Basis(n, qs, gs)
    rs = np.random.default_rng(16)
    D = ctx.num_decomp_digits(L)
    key = keyswitch_key(qs, n, D, rs)
    k = ctx.ksk_create(key, L - 1, K)
    lhs, rhs = rand_poly(qs[:L], n, rs, 2), rand_poly(qs[:L], n, rs, 2)
    o3 = empty((3, L, n))
    ctx.mult(dev(lhs), dev(rhs), L, 40, 40, o3)
    O.set_threads(16)
    quad = O.multiply_ciphertexts(qb, lhs, rhs)
    assert np.array_equal(o3.numpy(), quad)
    o = empty((2, L, n))
    ctx.relinearize(dev(quad), L, 80, k, o)
    assert np.array_equal(o.numpy(), O.relinearize(qb, eb, quad, key))
    gel = G.galois_element_for_rotation(n, 1 << 13)
    rk = ctx.ksk_create(key, L - 1, K, gel)
    ctx.rotate(dev(lhs), L, 40, rk, o)
    assert np.array_equal(o.numpy(), O.rotate(qb, eb, lhs, gel, key))


def test_wide_keygen_encrypt_decrypt_match_oracle():
    """Seeded keys, public key, encryption and decryption on 61-bit primes, draw for draw."""
    n, L, K = 1 << 12, 3, 2
    qs, gs = wide_primes(61, L + K)
    q, g = qs[:L], gs[:L]
    ctx = G.Context(poly_degree=n, primes=q, roots=g, aux_primes=qs[L:], aux_roots=gs[L:])
    gr, o = G.Rng(5), O.Rng(5)
    qb, eb = O.Basis(n, q, g), O.Basis(n, qs, gs)
    s = ctx.sample_sparse_ternary(gr, 32, L)
    so = O.sample_sparse_ternary(qb, o, 32)
    assert np.array_equal(s.numpy(), so)
    assert np.array_equal(ctx.sample_uniform(gr, L, K).numpy(), O.sample_uniform(eb, o))
    k = ctx.keygen_relin(gr, s, L - 1, sigma=3.2)
    want = O.keygen_hybrid(qb, eb, o, O.rns_multiply_ntt(qb, so, so), so, sigma=3.2)
    assert np.array_equal(ctx.ksk_export(k), want)
    pk = ctx.keygen_public(gr, s, L, sigma=3.2)
    pko, _ = O.keygen_public(qb, o, so, sigma=3.2)
    assert np.array_equal(pk.numpy(), pko)
    m = rand_poly(q, n, np.random.default_rng(1))
    ct = ctx.encrypt(gr, dev(m), pk, L, L, sigma=3.2)
    u = O.sample_ternary(qb, o)
    e0 = O.sample_error(qb, o, 3.2)
    e1 = O.sample_error(qb, o, 3.2)
    cto = O.encrypt_with_samples(qb, m, pko, u, e0, e1)
    assert np.array_equal(ct.numpy(), cto)
    assert np.array_equal(ctx.decrypt(ct, L, s).numpy(), O.decrypt(qb, cto, so))
    assert gr.next() == o.next()


def test_paths_without_an_integer_kernel_reject_wide_contexts():
    n, L = 1 << 12, 3
    qs, gs = wide_primes(60, L + 2)
    ctx = G.Context(poly_degree=n, primes=qs[:L], roots=gs[:L], aux_primes=qs[L:], aux_roots=gs[L:])
    with pytest.raises(G.RckError, match="at or above 2\\^50 \\+ 2\\^46"):
        ctx.set_arithmetic(G.ARITH_FP64)
    with pytest.raises(G.RckError, match="no 64-bit integer path"):
        G.Batch(ctx)
    ct = dev(rand_poly(qs[:L], n, np.random.default_rng(0), 2))
    with pytest.raises(G.RckError, match="no 64-bit integer path"):
        ctx.ps_eval(ct, L, 40, [0.5, 0.25], {}, empty((2, L, n)))
    a = dev(rand_poly(qs[:L], n, np.random.default_rng(1)))
    with pytest.raises(G.RckError, match="no 64-bit integer path"):
        ctx.poly_multiply_ntt_noalloc(a, a, empty((L, n)), L)
