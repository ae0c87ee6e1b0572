"""Seeded sampling, key generation, encryption and decryption on the device equal the oracle's
restatement draw for draw and limb by limb (same seed, same draw order)."""
import numpy as np
import pytest

from oracle import encoding as E
from oracle import oracle as O
from rnsckks import capi as G
from rnsckks import params as PR

pytestmark = pytest.mark.gpu

N = 1 << 12


def setup(aux=True, limbs=4):
    p = PR.config_batched(n=1 << 15, limbs=limbs, aux=2)
    ctx = G.Context(poly_degree=N, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes if aux else (),
                    aux_roots=p.aux_roots if aux else ())
    return p, ctx


def basis(p, k, aux=False):
    primes = p.primes[:k] + (p.aux_primes if aux else [])
    roots = p.roots[:k] + (p.aux_roots if aux else [])
    return O.Basis(N, primes, roots)


def test_sampling_matches_oracle():
    p, ctx = setup()
    g, o = G.Rng(7), O.Rng(7)
    L, K = len(p.primes), len(p.aux_primes)
    ev, qb = basis(p, L, True), basis(p, L)
    assert np.array_equal(ctx.sample_uniform(g, L, K).numpy(), O.sample_uniform(ev, o))
    assert np.array_equal(ctx.sample_error(g, L, K, 3.2).numpy(), O.sample_error(ev, o, 3.2))
    assert np.array_equal(ctx.sample_error(g, 2).numpy(), O.sample_error(basis(p, 2), o))  # odd count: spare kept
    assert np.array_equal(ctx.sample_ternary(g, L).numpy(), O.sample_ternary(qb, o))
    s = ctx.sample_sparse_ternary(g, 64, L).numpy()
    assert np.array_equal(s, O.sample_sparse_ternary(qb, o, 64))
    c = np.where(s[0] > p.primes[0] // 2, -1, s[0].astype(np.int64))
    assert np.count_nonzero(c) == 64 and set(np.unique(c)) <= {-1, 0, 1}
    assert g.next() == o.next()


@pytest.mark.parametrize("level", [3, 1])
def test_hybrid_relin_and_rotation_keys_match_oracle(level):
    p, ctx = setup()
    g, o = G.Rng(11), O.Rng(11)
    L = len(p.primes)
    s = ctx.sample_sparse_ternary(g, 64, L)
    so = O.sample_sparse_ternary(basis(p, L), o, 64)
    assert np.array_equal(s.numpy(), so)
    q, e = basis(p, level + 1), basis(p, level + 1, True)
    st = so[: level + 1]
    k = ctx.keygen_relin(g, s, level, sigma=3.2)
    want = O.keygen_hybrid(q, e, o, O.rns_multiply_ntt(q, st, st), st, sigma=3.2)
    assert (k.digits, k.level, k.aux, k.galois) == (want.shape[0], level, 2, 0)
    assert np.array_equal(ctx.ksk_export(k), want)
    for steps in (1, -3):
        r = ctx.keygen_rotation(g, s, steps, level, sigma=3.2)
        gel = O.galois_element_for_rotation(N, steps)
        want = O.keygen_hybrid(q, e, o, O.rns_galois(q, st, gel), st, sigma=3.2)
        assert r.galois == gel
        assert np.array_equal(ctx.ksk_export(r), want)
    assert g.next() == o.next()


def test_qonly_keys_public_key_encrypt_decrypt_match_oracle():
    p, ctx = setup(aux=False)
    g, o = G.Rng(3), O.Rng(3)
    L = len(p.primes)
    qb = basis(p, L)
    s = ctx.sample_ternary(g, L)
    so = O.sample_ternary(qb, o)
    pk = ctx.keygen_public(g, s, L, sigma=3.2)
    pko, _ = O.keygen_public(qb, o, so, sigma=3.2)
    assert np.array_equal(pk.numpy(), pko)
    level = 2
    q = basis(p, level + 1)
    st = so[: level + 1]
    k = ctx.keygen_relin(g, s, level, pk=pk, pk_limbs=L, sigma=3.2)
    want = O.keygen_qonly(q, o, O.rns_multiply_ntt(q, st, st), pko[:, : level + 1], sigma=3.2)
    assert k.aux == 0 and np.array_equal(ctx.ksk_export(k), want)
    r = ctx.keygen_rotation(g, s, 5, level, pk=pk, pk_limbs=L, sigma=3.2)
    want = O.keygen_qonly(q, o, O.rns_galois(q, st, O.galois_element_for_rotation(N, 5)), pko[:, : level + 1],
                          sigma=3.2)
    assert np.array_equal(ctx.ksk_export(r), want)
    # encryption at a truncated level, then decryption
    rs = np.random.default_rng(0)
    m = np.stack([rs.integers(0, qq, N, dtype=np.uint64) for qq in p.primes[:3]])
    ct = ctx.encrypt(g, G.DeviceBuffer.from_numpy(m), pk, L, 3, sigma=3.2)
    b3 = basis(p, 3)
    u = O.sample_ternary(b3, o)
    e0 = O.sample_error(b3, o, 3.2)
    e1 = O.sample_error(b3, o, 3.2)
    cto = O.encrypt_with_samples(b3, m, pko[:, :3], u, e0, e1)
    assert np.array_equal(ct.numpy(), cto)
    assert np.array_equal(ctx.decrypt(ct, 3, s).numpy(), O.decrypt(b3, cto, so[:3]))
    assert g.next() == o.next()


def test_device_keys_decrypt_relinearized_product():
    """Keys generated on the device make a working scheme: encrypt two encoded vectors, multiply,
    relinearize with the device relin key, rotate with a device rotation key, decrypt, decode."""
    p = PR.realistic_8_level()
    n = p.poly_degree
    ctx = G.Context(p)
    g = G.Rng(2024)
    L = len(p.primes)
    s = ctx.sample_sparse_ternary(g, 64, L)
    pk = ctx.keygen_public(g, s, L)
    rlk = ctx.keygen_relin(g, s, L - 1)
    rot = ctx.keygen_rotation(g, s, 1, L - 1)
    rs = np.random.default_rng(1)
    x = rs.uniform(-1, 1, n // 2)
    y = rs.uniform(-1, 1, n // 2)
    cx = ctx.encrypt(g, G.DeviceBuffer.from_numpy(E.encode(p.primes, n, 45, x.astype(complex))), pk, L, L)
    cy = ctx.encrypt(g, G.DeviceBuffer.from_numpy(E.encode(p.primes, n, 45, y.astype(complex))), pk, L, L)
    prod = G.DeviceBuffer((2, L, n))
    ctx.mult_relinearized(cx, cy, L, 45, 45, rlk, prod)
    Lr = L - 1
    dec = ctx.decrypt(prod, Lr, s).numpy()
    got = E.decode(p.primes[:Lr], n, 90 - 50, dec).real
    assert np.max(np.abs(got - x * y)) < 1e-3
    rx = G.DeviceBuffer((2, L, n))
    ctx.rotate(cx, L, 45, rot, rx)
    got = E.decode(p.primes, n, 45, ctx.decrypt(rx, L, s).numpy()).real
    assert np.max(np.abs(got - np.roll(x, -1))) < 1e-3
