"""Pins the CPU oracle against the reference crate's own known-answer tests, closed forms,
independent big-integer computations and exact invariants.  CPU only.

Every check here uses a method different from the oracle's (brute-force DFT, Python-int
schoolbook convolution, Python-int CRT instead of Garner, ...), so a dropped term, a wrong
sign or index, or a transposed operand in the oracle fails one of them.
"""
import json
import os
import random

import numpy as np
import pytest

from oracle import oracle as O
from oracle import encoding as E

HERE = os.path.dirname(os.path.abspath(__file__))
PRIMES = json.load(open(os.path.join(HERE, "golden", "primes.json")))


def chain(kind, n, count, skip=0):
    lst = PRIMES[kind][str(n)][skip:skip + count]
    return [p for p, _ in lst], [g for _, g in lst]


def negacyclic_py(a, b, q):
    n = len(a)
    r = [0] * n
    for i in range(n):
        for j in range(n):
            if i + j < n:
                r[i + j] += int(a[i]) * int(b[j])
            else:
                r[i + j - n] -= int(a[i]) * int(b[j])
    return [x % q for x in r]


def crt(moduli, residues):
    """Python-int CRT of one coefficient (independent of the oracle's Garner)."""
    Q = 1
    for m in moduli:
        Q *= m
    x = 0
    for m, r in zip(moduli, residues):
        part = Q // m
        x += int(r) * part * pow(part, -1, m)
    return x % Q, Q


# ------------------------------------------------------------------ NTT (src/math/ntt.rs)
def test_ntt_round_trip_kat():
    # ntt.rs:195-204
    plan = O.ntt_plan(8, 998244353, 3)
    v = np.arange(1, 9, dtype=np.uint64)
    assert list(O.ntt_inverse(plan, O.ntt_forward(plan, v))) == list(range(1, 9))
    # tests/math/ntt_tests.rs:4-16
    plan = O.ntt_plan(16, 998244353, 3)
    fib = [0, 1, 2, 3, 5, 8, 13, 21, 34, 55, 89, 144, 233, 377, 610, 987]
    assert list(O.ntt_inverse(plan, O.ntt_forward(plan, fib))) == fib


def test_ntt_plan_rejections():
    # ntt.rs:206-211, tests/math/ntt_tests.rs:18-28
    with pytest.raises(O.OracleError, match="2 \\* size must divide modulus - 1"):
        O.ntt_plan(8, 41, 6)
    with pytest.raises(O.OracleError):
        O.ntt_plan(12, 998244353, 3)
    with pytest.raises(O.OracleError):
        O.ntt_plan(7, 998244353, 3)


@pytest.mark.parametrize("n,q,g", [(8, 998244353, 3), (64, 998244353, 3), (32, 1125899906826241, 22),
                                   (128, 1125899908022273, 3)])
def test_ntt_forward_matches_closed_form_dft(n, q, g):
    """X_k = sum_j x_j w^{jk}, w = g^{(q-1)/N}: brute force O(N^2) with Python ints."""
    plan = O.ntt_plan(n, q, g)
    w = pow(g, (q - 1) // n, q)
    assert plan.root == w and plan.psi == pow(g, (q - 1) // (2 * n), q)
    rng = random.Random(n)
    x = [rng.randrange(q) for _ in range(n)]
    X = O.ntt_forward(plan, x)
    expect = [sum(x[j] * pow(w, j * k, q) for j in range(n)) % q for k in range(n)]
    assert [int(v) for v in X] == expect
    back = O.ntt_inverse(plan, X)
    ninv = pow(n, -1, q)
    expect_inv = [sum(expect[k] * pow(w, -j * k, q) for k in range(n)) * ninv % q for j in range(n)]
    assert [int(v) for v in back] == expect_inv == x


# ------------------------------------------------------------------ ModularPolynomial (src/poly/mod.rs)
def test_poly_kats():
    # poly/mod.rs:204-211
    assert list(O.poly_add([14, 1, 9, 0], [5, 16, 12, 3], 17)) == [2, 0, 4, 3]
    assert list(O.poly_sub([14, 1, 9, 0], [5, 16, 12, 3], 17)) == [9, 2, 14, 14]
    # tests/math/poly_tests.rs:5-8
    assert list(O.poly_new([17, 34, 35, 99], 17)) == [0, 0, 1, 14]
    # tests/math/poly_tests.rs:11-20
    a = O.poly_new([96, 10, 5, 42, 0, 1, 2, 3], 97)
    b = O.poly_new([5, 90, 90, 70, 1, 2, 3, 4], 97)
    assert list(O.poly_sub(O.poly_add(a, b, 97), b, 97)) == list(a)
    # poly/mod.rs:227-234
    assert list(O.poly_galois([1, 2, 3, 4], 3, 17)) == [1, 4, 14, 2]


def test_multiply_ntt_kat():
    # poly/mod.rs:214-224
    q = 998244353
    plan = O.ntt_plan(8, q, 3)
    got = O.poly_multiply_ntt(list(range(1, 9)), list(range(8, 0, -1)), plan)
    assert list(got) == [998244193, 998244243, 998244297, 0, 56, 110, 160, 204]
    assert list(got) == negacyclic_py(list(range(1, 9)), list(range(8, 0, -1)), q)
    # tests/math/poly_tests.rs:23-34
    lhs, rhs = [11, 22, 33, 44, 55, 66, 77, 88], [3, 1, 4, 1, 5, 9, 2, 6]
    assert list(O.poly_multiply_ntt(lhs, rhs, plan)) == list(O.poly_multiply_naive(lhs, rhs, q))


@pytest.mark.parametrize("n,q,g", [(16, 998244353, 3), (64, 1125899903827969, 11), (256, 1125899908022273, 3)])
def test_multiply_ntt_and_naive_match_python_convolution(n, q, g):
    plan = O.ntt_plan(n, q, g)
    rng = random.Random(q + n)
    a = [rng.randrange(q) for _ in range(n)]
    b = [rng.randrange(q) for _ in range(n)]
    want = negacyclic_py(a, b, q)
    assert [int(v) for v in O.poly_multiply_ntt(a, b, plan)] == want
    assert [int(v) for v in O.poly_multiply_naive(a, b, q)] == want


def test_galois_matches_definition():
    """sigma_g(a)(X) = a(X^g) mod X^N+1: coefficient i moves to i*g mod 2N with a sign flip past N."""
    n, q = 32, 1125899906826241
    rng = random.Random(5)
    a = [rng.randrange(q) for _ in range(n)]
    for g in (3, 5, 25, 2 * n - 1):
        want = [0] * n
        for i, c in enumerate(a):
            k = i * g % (2 * n)
            if k < n:
                want[k] = (want[k] + c) % q
            else:
                want[k - n] = (want[k - n] - c) % q
        assert [int(v) for v in O.poly_galois(a, g, q)] == want


# ------------------------------------------------------------------ RNS layer
SAMPLE = ([998244353, 1004535809, 469762049], [3, 3, 3])  # rns/tests.rs:14-24


def test_rns_sample_context_kats():
    b = O.Basis(8, *SAMPLE)
    lhs = np.zeros((3, 8), dtype=np.uint64)
    rhs = np.zeros((3, 8), dtype=np.uint64)
    lhs[:, :4] = [10, 20, 30, 40]
    rhs[:, :4] = [3, 4, 5, 6]
    s = O.rns_add(b, lhs, rhs)
    assert E.reconstruct(b.moduli, s)[:4] == [13, 24, 35, 46]  # rns/tests.rs:125-136
    lhs[:, :4] = [1, 2, 3, 4]
    rhs[:, :4] = [4, 3, 2, 1]
    prod = O.rns_multiply_ntt(b, lhs, rhs)  # rns/tests.rs:139-154
    for l, q in enumerate(b.moduli):
        assert list(prod[l]) == list(O.poly_multiply_naive(lhs[l], rhs[l], q))
    # rns/tests.rs:314-321
    assert [O.galois_element_for_rotation(8, s) for s in (0, 1, 2, -1)] == [1, 5, 9, 13]


def test_basis_rejects_duplicates_and_bad_primes():
    with pytest.raises(O.OracleError, match="duplicate RNS modulus 998244353"):
        O.Basis(8, [998244353, 998244353], [3, 3])
    with pytest.raises(O.OracleError, match="invalid RNS prime 41"):
        O.Basis(8, [41], [6])


def test_total_modulus_bits_and_digits():
    for n in (1 << 12, 1 << 16):
        ps, _ = chain("below_2^50", n, 24)
        for L in (1, 3, 8, 24):
            Q = 1
            for p in ps[:L]:
                Q *= p
            assert O.lib().orc_total_modulus_bits(np.array(ps[:L], dtype=np.uint64).ctypes.data_as(O.u64p),
                                                  L) == Q.bit_length()
    assert O.num_decomp_digits(1200) == 40 and O.num_decomp_digits(150) == 5 and O.num_decomp_digits(1201) == 41


def _random_poly(b, rng):
    return np.stack([np.array([rng.randrange(q) for _ in range(b.n)], dtype=np.uint64) for q in b.moduli])


def test_rescale_matches_bigint_formula():
    """c_i' = (c_i - centred(c_last)) * q_last^{-1} mod q_i  (ops.rs:135-186)."""
    ps, gs = chain("above_2^50", 1 << 12, 5)
    b = O.Basis(64, ps, gs)
    rng = random.Random(1)
    x = _random_poly(b, rng)
    got = O.rescale(b, x)
    ql = ps[-1]
    for c in range(64):
        last = int(x[-1, c])
        cen = last if last <= ql // 2 else last - ql
        for i in range(4):
            want = (int(x[i, c]) - cen) * pow(ql, -1, ps[i]) % ps[i]
            assert int(got[i, c]) == want


def balanced_digits_py(x_signed, D):
    """Balanced base-2^30 digits in [-2^29, 2^29) of x_signed mod 2^{30D} (independent method)."""
    v = x_signed % (1 << (30 * D))
    out = []
    for _ in range(D):
        d = v & ((1 << 30) - 1)
        v >>= 30
        if d >= 1 << 29:
            d -= 1 << 30
            v += 1
        out.append(d)
    return out


@pytest.mark.parametrize("kind,n,L", [("below_2^50", 1 << 12, 3), ("above_2^50", 1 << 14, 7),
                                      ("below_2^50", 1 << 16, 24)])
def test_decomposition_matches_bigint(kind, n, L):
    ps, gs = chain(kind, n, L)
    b = O.Basis(64, ps, gs) if False else None
    # small degree basis with the same primes (the primes support any N dividing their 2N)
    b = O.Basis(64, ps, gs)
    rng = random.Random(L)
    x = _random_poly(b, rng)
    # edge coefficients: 0, 1, Q-1, floor(Q/2), floor(Q/2)+1
    Q = 1
    for p in ps:
        Q *= p
    for c, v in enumerate([0, 1, Q - 1, Q // 2, Q // 2 + 1, (1 << 29), Q - (1 << 29)]):
        x[:, c] = [v % p for p in ps]
    D = O.num_decomp_digits(b.total_modulus_bits())
    sd = O.decompose_signed(b, x, D)
    res = O.decompose(b, x, D)
    top = (((1 << 29) - 1) * ((1 << (30 * D)) - 1)) // ((1 << 30) - 1)  # largest balanced D-digit value
    for c in range(64):
        v, _ = crt(ps, x[:, c])
        xs = v if v <= Q // 2 else v - Q
        want = balanced_digits_py(xs, D)
        assert [int(d) for d in sd[:, c]] == want
        rec = sum(d << (30 * i) for i, d in enumerate(want))
        assert (rec - xs) % (1 << (30 * D)) == 0
        # exact recomposition whenever xs fits D balanced digits; otherwise the reference drops the
        # final carry (keyswitching.rs:452, release builds) and so does the oracle
        assert (rec == xs) == (xs <= top)
        for i, d in enumerate(want):
            assert -(1 << 29) <= d < (1 << 29)
            for l, p in enumerate(ps):
                assert int(res[i, l, c]) == d % p


def test_decomposition_drops_final_carry_like_release_builds():
    """keyswitching.rs:452: with fewer digits than the value needs, the top carry is dropped."""
    ps, gs = chain("below_2^50", 1 << 12, 3)
    b = O.Basis(8, ps, gs)
    Q = ps[0] * ps[1] * ps[2]
    x = np.zeros((3, 8), dtype=np.uint64)
    vals = [Q // 2, Q // 2 + 1, 123456789 << 60, Q - (987654321 << 61)]
    for c, v in enumerate(vals):
        x[:, c] = [v % p for p in ps]
    sd = O.decompose_signed(b, x, 2)  # only 2 of the 5 digits
    for c, v in enumerate(vals):
        xs = v if v <= Q // 2 else v - Q
        assert [int(d) for d in sd[:, c]] == balanced_digits_py(xs, 2)


def test_lift_and_mod_down_match_bigint():
    ps, gs = chain("below_2^50", 1 << 12, 6)
    q = O.Basis(32, ps[:4], gs[:4])
    e = O.Basis(32, ps, gs)
    rng = random.Random(9)
    x = _random_poly(q, rng)
    lifted = O.lift(q, e, x)
    Q = ps[0] * ps[1] * ps[2] * ps[3]
    for c in range(32):
        v, _ = crt(ps[:4], x[:, c])
        cen = v if v <= Q // 2 else v - Q
        assert [int(t) for t in lifted[:, c]] == [cen % p for p in ps]
    y = _random_poly(e, rng)
    md = O.mod_down(q, e, y)
    P = ps[4] * ps[5]
    for c in range(32):
        xp, _ = crt(ps[4:], y[4:, c])
        cp = xp if xp <= P // 2 else xp - P
        for i in range(4):
            assert int(md[i, c]) == (int(y[i, c]) - cp) * pow(P, -1, ps[i]) % ps[i]


# ------------------------------------------------------------------ scheme-level invariants
def _keys(b, seed, hamming=None, sigma=3.19):
    rng = O.Rng(seed)
    s = O.sample_sparse_ternary(b, rng, hamming) if hamming else O.sample_ternary(b, rng)
    pk, err = O.keygen_public(b, rng, s, sigma)
    return rng, s, pk, err


def test_public_key_relation_exact():
    b = O.Basis(8, [998244353, 1004535809, 469762049], [3, 3, 3])
    rng, s, pk, err = _keys(b, 11, hamming=3)
    assert int((O.sample_sparse_ternary(b, O.Rng(2), 3)[0] != 0).sum()) == 3  # rns/tests.rs:557-569
    rel = O.rns_add(b, pk[0], O.rns_multiply_ntt(b, pk[1], s))  # rns/tests.rs:571-582
    assert np.array_equal(rel, err)


def test_noiseless_encryption_exact():
    """rns/tests.rs:652-677: zero errors and mask (1,0,1,0...) decrypt to m + e*mask exactly."""
    b = O.Basis(8, [998244353, 1004535809, 469762049], [3, 3, 3])
    rng, s, pk, err = _keys(b, 3, hamming=3)
    m = O.sample_uniform(b, rng)
    mask = b.zeros()
    mask[:, 0] = 1
    mask[:, 2] = 1
    z = b.zeros()
    ct = O.encrypt_with_samples(b, m, pk, mask, z, z)
    assert np.array_equal(O.decrypt(b, ct, s), O.rns_add(b, m, O.rns_multiply_ntt(b, err, mask)))


def test_tensor_identity_exact():
    """Dec2(ct_a (x) ct_b) = Dec(ct_a) * Dec(ct_b) in R_Q, exactly (ops.rs:84-113)."""
    ps, gs = chain("below_2^50", 1 << 12, 3)
    b = O.Basis(64, ps, gs)
    rng = O.Rng(7)
    s = O.sample_ternary(b, rng)
    a = np.stack([O.sample_uniform(b, rng), O.sample_uniform(b, rng)])
    c = np.stack([O.sample_uniform(b, rng), O.sample_uniform(b, rng)])
    quad = O.multiply_ciphertexts(b, a, c)
    assert np.array_equal(O.decrypt_quadratic(b, quad, s),
                          O.rns_multiply_ntt(b, O.decrypt(b, a, s), O.decrypt(b, c, s)))
    # c1 = a0 b1 + a1 b0 with the schoolbook product per limb
    for l, q in enumerate(ps):
        want = [(u + v) % q for u, v in zip(negacyclic_py(a[0, l], c[1, l], q), negacyclic_py(a[1, l], c[0, l], q))]
        assert [int(t) for t in quad[1, l]] == want


@pytest.mark.parametrize("hybrid", [False, True])
def test_noiseless_keyswitch_recovers_source_product(hybrid):
    """rns/tests.rs:437-456 (hybrid) and the Q-only analogue: decrypt(keyswitch(x)) == x * s' exactly."""
    if hybrid:
        p = (realistic := __import__("rnsckks.params", fromlist=["x"]).realistic_8_level())
        n = 256
        q = O.Basis(n, p.primes, p.roots)
        e = O.Basis(n, p.primes + p.aux_primes, p.roots + p.aux_roots)
    else:
        ps, gs = chain("below_2^50", 1 << 12, 3)
        n = 128
        q = O.Basis(n, ps, gs)
        e = None
    rng, s, pk, _ = _keys(q, 21, hamming=16)
    g = O.galois_element_for_rotation(n, 1)
    src = O.rns_galois(q, s, g)
    if hybrid:
        ksk = O.keygen_hybrid(q, e, rng, src, s, noiseless=True)
    else:
        ksk = O.keygen_qonly(q, rng, src, pk, noiseless=True)
    x = q.zeros()
    x[:, :4] = [1, 2, 3, 4]
    y = O.sample_uniform(q, rng)
    for inp in (x, y):
        ks = O.keyswitch(q, e, inp, ksk)
        assert np.array_equal(O.decrypt(q, ks, s), O.rns_multiply_ntt(q, inp, src))


def galois_py(a, g, q):
    """X -> X^g on one limb by the definition (coefficient i moves to i*g mod 2N, negated past N)."""
    n = len(a)
    out = [0] * n
    for i, v in enumerate(a):
        k = (i * g) % (2 * n)
        if k < n:
            out[k] = (out[k] + int(v)) % q
        else:
            out[k - n] = (out[k - n] - int(v)) % q
    return out


@pytest.mark.parametrize("hybrid,steps", [(False, 1), (False, -3), (True, 2), (True, "conj")])
def test_noiseless_rotation_decrypts_to_the_rotated_plaintext(hybrid, steps):
    """rotate (context/rotation.rs:104-127) with a noiseless key for s(X^g): decrypt(rotate(ct)) equals
    sigma_g(decrypt(ct)) exactly, sigma_g applied by its definition in Python integers."""
    if hybrid:
        p = __import__("rnsckks.params", fromlist=["x"]).realistic_8_level()
        n = 128
        q = O.Basis(n, p.primes[:3], p.roots[:3])
        e = O.Basis(n, p.primes[:3] + p.aux_primes, p.roots[:3] + p.aux_roots)
    else:
        ps, gs = chain("below_2^50", 1 << 12, 3)
        n = 64
        q = O.Basis(n, ps, gs)
        e = None
    rng, s, pk, _ = _keys(q, 33, hamming=8)
    g = 2 * n - 1 if steps == "conj" else O.galois_element_for_rotation(n, steps)
    src = O.rns_galois(q, s, g)
    ksk = O.keygen_hybrid(q, e, rng, src, s, noiseless=True) if hybrid else \
        O.keygen_qonly(q, rng, src, pk, noiseless=True)
    ct = np.stack([O.sample_uniform(q, rng), O.sample_uniform(q, rng)])
    rot = O.rotate(q, e, ct, g, ksk)
    dec = O.decrypt(q, ct, s)
    got = O.decrypt(q, rot, s)
    for l, m in enumerate(q.moduli):
        assert [int(v) for v in got[l]] == galois_py(dec[l], g, m), f"limb {l}"


def test_multiply_plain_matches_schoolbook():
    """multiply_plain (rns/ops.rs:54-66): each component times the plaintext, by a Python-int schoolbook
    negacyclic product per limb."""
    ps, gs = chain("below_2^50", 1 << 12, 3)
    n = 32
    b = O.Basis(n, ps, gs)
    rng = O.Rng(12)
    ct = np.stack([O.sample_uniform(b, rng), O.sample_uniform(b, rng)])
    pt = O.sample_uniform(b, rng)
    out = O.multiply_plain(b, ct, pt)
    assert out.shape == ct.shape
    for c in range(2):
        for l, m in enumerate(ps):
            assert [int(v) for v in out[c, l]] == negacyclic_py(ct[c, l], pt[l], m), (c, l)


def test_keyswitch_rejects_short_key():
    ps, gs = chain("below_2^50", 1 << 12, 3)
    q = O.Basis(16, ps, gs)
    with pytest.raises(O.OracleError):
        O.keyswitch(q, None, q.zeros(), np.zeros((4, 2, 3, 16), dtype=np.uint64))


# ------------------------------------------------------------------ approximate CKKS (decode error)
def test_encode_decode_round_trip():
    ps, gs = chain("below_2^50", 1 << 12, 3)
    slots = np.array([0.125 - 0.25j, -0.75 + 0.5j, 0.3125j, 0.625 - 0.4375j])
    pt = E.encode(ps, 8, 40, slots)
    assert np.abs(E.decode(ps, 8, 40, pt) - slots).max() < 1e-3  # rns/tests.rs:585-600


def test_mult_relin_decode_error_config1_shape():
    """The crate's Hadamard-product flow at N=2^12, 3 primes < 2^50, Q-only keys, scale 2^40:
    decode(decrypt(relin(mult(enc a, enc b)))) ~ a*b at scale 2^80 (examples rns_hadamard_product.rs)."""
    n = 1 << 12
    ps, gs = chain("below_2^50", n, 3)
    q = O.Basis(n, ps, gs)
    rs = np.random.default_rng(0)
    za = rs.uniform(-1, 1, n // 2) + 1j * rs.uniform(-1, 1, n // 2)
    zb = rs.uniform(-1, 1, n // 2) + 1j * rs.uniform(-1, 1, n // 2)
    rng, s, pk, _ = _keys(q, 0, sigma=3.2)
    rlk = O.keygen_qonly(q, rng, O.rns_multiply_ntt(q, s, s), pk, sigma=3.2)
    cts = []
    for z in (za, zb):
        u, e0, e1 = O.sample_ternary(q, rng), O.sample_error(q, rng, 3.2), O.sample_error(q, rng, 3.2)
        cts.append(O.encrypt_with_samples(q, E.encode(ps, n, 40, z), pk, u, e0, e1))
    quad = O.multiply_ciphertexts(q, cts[0], cts[1])
    ct = O.relinearize(q, None, quad, rlk)
    got = E.decode(ps, n, 80, O.decrypt(q, ct, s))
    assert np.abs(got - za * zb).max() < 1.8e-2
