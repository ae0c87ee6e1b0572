"""GPU parity over a seeded sweep of shapes: each case draws a ring degree, a level, a key kind
(hybrid with the realistic_8_level aux primes, or Q-only) and a keyswitching operation, and checks
the C-ABI result against the CPU oracle bit for bit.  The shapes cover levels below the context's
full chain (the key and ciphertext at a prefix of the primes, context/ops.rs:97-122,
context/rotation.rs:104-135), single-limb ciphertexts and ring degrees drawn from 2^11 (one-pass transform) .. 2^16.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from rnsckks import capi as G
from rnsckks import params as PR

from test_gpu_parity import chain, dev, empty, rand_poly

OPS = ("relinearize", "keyswitch", "rotate", "conjugate", "mult_relinearized")
SEEDS = int(os.environ.get("RCK_SWEEP_SEEDS", "18"))  # a longer sweep: RCK_SWEEP_SEEDS=1500


def _case(seed):
    rs = np.random.default_rng(1000 + seed)
    hybrid = bool(rs.integers(0, 2))
    logn = int(rs.integers(11, 16 if hybrid else 17))  # realistic_8_level's primes serve N <= 2^15
    limbs = int(rs.integers(1 if hybrid else 2, 9))
    op = OPS[int(rs.integers(0, len(OPS)))]
    return hybrid, 1 << logn, limbs, op, rs


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(SEEDS))
def test_seeded_shape_sweep_matches_oracle(seed):
    hybrid, n, L, op, rs = _case(seed)
    if op == "mult_relinearized" and L < 2:
        L = 2  # the rescale drops a limb
    if hybrid:
        params = PR.realistic_8_level()
        q, g = params.primes, params.roots
        ctx = G.Context(poly_degree=n, primes=q, roots=g, aux_primes=params.aux_primes,
                        aux_roots=params.aux_roots, scale_bits=params.scale_bits)
        qb = O.Basis(n, q[:L], g[:L])
        eb = O.Basis(n, q[:L] + params.aux_primes, g[:L] + params.aux_roots)
        K = len(params.aux_primes)
    else:
        q, g = chain("below_2^50", 1 << 16, 8)
        ctx = G.Context(poly_degree=n, primes=q, roots=g, scale_bits=40)
        qb, eb, K = O.Basis(n, q[:L], g[:L]), None, 0
    print(f"[sweep] seed={seed} hybrid={hybrid} n={n} L={L} op={op}", flush=True)
    rng = O.Rng(seed)
    s = O.sample_ternary(qb, rng)
    pk, _ = O.keygen_public(qb, rng, s)

    def keygen(source):
        if hybrid:
            return O.keygen_hybrid(qb, eb, rng, source, s)
        return O.keygen_qonly(qb, rng, source, pk)

    out = empty((2, L, n))
    if op in ("rotate", "conjugate"):
        if op == "rotate":
            steps = int(rs.integers(1, n // 2)) * (1 if rs.integers(0, 2) else -1)
            gel = O.galois_element_for_rotation(n, steps)
        else:
            gel = 2 * n - 1
        rk = keygen(O.rns_galois(qb, s, gel))
        drk = ctx.ksk_create(rk, L - 1, K, gel)
        ct = rand_poly(qb.moduli, n, rs, 2)
        assert ctx.rotate(dev(ct), L, 45, drk, out) == 45
        assert np.array_equal(out.numpy(), O.rotate(qb, eb, ct, gel, rk))
        return
    rlk = keygen(O.rns_multiply_ntt(qb, s, s))
    dk = ctx.ksk_create(rlk, L - 1, K)
    if op == "relinearize":
        quad = rand_poly(qb.moduli, n, rs, 3)
        assert ctx.relinearize(dev(quad), L, 90, dk, out) == 90
        assert np.array_equal(out.numpy(), O.relinearize(qb, eb, quad, rlk))
    elif op == "keyswitch":
        poly = rand_poly(qb.moduli, n, rs)
        assert ctx.keyswitch(dev(poly), L, dk, out) == L  # keyswitching.rs:156 scale quirk
        assert np.array_equal(out.numpy(), O.keyswitch(qb, eb, poly, rlk))
    else:
        lhs, rhs = rand_poly(qb.moduli, n, rs, 2), rand_poly(qb.moduli, n, rs, 2)
        o = empty((2, L - 1, n))
        scale = ctx.mult_relinearized(dev(lhs), dev(rhs), L, 45, 45, dk, o)
        assert scale == 90 - (int(qb.moduli[-1]).bit_length() - 1)
        want = O.rescale(qb, O.relinearize(qb, eb, O.multiply_ciphertexts(qb, lhs, rhs), rlk))
        assert np.array_equal(o.numpy(), want)


def test_sweep_draws_every_operation_and_key_kind():
    """The seeds above cover each operation and both key kinds (a CPU-side check of the draw)."""
    drawn = [_case(seed)[:4] for seed in range(18)]
    assert {op for *_, op in drawn} == set(OPS)
    assert {h for h, *_ in drawn} == {True, False}
