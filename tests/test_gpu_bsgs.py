"""Batched baby-step giant-step diagonal transform on the device equals the oracle's restatement
of the crate's per-ciphertext algorithm bit for bit (same plaintext diagonals and keys), and
decrypts to the FP64 matrix-vector product within the crate's tolerance."""
import numpy as np
import pytest

from oracle import bsgs as OB
from oracle import encoding as E
from rnsckks import capi as G
from rnsckks import params as PR

pytestmark = pytest.mark.gpu


def uniform(moduli, n, rs, polys):
    out = np.empty((polys, len(moduli), n), dtype=np.uint64)
    for l, q in enumerate(moduli):
        out[:, l] = rs.integers(0, q, size=(polys, n), dtype=np.uint64)
    return out


@pytest.mark.parametrize("dim,batch", [(16, 3), (8, 2)])
def test_bsgs_matches_oracle_bit_exact(dim, batch):
    n, L = 1 << 12, 5
    p = PR.config_batched(n=1 << 15, limbs=L, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    rs = np.random.default_rng(dim)
    matrix = rs.uniform(-1, 1, (dim, dim))
    diag = G.Diag(ctx, matrix, L - 1, 40)
    assert (diag.baby_step, diag.giant_steps) == (OB.baby_step(dim), -(-dim // OB.baby_step(dim)))
    plains = diag.export()
    # the library's encodings are the oracle encoder's values (within one unit of rounding)
    for g in range(diag.giant_steps):
        for b in range(diag.baby_step):
            step = g * diag.baby_step + b
            if step >= dim:
                assert not plains[g, b].any()
                continue
            slots = OB.packed_bsgs_diagonal_slots(matrix, step, g * diag.baby_step, n // 2)
            want = E.encode_integer_coeffs(n, 2.0 ** 40, slots.astype(complex))
            got = E.centered(E.reconstruct(p.primes[:L], plains[g, b]), E.total_modulus(p.primes[:L]))
            assert max(abs(a - w) for a, w in zip(got, want)) <= 1
    keys_np, keys = {}, {}
    D = ctx.num_decomp_digits(L)
    for st in G.bsgs_rotation_steps(dim):
        keys_np[st] = np.stack([uniform(p.primes + p.aux_primes, n, rs, 2) for _ in range(D)])
        keys[st] = ctx.ksk_create(keys_np[st], L - 1, 2, G.galois_element_for_rotation(n, st))
    cts = uniform(p.primes, n, rs, 2 * batch).reshape(batch, 2, L, n)
    out = G.DeviceBuffer((batch, 2, L - 1, n))
    sc = ctx.bsgs_matvec(G.DeviceBuffer.from_numpy(cts), batch, L, 40, diag, keys, out)
    got = out.numpy()
    rows = [[plains[g, b] if g * diag.baby_step + b < dim else None for b in range(diag.baby_step)]
            for g in range(diag.giant_steps)]
    for v in range(batch):
        want, wsc = OB.apply(n, p.primes, p.roots, p.aux_primes, p.aux_roots, cts[v], 40, rows, keys_np)
        assert sc == wsc
        assert np.array_equal(got[v], want), f"ciphertext {v}"


def test_bsgs_reference_matrix_decrypts_within_tolerance():
    """packed_diagonal_matvec_matches_expected_on_realistic_params (packed_diagonal.rs:98-161) with
    device keys: the first 4 slots and their periodic repeat within 2e-2."""
    p = PR.realistic_8_level()
    n, L = p.poly_degree, len(p.primes)
    ctx = G.Context(p)
    g = G.Rng(64)
    s = ctx.sample_sparse_ternary(g, 64, L)
    pk = ctx.keygen_public(g, s, L)
    matrix = np.array([[1.0, 0.5, -0.25, 0.0], [0.0, -1.0, 0.75, 0.25], [0.5, 0.0, 1.25, -0.5],
                       [-0.25, 0.125, 0.0, 0.75]])
    vec = np.array([0.5, -0.25, 0.125, 0.75])
    keys = {st: ctx.keygen_rotation(g, s, st, L - 1) for st in G.bsgs_rotation_steps(4)}
    packed = np.tile(vec, n // 2 // 4)
    ct = ctx.encrypt(g, ctx.encode(packed, L, 45), pk, L, L)
    diag = G.Diag(ctx, matrix, L - 1, 45)
    out = G.DeviceBuffer((1, 2, L - 1, n))
    sc = ctx.bsgs_matvec(ct, 1, L, 45, diag, keys, out)
    got = ctx.decode(ctx.decrypt(out, L - 1, s), L - 1, sc).real
    want = matrix @ vec
    assert np.max(np.abs(got[:4] - want)) <= 2e-2
    assert np.max(np.abs(got[4:8] - want)) <= 2e-2


def test_bsgs_batch_decode_error():
    """dim 64 (baby step 8, 8 giant steps), 4 vectors at N=2^13, 6 limbs: every output block of every
    vector matches FP64 to 1e-4."""
    n, L, dim, B = 1 << 13, 6, 64, 4
    p = PR.config_batched(n=1 << 15, limbs=L, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    g = G.Rng(8)
    s = ctx.sample_ternary(g, L)
    pk = ctx.keygen_public(g, s, L)
    keys = {st: ctx.keygen_rotation(g, s, st, L - 1) for st in G.bsgs_rotation_steps(dim)}
    rs = np.random.default_rng(4)
    matrix = rs.uniform(-1, 1, (dim, dim))
    vecs = rs.uniform(-1, 1, (B, dim))
    cts = G.DeviceBuffer((B, 2, L, n))
    for v in range(B):
        ct = ctx.encrypt(g, ctx.encode(np.tile(vecs[v], n // 2 // dim), L, 40), pk, L, L)
        G.lib().rck_memcpy_dtod(cts.ptr + v * 2 * L * n * 8, ct.ptr, 2 * L * n * 8, None)
    diag = G.Diag(ctx, matrix, L - 1, 40)
    out = G.DeviceBuffer((B, 2, L - 1, n))
    sc = ctx.bsgs_matvec(cts, B, L, 40, diag, keys, out)
    P = 2 * (L - 1) * n
    for v in range(B):
        dec = ctx.decrypt(out.ptr + v * P * 8, L - 1, s)
        got = ctx.decode(dec, L - 1, sc).real
        want = np.tile(matrix @ vecs[v], n // 2 // dim)
        assert np.max(np.abs(got - want)) <= 1e-4, f"vector {v}"


def test_bsgs_errors():
    n, L = 1 << 12, 4
    p = PR.config_batched(n=1 << 15, limbs=L, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    with pytest.raises(G.RckError, match="power of two"):
        G.Diag(ctx, np.ones((6, 6)), L - 1, 40)
    diag = G.Diag(ctx, np.ones((4, 4)), L - 1, 40)
    ct = G.DeviceBuffer((1, 2, L, n))
    out = G.DeviceBuffer((1, 2, L, n))
    with pytest.raises(G.RckError, match="missing input rotation key for BSGS baby step 1"):
        ctx.bsgs_matvec(ct, 1, L, 40, diag, {}, out)
    with pytest.raises(G.RckError, match="must match the ciphertext scale"):
        ctx.bsgs_matvec(ct, 1, L, 45, diag, {}, out)
