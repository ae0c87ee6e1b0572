"""Baby-step giant-step diagonal transform restated over the oracle (TEST INFRASTRUCTURE ONLY).

Follows src/ckks/bootstrapping/linear_transform.rs:22-253 (and src/algorithms/packed_diagonal.rs):
baby rotations of the input, per giant step the sum of mult_plain(baby_b, diag[g][b]) left
unrescaled, the giant rotation of every inner sum but the first, the sum, one rescale.  The plaintext
diagonals are inputs (coefficient form), so the library's encodings can be checked through the same
arithmetic bit for bit; packed_bsgs_diagonal_slots is restated for the encoding check.

Parity pins: rotation steps against the crate's unit tests (:309-320); the whole transform against
the library bit for bit; decryption against FP64 matvec within the crate's tolerance (:322-376).
"""
import numpy as np

from . import oracle as O


def baby_step(dim):
    """diagonal_transform_bsgs_baby_step (:146-152)."""
    b = 1
    while b * b < dim:
        b += 1
    return b


def rotation_steps(dim):
    """diagonal_transform_bsgs_rotation_steps (:27-47)."""
    if dim <= 0:
        raise ValueError("transform dimension must be positive")
    if dim == 1:
        return []
    bs = baby_step(dim)
    gc = -(-dim // bs)
    steps = set(range(1, min(bs, dim)))
    steps.update(g * bs for g in range(1, gc) if g * bs < dim)
    return sorted(steps)


def packed_bsgs_diagonal_slots(matrix, step, shift, slots):
    """packed_bsgs_diagonal_slots (:255-280)."""
    dim = len(matrix)
    shift %= dim
    out = np.empty(slots)
    for slot in range(slots):
        row = slot % dim
        src = (row + dim - shift) % dim
        out[slot] = matrix[src][(step + src) % dim]
    return out


def apply(n, primes, roots, aux, aux_roots, ct, scale, diags, keys):
    """apply_diagonal_linear_transform (:118-144).  diags[g][b]: coefficient-form plaintext or None;
    keys: {step: coefficient-form rotation key}.  Returns (array[2][limbs-1][N], scale)."""
    limbs = ct.shape[1]
    q = O.Basis(n, primes[:limbs], roots[:limbs])
    e = O.Basis(n, primes[:limbs] + aux, roots[:limbs] + aux_roots) if aux else None
    bs = len(diags[0])
    baby = [ct]
    for b in range(1, bs):
        baby.append(O.rotate(q, e, ct, O.galois_element_for_rotation(n, b), keys[b]))
    acc = None
    for g, row in enumerate(diags):
        inner = None
        for b, d in enumerate(row):
            if d is None:
                continue
            t = O.multiply_plain(q, baby[b], d)
            inner = t if inner is None else np.stack([O.rns_add(q, inner[0], t[0]), O.rns_add(q, inner[1], t[1])])
        if inner is None:
            continue
        if g > 0:
            inner = O.rotate(q, e, inner, O.galois_element_for_rotation(n, g * bs), keys[g * bs])
        acc = inner if acc is None else np.stack([O.rns_add(q, acc[0], inner[0]), O.rns_add(q, acc[1], inner[1])])
    drop = int(primes[limbs - 1]).bit_length() - 1
    return O.rescale(q, acc), 2 * scale - drop
