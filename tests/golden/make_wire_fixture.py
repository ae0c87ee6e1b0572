"""Writes tests/golden/wire_rck1.json: golden "RCK1" byte images of a ciphertext and a hybrid
keyswitching key at N=16, laid out here with `struct` from the format's written description
(little-endian header of 56 bytes, then canonical u64 residues limb-major), independently of the
library.  The GPU test serializes the same residues through the C ABI and compares byte for byte;
the CPU test checks the fixture's own header against its fields.

  python tests/golden/make_wire_fixture.py
"""
import json
import os
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "rns-ckks-rust_amd"))

from rnsckks import params as PR  # noqa: E402

N = 16
Q = PR.primes_below(30, N, 3)            # three Q primes q = 1 mod 2N below 2^30
P = PR.primes_below(30, N, 1, skip=3)    # one aux prime, the next of the same sequence
SCALE = 20
GALOIS = 5


def fnv1a(words):
    h = 1469598103934665603
    for w in words:
        for b in range(8):
            h ^= (w >> (8 * b)) & 0xFF
            h = (h * 1099511628211) & (2**64 - 1)
    return h


def header(kind, count, nq, np_, scale, galois, payload):
    fp = fnv1a([N] + Q[:nq] + P[:np_])
    return b"RCK1" + struct.pack("<7I3Q", 1, kind, N.bit_length() - 1, count, nq, np_, scale, galois, fp, payload)


def residues(moduli, groups, seed):
    """Deterministic canonical residues [groups][limbs][N]: x = (seed * 7919 + idx * 104729) mod q."""
    out = []
    for g in range(groups):
        for l, q in enumerate(moduli):
            out.append([((seed + g) * 7919 + (l * N + i) * 104729 + i * i) % q for i in range(N)])
    return out


def main():
    ct = residues(Q, 2, 11)                      # a ciphertext [c0, c1] at 3 limbs
    ct_bytes = header(1, 2, 3, 0, SCALE, 0, 2 * 3 * N * 8) + b"".join(struct.pack(f"<{N}Q", *r) for r in ct)
    digits = 3                                   # key digits (any count works for the format)
    key = residues(Q + P, digits * 2, 29)        # [digit][comp][Q limbs + aux limbs][N]
    key_bytes = header(2, digits, 3, 1, 2, GALOIS, digits * 2 * 4 * N * 8) + \
        b"".join(struct.pack(f"<{N}Q", *r) for r in key)
    fixture = {
        "N": N, "primes": Q, "aux_primes": P, "roots": [PR.smallest_primitive_root(q) for q in Q],
        "aux_roots": [PR.smallest_primitive_root(q) for q in P],
        "ciphertext": {"scale_bits": SCALE, "limbs": 3, "residues": ct, "bytes_hex": ct_bytes.hex()},
        "key": {"digits": digits, "level": 2, "aux": 1, "galois": GALOIS, "residues": key,
                "bytes_hex": key_bytes.hex()},
    }
    with open(os.path.join(HERE, "wire_rck1.json"), "w") as f:
        json.dump(fixture, f, indent=1)
        f.write("\n")
    print(f"wrote wire_rck1.json: ciphertext {len(ct_bytes)} bytes, key {len(key_bytes)} bytes")


if __name__ == "__main__":
    main()
