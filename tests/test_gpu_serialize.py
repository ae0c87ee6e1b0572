"""Versioned binary format: ciphertexts and keys round-trip bit for bit, a deserialized key
keyswitches exactly like the original, and foreign / damaged data is rejected."""
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


def test_round_trips_and_rejections():
    n, L = 1 << 12, 5
    p = PR.config_batched(n=1 << 15, limbs=L, aux=2)
    ctx = G.Context(poly_degree=n, primes=p.primes, roots=p.roots, aux_primes=p.aux_primes, aux_roots=p.aux_roots)
    rs = np.random.default_rng(2)
    ct = uniform(p.primes[:4], n, rs, 2)
    data = ctx.serialize_polys(G.DeviceBuffer.from_numpy(ct), 2, 4, 0, 40)
    assert data[:4] == b"RCK1" and len(data) == 56 + ct.nbytes
    back, meta = ctx.deserialize_polys(data)
    assert meta == {"count": 2, "nq": 4, "np": 0, "scale_bits": 40}
    assert np.array_equal(back.numpy().reshape(ct.shape), ct)
    # keys: coefficient form round trip, and the restored key switches identically
    D = ctx.num_decomp_digits(L)
    key = np.stack([uniform(p.primes + p.aux_primes, n, rs, 2) for _ in range(D)])
    k = ctx.ksk_create(key, L - 1, 2, 5)
    kd = ctx.ksk_serialize(k)
    k2 = ctx.ksk_deserialize(kd)
    assert (k2.digits, k2.level, k2.aux, k2.galois) == (D, L - 1, 2, 5)
    assert np.array_equal(ctx.ksk_export(k2), key)
    quad = G.DeviceBuffer.from_numpy(uniform(p.primes, n, rs, 3))
    o1, o2 = G.DeviceBuffer((2, L, n)), G.DeviceBuffer((2, L, n))
    ctx.relinearize(quad, L, 80, k, o1)
    ctx.relinearize(quad, L, 80, k2, o2)
    assert np.array_equal(o1.numpy(), o2.numpy())
    # rejections
    with pytest.raises(G.RckError, match="bad magic"):
        ctx.deserialize_polys(b"XXXX" + data[4:])
    with pytest.raises(G.RckError, match="truncated"):
        ctx.deserialize_polys(data[:100])
    with pytest.raises(G.RckError, match="not a key"):
        ctx.ksk_deserialize(data)
    other = G.Context(poly_degree=n, primes=p.primes[1:], roots=p.roots[1:])
    with pytest.raises(G.RckError, match="another prime chain"):
        other.deserialize_polys(data)
    small = G.Context(poly_degree=n // 2, primes=p.primes, roots=p.roots)
    with pytest.raises(G.RckError, match="another ring degree"):
        small.deserialize_polys(data)


def _golden():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "wire_rck1.json")) as f:
        return json.load(f)


def test_serialized_bytes_match_the_golden_layout():
    """The library's bytes equal the golden RCK1 images written from the format description alone
    (tests/golden/make_wire_fixture.py), and the golden images load back to the same residues."""
    fx = _golden()
    n = fx["N"]
    ctx = G.Context(poly_degree=n, primes=fx["primes"], roots=fx["roots"], aux_primes=fx["aux_primes"],
                    aux_roots=fx["aux_roots"])
    c = fx["ciphertext"]
    ct = np.array(c["residues"], dtype=np.uint64).reshape(2, c["limbs"], n)
    got = ctx.serialize_polys(G.DeviceBuffer.from_numpy(ct), 2, c["limbs"], 0, c["scale_bits"])
    assert got.hex() == c["bytes_hex"]
    back, meta = ctx.deserialize_polys(bytes.fromhex(c["bytes_hex"]))
    assert meta == {"count": 2, "nq": c["limbs"], "np": 0, "scale_bits": c["scale_bits"]}
    assert np.array_equal(back.numpy().reshape(ct.shape), ct)
    k = fx["key"]
    limbs = k["level"] + 1 + k["aux"]
    key = np.array(k["residues"], dtype=np.uint64).reshape(k["digits"], 2, limbs, n)
    dk = ctx.ksk_create(key, k["level"], k["aux"], k["galois"])
    assert ctx.ksk_serialize(dk).hex() == k["bytes_hex"]
    dk2 = ctx.ksk_deserialize(bytes.fromhex(k["bytes_hex"]))
    assert (dk2.digits, dk2.level, dk2.aux, dk2.galois) == (k["digits"], k["level"], k["aux"], k["galois"])
    assert np.array_equal(ctx.ksk_export(dk2), key)


def test_non_canonical_residues_are_rejected():
    """A payload whose header and fingerprint are intact but holds a residue at or above its modulus is
    refused before upload, for polynomial sets and keys."""
    fx = _golden()
    n = fx["N"]
    ctx = G.Context(poly_degree=n, primes=fx["primes"], roots=fx["roots"], aux_primes=fx["aux_primes"],
                    aux_roots=fx["aux_roots"])
    import struct
    raw = bytearray(bytes.fromhex(fx["ciphertext"]["bytes_hex"]))
    q1 = fx["primes"][1]
    raw[56 + (1 * n + 5) * 8: 56 + (1 * n + 6) * 8] = struct.pack("<Q", q1)  # limb 1, coefficient 5 := q_1
    with pytest.raises(G.RckError, match="not below its modulus"):
        ctx.deserialize_polys(bytes(raw))
    raw = bytearray(bytes.fromhex(fx["key"]["bytes_hex"]))
    pa = fx["aux_primes"][0]
    off = 56 + ((1 * 2 + 1) * 4 + 3) * n * 8  # digit 1, component 1, the aux limb, coefficient 0
    raw[off: off + 8] = struct.pack("<Q", pa + 7)
    with pytest.raises(G.RckError, match="not below its modulus"):
        ctx.ksk_deserialize(bytes(raw))
