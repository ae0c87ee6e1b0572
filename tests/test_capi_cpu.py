"""CPU-side checks of the C ABI library: it loads and exports every symbol include/*.h declares
(no compute calls: there is no GPU here), plus host-only logic of the package."""
import glob
import os
import re

import pytest

from rnsckks import capi
from rnsckks import params as PR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(rck_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("librnsckks.so not built")
    lib = capi.lib()
    names = declared_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(capi.EXPORTS) == names


def test_galois_element_for_rotation_host_only():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("librnsckks.so not built")
    # rns/tests.rs:314-321
    assert [capi.galois_element_for_rotation(8, s) for s in (0, 1, 2, -1)] == [1, 5, 9, 13]
    n = 1 << 15
    for s in (1, -1, 4096, -8192, 13):
        g = capi.galois_element_for_rotation(n, s)
        want = pow(5, s, 2 * n) if s >= 0 else pow(pow(5, -1, 2 * n), -s, 2 * n)
        assert g == want


def test_prime_chains_match_golden():
    import json
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "primes.json")))
    for n in (1 << 12, 1 << 14, 1 << 15, 1 << 16):
        below = PR.primes_below(50, n, 28)
        above = PR.primes_above(50, n, 28)
        assert below == [p for p, _ in gold["below_2^50"][str(n)]]
        assert above == [p for p, _ in gold["above_2^50"][str(n)]]
        assert [PR.smallest_primitive_root(q) for q in below[:6]] == [g for _, g in gold["below_2^50"][str(n)][:6]]
    p = PR.config_per_op_gpu()
    assert len(p.primes) == 24 and len(p.aux_primes) == 2
    assert p.primes[:3] == [1125899903827969, 1125899902124033, 1125899887312897]


def test_tune_table_round_trip_and_validation():
    """The keyswitch tuning table (host-side state) exports what was imported and rejects
    malformed or out-of-range launch shapes without changing the table."""
    G = capi
    before = G.tune_table_export()
    rows = ("9 5 7 9 3 7 2 16 1 16 1.250000\n9 5 7 9 2 4 1 4 0 16 0.500000\n9 5 7 9 4 7 1 8 0 8 0.750000\n"
            "9 5 7 9 3 7 1 4 7 16 0.125000\n")  # the last: an exact-convolution shape (load schedule 7)
    assert G.tune_table_import(rows) == 4
    table = G.tune_table_export()
    for line in rows.splitlines():
        assert line in table.splitlines()
    for bad in ("9 5 7 9 3 7 2 8 1 16 1.0\n", "9 5 7 9 3 8 2 16 1 16 1.0\n", "9 5 7 9\n",
                "9 5 7 9 3 7 2 4 0 8 1.0\n", "9 5 7 9 3 7 2 16 1 1.0\n", "9 5 7 9 3 7 2 16 5 16 1.0\n", "9 5 7 9 3 7 2 16 6 16 1.0\n", "9 5 7 9 3 7 2 4 7 16 1.0\n",
                "9 5 7 9 3 7 1 3 7 16 1.0\n", "9 5 7 9 3 6 1 4 7 16 1.0\n", "9 5 7 9 3 7 1 4 7 8 1.0\n"):
        with pytest.raises(G.RckError):
            G.tune_table_import(bad)
        assert G.tune_table_export() == table
    assert set(before.splitlines()) <= set(table.splitlines())


def test_rng_stream_matches_oracle():
    """The library's seeded stream (host side, no GPU needed) is the oracle's xoshiro256** /
    splitmix64 stream draw for draw, and its first outputs match the published xoshiro256**
    reference values for the splitmix64-expanded seed 0 state."""
    from oracle import oracle as O
    for seed in (0, 1, 12345, (1 << 64) - 1):
        a, b = capi.Rng(seed), O.Rng(seed)
        assert [a.next() for _ in range(1000)] == [b.next() for _ in range(1000)]
    # splitmix64(0) expansion and xoshiro256** step, recomputed here with Python integers
    M = (1 << 64) - 1

    def sm(x):
        x = (x + 0x9E3779B97F4A7C15) & M
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return x, z ^ (z >> 31)

    x, s = 0, []
    for _ in range(4):
        x, z = sm(x)
        s.append(z)
    rotl = lambda v, k: ((v << k) | (v >> (64 - k))) & M
    want = []
    for _ in range(5):
        want.append(rotl((s[1] * 5) & M, 7) * 9 & M)
        t = (s[1] << 17) & M
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45)
    r = capi.Rng(0)
    assert [r.next() for _ in range(5)] == want


def test_ps_block_size_and_required_levels_kats():
    """paterson_stockmeyer_block_size_is_ceil_sqrt / ..._required_levels_cover_all_multiplication_levels
    (src/algorithms/paterson_stockmeyer.rs:353-366), for the library and the oracle restatement."""
    from oracle import ps as OPS
    for t, b in ((1, 1), (2, 2), (4, 2), (5, 3), (9, 3), (10, 4), (64, 8)):
        assert capi.ps_block_size(t) == b == OPS.block_size(t)
    assert capi.ps_required_levels(7, 1) == [] == OPS.required_levels(7, 1)
    assert capi.ps_required_levels(7, 5) == [7, 6, 5] == OPS.required_levels(7, 5)
    # degree 63: block 8, 8 blocks -> 14 multiplication levels, needs level >= 14
    assert capi.ps_required_levels(14, 64) == list(range(14, 0, -1)) == OPS.required_levels(14, 64)
    with pytest.raises(capi.RckError, match="polynomial degree is too large"):
        capi.ps_required_levels(13, 64)
    with pytest.raises(OPS.PsError, match="polynomial degree is too large"):
        OPS.required_levels(13, 64)


def test_bsgs_rotation_steps_kats():
    """diagonal_transform_bsgs_steps_cover_baby_and_giant_rotations (linear_transform.rs:315-320)."""
    from oracle import bsgs as OB
    for dim, want in ((1, []), (4, [1, 2]), (8, [1, 2, 3, 6])):
        assert capi.bsgs_rotation_steps(dim) == want == OB.rotation_steps(dim)
    steps = capi.bsgs_rotation_steps(4096)
    assert len(steps) == 126 and steps == OB.rotation_steps(4096)
    assert steps[:63] == list(range(1, 64)) and steps[63:] == [64 * g for g in range(1, 64)]


@pytest.mark.parametrize("gpus", [2, 8])
def test_bench_refuses_more_than_one_gpu(gpus):
    """bench.py measures one GPU only: --gpus N > 1 exits with status 2 and says so, before
    anything touches a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(gpus)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_rust_binding_covers_every_declared_symbol():
    """INTEGRATION.md's Rust extern block has one `pub fn` per entry point of include/*.h."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    bound = set(re.findall(r"pub fn (rck_\w+)\(", text))
    assert declared_symbols() == bound


def test_wire_fixture_is_self_consistent():
    """The golden RCK1 images (tests/golden/wire_rck1.json) carry the header fields their JSON states,
    payload sizes that match, and the stated residues limb-major as little-endian u64, all canonical."""
    import json
    import struct
    with open(os.path.join(os.path.dirname(__file__), "golden", "wire_rck1.json")) as f:
        fx = json.load(f)
    n = fx["N"]
    for kind, rec, nq, np_, count, groups in (
            (1, fx["ciphertext"], fx["ciphertext"]["limbs"], 0, 2, 2),
            (2, fx["key"], fx["key"]["level"] + 1, fx["key"]["aux"], fx["key"]["digits"], 2 * fx["key"]["digits"])):
        raw = bytes.fromhex(rec["bytes_hex"])
        assert raw[:4] == b"RCK1"
        ver, k, logn, cnt, hq, hp, sc = struct.unpack("<7I", raw[4:32])
        gal, _fp, payload = struct.unpack("<3Q", raw[32:56])
        assert (ver, k, 1 << logn, cnt, hq, hp) == (1, kind, n, count, nq, np_)
        assert sc == (rec["scale_bits"] if kind == 1 else rec["level"])
        assert gal == (0 if kind == 1 else rec["galois"])
        assert payload == len(raw) - 56 == groups * (nq + np_) * n * 8
        vals = struct.unpack(f"<{payload // 8}Q", raw[56:])
        assert list(vals) == [v for row in rec["residues"] for v in row]
        moduli = fx["primes"][:nq] + fx["aux_primes"][:np_]
        for g in range(groups):
            for l, q in enumerate(moduli):
                assert all(v < q for v in rec["residues"][g * len(moduli) + l])
