"""bench.py's measurement accounting on the CPU: the exact MAC's byte and matrix-core counts against
their piecewise sums, and the counter ingestion (rocprofv3 --pmc CSVs -> the counters record) on
synthetic counter files."""
import csv
import importlib.util
import json
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_exact_mac_io_is_key_plus_digits_plus_accumulators(bench):
    # n=4 slots, B=3 ciphertexts, D=5 digits (8 once rounded), 6 key columns, 2 groups, 2 aux primes
    key = 2 * 4 * 7 * 6 * 8          # byte planes: prime x slot x plane x column x rounded digits
    digits = 2 * 4 * 3 * 8 * 8       # digit words: prime x slot x ciphertext x rounded digits x 8 bytes
    acc = 2 * 2 * 4 * 3 * 6 * 8      # accumulators: group x prime x slot x ciphertext x column x 8 bytes
    assert bench.exact_mac_io(4, 3, 5, 6, 2) == key + digits + acc == 6528


def test_exact_mac_mfma_counts_tiles(bench):
    # 17 ciphertexts -> 2 row tiles, 52 columns -> 4 column tiles, 49 plane products per tile and slot
    r = bench.exact_mac_mfma(4, 17, 40, 52, 1)
    assert r["mfma_per_launch"] == 2 * 4 * 2 * 4 * 49
    assert r["int8_ops_issued"] == r["mfma_per_launch"] * 16 * 16 * 64 * 2
    assert r["int8_ops_useful"] == 2 * 4 * 17 * 40 * 52 * 49 * 2
    assert r["int8_ops_useful"] < r["int8_ops_issued"]


def _write_pass(d, counter, values):
    """One --pmc pass: a warm-up dispatch, the marker copy, then (kernel, value) dispatches."""
    os.makedirs(os.path.join(d, "host"), exist_ok=True)
    with open(os.path.join(d, "host", "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        rows = [("rck::k_fwd_cols<8>(x)", 5.0), ("rck::k_copy(double4 const*, double4*, long long)", 1.0)] + values
        for i, (k, v) in enumerate(rows, start=1):
            w.writerow([i, k, counter, v])


def test_ingest_counters(bench, tmp_path, monkeypatch):
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    monkeypatch.setattr(bench, "source_sha", lambda: "feedface")
    os.makedirs(tmp_path / "profiles")
    mac, dec = "rck::k_ks_mac_mfma(rck::MfmaArgs, int, int, int const*)", "void rck::k_decompose<40, 24>(x)"
    gated = mac  # the gated recomputation's launch: same kernel, no work while the gate is clear
    steps = 2
    # per step: the decomposition, the working MAC, the gated MAC that does nothing
    fetch = [(dec, 100.0), (mac, 1000.0), (gated, 0.5)] * steps
    write = [(dec, 50.0), (mac, 400.0), (gated, 0.0)] * steps
    sq = [(dec, 7.0), (mac, 3.0), (gated, 0.0)] * steps
    _write_pass(tmp_path / "p1", "FETCH_SIZE", fetch)
    _write_pass(tmp_path / "p2", "WRITE_SIZE", write)
    _write_pass(tmp_path / "p3", "SQ_INSTS_VALU_MFMA_MOPS_I8", sq)
    rec = bench.ingest_counters([str(tmp_path / p) for p in ("p1", "p2", "p3")], "batched", "W", 16, steps)
    dom = rec["dominant"]
    # the marker and what precedes it are left out; the idle launch is not averaged into the MAC
    assert dom["kernel"] == mac and dom["dispatches_per_step"] == 1.0
    assert dom["fetch_bytes"] == round(1000 * 1024 * 1.9913)
    assert dom["write_bytes"] == round(400 * 1024 * 0.9666)
    assert dom["mfma_i8_ops"] == 3 * 512
    # per step: every dispatch after the marker, divided by the steps
    assert rec["per_step"]["fetch_bytes"] == round((100 + 1000 + 0.5) * 1024 * 1.9913)
    assert rec["source_sha"] == "feedface" and rec["workload"] == "W"
    with open(tmp_path / "profiles" / "counters.json") as f:
        assert json.load(f)["batched"]["dominant"]["kernel"] == mac


def test_ingest_rejects_mismatched_passes(bench, tmp_path, monkeypatch):
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    monkeypatch.setattr(bench, "source_sha", lambda: "x")
    os.makedirs(tmp_path / "profiles")
    _write_pass(tmp_path / "a", "FETCH_SIZE", [("rck::k_ks_mac_mfma(x)", 1.0)])
    _write_pass(tmp_path / "b", "WRITE_SIZE", [("rck::k_decompose<8, 8>(x)", 1.0)])
    with pytest.raises(SystemExit):
        bench.ingest_counters([str(tmp_path / "a"), str(tmp_path / "b")], "batched", "W", 1, 1)


def test_counter_record_needs_matching_sources(bench, tmp_path, monkeypatch):
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    os.makedirs(tmp_path / "profiles")
    n, L, K, D, B = 1 << 16, 24, 2, 40, 16
    rec = {"batched": {"workload": f"N=2^16 L={L} K={K} D={D} B={B}", "source_sha": "aaaa", "dominant": {}}}
    with open(tmp_path / "profiles" / "counters.json", "w") as f:
        json.dump(rec, f)
    monkeypatch.setattr(bench, "source_sha", lambda: "aaaa")
    assert bench.counter_record("batched", n, L, K, D, B) is not None
    assert bench.counter_record("batched", n, L, K, D, 4) is None  # other workload
    monkeypatch.setattr(bench, "source_sha", lambda: "bbbb")
    assert bench.counter_record("batched", n, L, K, D, B) is None  # other library sources
