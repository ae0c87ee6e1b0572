"""ct x ct mult + relinearize throughput at N=2^16, L=24 (+2 aux primes, 40 digits) on one MI355X.

A step is one RnsCkksContext::mult followed by one relinearize (the crate's Hadamard-product
flow, context/ops.rs:63-70, 111-122) on a ciphertext pair already resident in HBM, with the
relinearization key resident in NTT form.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode per-op|batched] [--batch B]
                  [--no-cpu-baseline]
"""
import argparse
import hashlib
import json
import math
import os
import re
import subprocess
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "rns-ckks-rust_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# MI355X: 256 CUs, each issuing one 64-lane FP64 wave instruction per clock (78.6 TFLOP/s of FMA at
# 2.4 GHz); every count below is in wave instructions
CUS, NOMINAL_GHZ = 256, 2.4
FP64_PER_BUTTERFLY = 8  # one exact FP64 modular product (6) + the add and the subtract


KS_METHOD = {"exact": 0, "digit-ntt": 1, "auto": 2}  # rck_context_set_keyswitch_method values


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def uniform(moduli, n, rs, polys):
    out = np.empty((polys, len(moduli), n), dtype=np.uint64)
    for l, q in enumerate(moduli):
        out[:, l] = rs.integers(0, q, size=(polys, n), dtype=np.uint64)
    return out


def time_it(G, fn, steps, stream=None):
    """HIP events on the stream the work is launched on, bracketing exactly `steps` steps."""
    e0, e1 = G.Event(), G.Event()
    G.synchronize()
    e0.record(stream.ptr if stream is not None else None)
    for _ in range(steps):
        fn()
    e1.record(stream.ptr if stream is not None else None)
    e1.synchronize()
    G.synchronize()
    return e0.elapsed_ms(e1)


def mac_bytes(n, D, G_, B, dsplit, fold_limbs=0):
    """Algorithmic HBM bytes of one digit-MAC launch: the column-pass output of B*D digits for its
    G_ eval limbs, the key rows (2 per digit per limb, read once), the partial sums written and, when
    the relinearization terms are folded into the accumulators, their NTT-form rows (2 per ciphertext
    for each of the launch's fold_limbs Q limbs)."""
    return (B * D * G_ + 2 * D * G_ + dsplit * B * 2 * G_ + 2 * B * fold_limbs) * n * 8


def alg_bytes_per_op(n, L, K, D, B):
    """Algorithmic HBM bytes of one mult+relin when B ops share one key stream: the two input
    ciphertexts and the output ciphertext once (6 L limbs) plus the key (2 D (L+K) limbs) read once per
    batch -- 8 N [6 L + 2 D (L+K) / B]; B = 1 is the per-op floor 8 N [4 L + 2 D (L+K) + 2 L]."""
    return 8 * n * (6 * L + 2 * D * (L + K) / B)


def modmul_per_op(n, L, K, D):
    """64-bit modular multiply(-add)s of one mult+relin: NTT butterflies of the NTT-resident algorithm
    ((N/2) log2 N [4L + 3L + D(L+K) + 2(L+K)] limb transforms), the key MAC 2 D (L+K) N, the
    decomposition N [L(L-1)/2 + L D/2 + D(L+K)], the mod-down 2 N L (K+2) and the tensor 4 L N."""
    logn = n.bit_length() - 1
    bfly = (n // 2) * logn * (4 * L + 3 * L + D * (L + K) + 2 * (L + K))
    return {"butterflies": bfly, "key_mac": 2 * D * (L + K) * n,
            "decompose": n * (L * (L - 1) // 2 + L * D // 2 + D * (L + K)),
            "mod_down": 2 * n * L * (K + 2), "tensor": 4 * L * n}



MFMA_I8_PEAK = 5.0e15  # MI355X dense int8 matrix-core peak (ops/s, 2 per multiply-accumulate)


def exact_mac_io(n, B, D, cols, G_):
    """HBM bytes of one exact-convolution MAC launch (matrix cores): the key's byte planes read once
    (7 planes x cols x D rounded up to 8 bytes per slot, two auxiliary primes), the digit words (B x D rounded
    up to 8, 8 bytes each, two primes) and the accumulators written (G_ groups x B x cols doubles, two primes)."""
    d8 = 8 * ((D + 7) // 8)
    return 2 * n * (7 * cols * d8 + B * d8 * 8 + G_ * B * cols * 8)


def exact_mac_mfma(n, B, D, cols, G_):
    """int8 matrix-core work of one exact-convolution MAC launch: every 16 x 16 (ciphertexts x key columns)
    tile of every slot takes 49 v_mfma_i32_16x16x64_i8 (one per key plane x digit plane); the useful part
    is the 49 byte products of each digit x key pair."""
    mfma = G_ * 2 * n * (-(-B // 16)) * (-(-cols // 16)) * 49
    return {"mfma_per_launch": mfma, "int8_ops_issued": mfma * 16 * 16 * 64 * 2,
            "int8_ops_useful": 2 * n * B * D * cols * 49 * 2}


def ingest_counters(dirs, mode, workload, batch, steps):
    """profiles/counters.json record of `mode` from rocprofv3 --pmc passes (one directory each) over
    `bench.py --mode MODE --counter-run --steps STEPS`: the dispatches after the 1 MiB marker copy,
    FETCH_SIZE / WRITE_SIZE (KiB) times the copy calibration (profiles/copy_calibration.md), FP64 and all
    VALU wave instructions, matrix-core instructions and busy cycles; per step and per launch of each kernel.
    A kernel's launches that do no work (the gated recomputation when its flag is clear) are left out of
    its per-launch figures; the dominant kernel is the keyswitch multiply-accumulate."""
    import csv
    import glob
    cal = {"fetch": 1.9913, "write": 0.9666}
    disp = {}  # dispatch id -> [kernel, {counter: value}]
    for d in dirs:
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if len(files) != 1:
            raise SystemExit(f"ingest: expected one counter_collection.csv under {d}, found {len(files)}")
        pass_disp = {}
        with open(files[0]) as f:
            for r in csv.DictReader(f):
                pass_disp.setdefault(int(r["Dispatch_Id"]), [r["Kernel_Name"], {}])[1][r["Counter_Name"]] = \
                    float(r["Counter_Value"])
        ids = sorted(pass_disp)
        marks = [i for i in ids if pass_disp[i][0].split("(")[0].split("::")[-1].strip() == "k_copy"]
        if not marks:
            raise SystemExit(f"ingest: no marker copy in {d}")
        after = [pass_disp[i] for i in ids if i > marks[-1]]
        for k, (name, cs) in enumerate(after):
            e = disp.setdefault(k, [name, {}])
            if e[0] != name:
                raise SystemExit("ingest: the passes dispatched different kernels")
            e[1].update(cs)

    def conv(cs):
        out = {}
        if "FETCH_SIZE" in cs:
            out["fetch_bytes"] = cs["FETCH_SIZE"] * 1024 * cal["fetch"]
        if "WRITE_SIZE" in cs:
            out["write_bytes"] = cs["WRITE_SIZE"] * 1024 * cal["write"]
        f64 = [cs[k] for k in ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64",
                               "SQ_INSTS_VALU_TRANS_F64") if k in cs]
        if f64:
            out["valu_f64"] = sum(f64)
        for src, dst, scale in (("SQ_INSTS_VALU", "valu", 1), ("SQ_INSTS_VALU_MFMA_MOPS_I8", "mfma_i8_ops", 512),
                                ("SQ_VALU_MFMA_BUSY_CYCLES", "mfma_busy_cycles", 1),
                                ("GRBM_GUI_ACTIVE", "gui_active_cycles", 1), ("MfmaUtil", "mfma_util_pct", 1)):
            if src in cs:
                out[dst] = cs[src] * scale
        return out

    per_step, kernels = {}, {}
    by_name = {}
    for k in sorted(disp):
        name, cs = disp[k]
        v = conv(cs)
        for key, x in v.items():
            if key != "mfma_util_pct":  # a ratio: per kernel only
                per_step[key] = per_step.get(key, 0.0) + x / steps
        by_name.setdefault(name, []).append(v)
    for name, vs in by_name.items():
        top = max(v.get("fetch_bytes", 0.0) + v.get("write_bytes", 0.0) for v in vs)
        work = [v for v in vs if v.get("fetch_bytes", 0.0) + v.get("write_bytes", 0.0) > 0.01 * top] or vs
        rec = {"dispatches_per_step": len(work) / steps}
        for key in work[0]:
            rec[key] = sum(v.get(key, 0.0) for v in work) / len(work)
        kernels[name] = rec
    dom = [k for k in kernels if "k_ks_mac" in k or "k_digit_mac" in k]
    if not dom:
        raise SystemExit("ingest: no keyswitch multiply-accumulate among the dispatches")
    dname = max(dom, key=lambda k: kernels[k].get("fetch_bytes", 0.0))
    rnd = lambda r: {k: (round(v, 3) if k in ("dispatches_per_step", "mfma_util_pct") else int(round(v)))
                     for k, v in r.items()}
    record = {"workload": workload, "source_sha": source_sha(), "steps_profiled": steps, "batch": batch,
              "calibration": cal, "dominant": dict(rnd(kernels[dname]), kernel=dname),
              "per_step": rnd(per_step), "kernels": {k: rnd(v) for k, v in kernels.items()}}
    path = os.path.join(ROOT, "profiles", "counters.json")
    allrec = {}
    if os.path.exists(path):
        with open(path) as f:
            allrec = json.load(f)
    allrec["_note"] = ("rocprofv3 --pmc runs of bench.py --counter-run (dispatches after the marker copy), "
                       "aggregated by bench.py --ingest-counters: HBM bytes = FETCH_SIZE / WRITE_SIZE KiB x copy "
                       "calibration; FP64 = SQ_INSTS_VALU_{FMA,ADD,MUL,TRANS}_F64 wave instructions; per launch of "
                       "each kernel (launches that do no work left out) and per step (one batch); int8 matrix-core ops = "
                       "SQ_INSTS_VALU_MFMA_MOPS_I8 x 512; mfma_util_pct = the MfmaUtil derived counter (its own pass)")
    allrec[mode] = record
    with open(path, "w") as f:
        json.dump(allrec, f, indent=1)
    return record

# the sources the measured mult+relinearize step is compiled from (kernels, their launchers and the
# batched executor): counter records taken on other versions of these files are not reused
MEASURED_SOURCES = ("modarith.hpp", "context.hpp", "ntt.hpp", "ntt.hip", "ops.hpp", "ops.hip", "pipeline.hpp",
                    "pipeline.hip", "moddown.hpp", "keyswitch.hip", "ks_exact.hip", "batch.hip")


def source_sha():
    """Hash of the measured path's sources."""
    h = hashlib.sha256()
    for f in MEASURED_SOURCES:
        h.update(f.encode())
        with open(os.path.join(ROOT, "rns-ckks-rust_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def counter_record(mode, n, L, K, D, B):
    """rocprofv3 counter summary of this exact workload and library source (profiles/counters.json):
    copy-calibrated FETCH_SIZE/WRITE_SIZE and FP64 / all VALU wave instructions, per launch of the
    dominant kernel and per step.  None when the workload or the sources differ."""
    path = os.path.join(ROOT, "profiles", "counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get(mode)
    if not rec or rec.get("workload") != f"N=2^{n.bit_length() - 1} L={L} K={K} D={D} B={B}":
        return None
    if rec.get("source_sha") != source_sha():
        log(f"[bench] profiles/counters.json ({mode}) was taken on other library sources: not used")
        return None
    return rec


def cpu_info():
    """Host cores this process may use (the job's CPU share) and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else affinity
    return {"threads": threads, "affinity": affinity, "visible": os.cpu_count(), "model": model}


class ClockSampler:
    """Samples the GPU engine clock (rocm-smi, a child process) every 0.25 s while work runs."""

    def __init__(self):
        self.samples, self._stop = [], threading.Event()
        self._t = threading.Thread(target=self._run, daemon=True)

    def _run(self):
        t0 = time.perf_counter()
        while not self._stop.is_set():
            try:
                r = subprocess.run(["rocm-smi", "--showgpuclocks"], capture_output=True, text=True, timeout=10)
                m = re.search(r"sclk[^\n]*?\((\d+)\s*mhz\)", r.stdout, re.I)
                if m:
                    self.samples.append((time.perf_counter() - t0, int(m.group(1))))
            except (OSError, subprocess.SubprocessError):
                return
            self._stop.wait(0.25)

    def __enter__(self):
        self._t.start()
        return self

    def __exit__(self, *exc):
        self._stop.set()
        self._t.join(15)

    def sustained_ghz(self, settle_s=1.5):
        late = [mhz for t, mhz in self.samples if t >= settle_s]
        if not late:
            return None
        late.sort()
        return late[len(late) // 2] / 1000.0


def brief(rec, keep=4):
    """A tuning report with each candidate list cut to its `keep` fastest entries (and their count)."""
    if not isinstance(rec, dict):
        return rec
    out = {}
    for k, v in rec.items():
        if k == "candidates" and isinstance(v, list):
            out["candidates_timed"] = len(v)
            out[k] = sorted(v, key=lambda c: c.get("ms", 0))[:keep]
        elif isinstance(v, list):
            out[k] = [brief(x, keep) for x in v]
        else:
            out[k] = v
    return out


def per_op_entry(table, n, limbs, E, D):
    """The tuning-table entry of one ciphertext at this workload, in the tuner's report form."""
    logn = n.bit_length() - 1
    for line in table.splitlines():
        v = line.split()
        if [int(x) for x in v[:5]] == [logn, limbs, E, D, 1]:
            return {"limbs": limbs, "batch": 1, "eval_limb_group": int(v[5]), "digit_splits": int(v[6]),
                    "mac_rows": int(v[7]), "key_prefetch": int(v[8]), "mac_width": int(v[9]),
                    "keyswitch_ms": float(v[10]), "candidates": [], "from_table": True}
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed batches (batched) / ops (per-op)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mode", default="batched", choices=["batched", "per-op", "serve", "mixed", "bsgs", "wide", "ntt"])
    ap.add_argument("--cts", type=int, default=256, help="mixed: ciphertexts in the batch")
    ap.add_argument("--dim", type=int, default=4096, help="bsgs: matrix dimension")
    ap.add_argument("--vectors", type=int, default=32, help="bsgs: encrypted vectors per batch")
    ap.add_argument("--requests", type=int, default=128, help="serve: requests in the Poisson stream")
    ap.add_argument("--qps", type=float, default=250.0, help="serve: mean arrival rate (requests/s)")
    ap.add_argument("--serve-budget-gb", type=float, default=32.0, help="serve: working-set budget (GB)")
    ap.add_argument("--lanes", type=int, default=4, help="serve: concurrent stream lanes")
    ap.add_argument("--batch", type=int, default=16, help="ciphertext pairs per batch")
    ap.add_argument("--method", default="auto", choices=["auto", "exact", "digit-ntt"],
                    help="keyswitch method of the measured keys (auto: keys hold both forms and each keyswitch "
                         "takes the faster; the single methods are timed beside it)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=None,
                    help="oracle threads (default: the host cores this process may use)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuners (automatic launch shapes)")
    ap.add_argument("--tune-table", default=None, metavar="PATH",
                    help="load autotuner results from PATH when it exists (tuned shapes are not re-timed) and "
                         "save the table there afterwards")
    ap.add_argument("--clock-probe-s", type=float, default=3.0,
                    help="batched: seconds of back-to-back replays while the engine clock is sampled (0: skip)")
    ap.add_argument("--counter-run", action="store_true",
                    help="only the measured work for rocprofv3 --pmc: after a 1 MiB marker copy, exactly --steps "
                         "batches (or per-op steps) and nothing else")
    ap.add_argument("--ingest-counters", nargs="+", default=None, metavar="DIR",
                    help="write the profiles/counters.json record of --mode (batched / per-op at --batch, --steps) "
                         "from rocprofv3 --pmc output directories of its --counter-run, and exit")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB MiB (5 times): the known-volume reference "
                         "for rocprofv3 FETCH_SIZE/WRITE_SIZE calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the path has no multi-GPU form here); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    host = cpu_info()
    if args.cpu_threads is None:
        args.cpu_threads = host["threads"]

    from rnsckks import capi as G
    from rnsckks import params as PR

    if args.ingest_counters:
        p = PR.config_per_op_gpu()
        n, L, K = p.poly_degree, len(p.primes), len(p.aux_primes)
        B = args.batch if args.mode == "batched" else 1
        if args.mode not in ("batched", "per-op"):
            raise SystemExit("--ingest-counters: --mode batched or per-op")
        # decomposition digits: 30-bit words of Q_L (the library's rule; no device needed)
        D = (math.prod(p.primes[:L]).bit_length() + 29) // 30
        rec = ingest_counters(args.ingest_counters, args.mode, f"N=2^{n.bit_length() - 1} L={L} K={K} D={D} B={B}",
                              B, args.steps)
        print(json.dumps({"ingested": args.mode, "dominant": rec["dominant"], "per_step": rec["per_step"]}))
        return

    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        src, dst = G.DeviceBuffer(nbytes // 8), G.DeviceBuffer(nbytes // 8)
        for _ in range(5):
            G.diag_copy(dst, src, nbytes)
        G.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "launches": 5}), flush=True)
        return

    if args.mode == "serve":
        serve_main(args, G, PR)
        return
    if args.mode == "mixed":
        mixed_main(args, G, PR)
        return
    if args.mode == "bsgs":
        bsgs_main(args, G, PR)
        return
    if args.mode == "wide":
        wide_main(args, G, PR)
        return
    if args.mode == "ntt":
        ntt_main(args, G, PR)
        return

    p = PR.config_per_op_gpu()
    n, L, K = p.poly_degree, len(p.primes), len(p.aux_primes)
    B = args.batch if args.mode == "batched" else 1
    log(f"[bench] context N={n} L={L} K={K} mode={args.mode} B={B}")
    ctx = G.Context(p)
    D = ctx.num_decomp_digits(L)
    rs = np.random.default_rng(0)
    P = L * n
    lhs = G.DeviceBuffer.from_numpy(uniform(p.primes, n, rs, 2 * B))
    rhs = G.DeviceBuffer.from_numpy(uniform(p.primes, n, rs, 2 * B))
    log(f"[bench] relin key: {D} digits x 2 x {L + K} limbs")
    key = np.stack([uniform(p.primes + p.aux_primes, n, rs, 2) for _ in range(D)])
    # the measured keys use --method; the single-method keys are timed beside it (not in counter runs)
    ctx.set_keyswitch_method(KS_METHOD[args.method])
    ksk = ctx.ksk_create(key, L - 1, K)
    alts = {}
    if not args.counter_run:
        for other in ("exact", "digit-ntt"):
            if other != args.method:
                ctx.set_keyswitch_method(KS_METHOD[other])
                alts[other] = ctx.ksk_create(key, L - 1, K)
        ctx.set_keyswitch_method(KS_METHOD[args.method])
    del key
    quad = G.DeviceBuffer((3, L, n))
    out = G.DeviceBuffer((B, 2, L, n))
    one = G.DeviceBuffer((2, L, n))

    def per_op_pair(b, dst, k=None):
        ctx.mult(lhs.offset(b * 2 * P), rhs.offset(b * 2 * P), L, 40, 40, quad)
        ctx.relinearize(quad, L, 80, k if k is not None else ksk, dst)

    def per_op_step():
        per_op_pair(0, one)

    # autotune the per-op keyswitch launch shape (results are identical for every shape)
    tune1 = {"digit_splits": None, "candidates": []}
    loaded = 0
    if args.tune_table and os.path.exists(args.tune_table):
        with open(args.tune_table) as f:
            loaded = G.tune_table_import(f.read())
        log(f"[bench] tuning table {args.tune_table}: {loaded} entries loaded")
        tune1 = per_op_entry(G.tune_table_export(), n, L, L + K, D) or tune1
    if not args.no_tune and not loaded:
        tune1 = ctx.tune_keyswitch(L, ksk, 1)
        log(f"[bench] per-op keyswitch tuned: { {k: v for k, v in tune1.items() if k != 'candidates'} }")
        for cand in sorted(tune1["candidates"], key=lambda c: c["ms"])[:8]:
            log(f"[bench]     candidate {cand}")
        for k_alt in alts.values():
            ctx.tune_keyswitch(L, k_alt, 1)

    bt = None
    if args.mode == "batched":
        bt = G.Batch(ctx)
        for b in range(B):
            bt.add(G.OP_MULT_RELIN, L, lhs.offset(b * 2 * P), rhs.offset(b * 2 * P), ksk, out.offset(b * 2 * P))
        t = time.perf_counter()
        bt.compile(tune=not args.no_tune)
        log(f"[bench] batch compiled{'' if args.no_tune else '+tuned'} in {time.perf_counter() - t:.1f} s")
    stream = G.Stream()

    if args.counter_run:
        # the profiled region: a 1 MiB marker copy, then exactly `steps` units of measured work
        mk = G.DeviceBuffer(1 << 17)
        G.synchronize()
        G.diag_copy(mk, mk.offset(1 << 16), 1 << 19)
        for _ in range(args.steps):
            if bt is not None:
                bt.run(stream)
            else:
                per_op_step()
        G.synchronize()
        print(json.dumps({"counter_run": args.mode, "steps": args.steps, "batch": B,
                          "workload": f"N=2^{n.bit_length() - 1} L={L} K={K} D={D} B={B}",
                          "source_sha": source_sha()}), flush=True)
        return

    # per-op path (always measured: the batched form is reported against it)
    for _ in range(args.warmup):
        per_op_step()
    per_op_steps = args.steps if args.mode == "per-op" else max(4, args.steps // 2)
    per_op_ms = time_it(G, per_op_step, per_op_steps) / per_op_steps
    per_op_by = {args.method: round(1000.0 / per_op_ms, 2)}
    for other, k_alt in alts.items():
        for _ in range(args.warmup):
            per_op_pair(0, one, k_alt)
        ms_alt = time_it(G, lambda: per_op_pair(0, one, k_alt), per_op_steps) / per_op_steps
        per_op_by[other] = round(1000.0 / ms_alt, 2)
    # the plain per-op path: the fastest of the methods
    per_op_value = max(per_op_by.values())
    log(f"[bench] per-op: {per_op_by} ops/s")

    prof_steps = 2
    report = None
    batched_by = None
    value_rescale = None
    bit_exact = None
    sclk = None
    if args.mode == "batched":
        report = bt.report()
        for grp in report["groups"]:
            log(f"[bench]   group { {k: v for k, v in grp.items() if k != 'candidates'} }")
            for cand in sorted(grp["candidates"], key=lambda c: c["ms"])[:8]:
                log(f"[bench]     candidate {cand}")

        def step():
            bt.run(stream)

        for _ in range(args.warmup):
            step()
        ms = time_it(G, step, args.steps, stream)
        ms_per_step = ms / args.steps
        value = B * 1000.0 / ms_per_step
        log(f"[bench] batched: {ms_per_step:.3f} ms per batch of {B} = {value:.1f} ops/s "
            f"({value / per_op_value:.2f}x per-op)")
        # bit-exactness of the timed batch: every output of the last replay against its per-op call
        stream.synchronize()
        got = out.numpy()
        bit_exact = True
        for b in range(B):
            per_op_pair(b, one)
            if not np.array_equal(one.numpy(), got[b]):
                bit_exact = False
                log(f"[bench] MISMATCH: batched output {b} differs from its per-op call")
        log(f"[bench] batched outputs == per-op outputs for all {B} pairs: {bit_exact}")
        prof = profile_kernels(G, bt.run_eager, prof_steps)

        # also reported: the same batch with the rescale included (mult_relinearized,
        # context/ops.rs:124-133); it shares the key-switch launch shape tuned above
        Pr = (L - 1) * n
        out_r = G.DeviceBuffer((B, 2, L - 1, n))
        btr = G.Batch(ctx)
        for b in range(B):
            btr.add(G.OP_MULT_RELIN_RESCALE, L, lhs.offset(b * 2 * P), rhs.offset(b * 2 * P), ksk,
                    out_r.offset(b * 2 * Pr))
        btr.compile(tune=not args.no_tune)
        for _ in range(args.warmup):
            btr.run(stream)
        ms_r = time_it(G, lambda: btr.run(stream), args.steps, stream) / args.steps
        value_rescale = B * 1000.0 / ms_r
        log(f"[bench] batched with rescale: {ms_r:.3f} ms per batch of {B} = {value_rescale:.1f} ops/s")
        del btr, out_r
        # the same batch with the single-method keys (same outputs)
        batched_by = {args.method: round(value, 2)}
        for other, k_alt in alts.items():
            out_a = G.DeviceBuffer((B, 2, L, n))
            bta = G.Batch(ctx)
            for b in range(B):
                bta.add(G.OP_MULT_RELIN, L, lhs.offset(b * 2 * P), rhs.offset(b * 2 * P), k_alt,
                        out_a.offset(b * 2 * P))
            bta.compile(tune=not args.no_tune)
            for _ in range(args.warmup):
                bta.run(stream)
            ms_a = time_it(G, lambda: bta.run(stream), args.steps, stream) / args.steps
            batched_by[other] = round(B * 1000.0 / ms_a, 2)
            bit_exact = bit_exact and np.array_equal(out_a.numpy(), got)
            del bta, out_a
        log(f"[bench] batched by method: {batched_by} ops/s; outputs equal: {bit_exact}")
    else:
        ms_per_step, value = per_op_ms, per_op_value
        prof = profile_kernels(G, per_op_step, prof_steps)

    if args.tune_table:
        with open(args.tune_table, "w") as f:
            f.write(G.tune_table_export())

    ntt = ntt_entry(G, ctx, p, n, L, K)

    if args.mode == "batched" and args.clock_probe_s > 0:
        # the engine clock under this load (it settles below the nominal clock after ~1.5 s)
        with ClockSampler() as cs:
            t_end = time.perf_counter() + args.clock_probe_s
            while time.perf_counter() < t_end:
                for _ in range(20):
                    bt.run(stream)
                stream.synchronize()
        sclk = cs.sustained_ghz()
        log(f"[bench] engine clock under the batched load: {len(cs.samples)} samples, sustained "
            f"{sclk if sclk else 'n/a'} GHz")
    top = sorted(prof.items(), key=lambda kv: -kv[1][1])
    for k, (cnt, tot) in top:
        log(f"[bench]   {k:16s} launches={cnt:5d} total={tot:9.3f} ms avg={tot / cnt * 1e3:9.2f} us")
    kern_ms = {k: v[1] / prof_steps for k, v in top}
    if args.mode == "batched" and "ks_mac" not in kern_ms:
        # the batched mult's transforms (digit-NTT keyswitch): 4 L forward + L inverse (d2) limb NTTs per op,
        # 16 N bytes each, over its forward column pass, fused rows + tensor, and inverse column passes
        # (with the exact convolution the inverse column passes also carry the keyswitch accumulators)
        t_ms = sum(kern_ms.get(k, 0.0) for k in ("ntt_fwd_cols", "mult_rows", "ntt_inv_cols"))
        if t_ms > 0:
            nb = B * 5 * L * 16 * n
            ntt["batched_mult_transforms"] = {
                "limb_ntts": B * 5 * L, "alg_bytes": nb, "ms": round(t_ms, 4),
                "kernels": "ntt_fwd_cols + mult_rows (incl. tensor) + ntt_inv_cols (incl. the aux-limb inverse "
                           "columns of the keyswitch)",
                "achieved_gbs": round(nb / t_ms / 1e6, 1), "frac": round(nb / t_ms / 1e6 / HBM_PEAK_GBS, 4)}

    # the dominant kernel: the keyswitch multiply-accumulate when it ran, else the longest one
    name, (cnt, tot) = next((kv for kv in top if kv[0] in ("ks_mac", "digit_mac")), top[0])
    avg_s = tot / cnt / 1e3
    ops_per_step = B
    launches = max(1, cnt // prof_steps)
    rec = counter_record(args.mode, n, L, K, D, B)
    roofline = None
    alg_op = alg_bytes_per_op(n, L, K, D, B)
    if name in ("digit_mac", "ks_mac"):
        if name == "ks_mac":
            # exact-convolution MAC on the matrix cores: key byte planes read once, digit words read,
            # accumulators written (each group once)
            groups = ctx.exact_group_size(L)[1]
            io = int(exact_mac_io(n, B, D, 2 * (L + K), groups) / launches)
            mf = exact_mac_mfma(n, B, D, 2 * (L + K), groups)
        else:
            G_ = -(-(L + K) // launches)  # eval limbs per launch
            ds = report["groups"][0]["digit_splits"] if report else tune1["digit_splits"]
            if ds is None:  # automatic shape without tuning: the library's heuristic for one ciphertext
                ds = 1
            folded = report is not None and report["groups"][0]["load_schedule"] in (5, 6) and ds == 1
            io = int(mac_bytes(n, D, G_, B, ds, L / launches if folded else 0))
        alg = int(alg_op * ops_per_step / launches)
        achieved = alg / avg_s / 1e9
        roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "avg_launch_us": round(avg_s * 1e6, 2), "alg_bytes_per_launch": alg,
                    "alg_bytes_basis": f"8N[6L+2D(L+K)/B] = {alg_op / 1e6:.2f} MB per op x {ops_per_step / launches:g} "
                                       "ops per launch",
                    "kernel_io_bytes_per_launch": io,
                    "kernel_io_gbs": round(io / avg_s / 1e9, 1)}
        if name == "ks_mac":
            ops_launch = mf["int8_ops_issued"] / launches
            roofline["mfma"] = {
                "instr": "v_mfma_i32_16x16x64_i8", "instr_per_launch": int(mf["mfma_per_launch"] / launches),
                "int8_ops_issued_per_launch": int(ops_launch),
                "int8_ops_useful_per_launch": int(mf["int8_ops_useful"] / launches),
                "achieved_tops": round(ops_launch / avg_s / 1e12, 1), "peak_tops": MFMA_I8_PEAK / 1e12,
                "frac": round(ops_launch / avg_s / MFMA_I8_PEAK, 4),
                "basis": "49 byte-plane products per 16 x 16 tile (ciphertexts x key columns) per NTT slot, "
                         "K = 64 digits; issued = the tiles' full size, useful = B D 2(L+K) byte products"}
        if rec:
            dom = rec["dominant"]
            roofline["traffic"] = int(dom["fetch_bytes"] + dom["write_bytes"])
            roofline["traffic_vs_alg"] = round(roofline["traffic"] / alg, 3)
            roofline["traffic_vs_kernel_io"] = round(roofline["traffic"] / io, 3)
            roofline["traffic_gbs"] = round(roofline["traffic"] / avg_s / 1e9, 1)
            if "mfma_i8_ops" in dom and "mfma" in roofline:
                roofline["mfma"]["int8_ops_counted_per_launch"] = int(dom["mfma_i8_ops"])
                roofline["mfma"]["busy_pct_counted"] = dom.get("mfma_util_pct")
            f64 = dom.get("valu_f64")
            if f64 is not None:
                roofline["fp64_issue"] = {
                    "wave_instr_per_launch": int(f64),
                    "frac_at_nominal_2.4GHz": round(f64 / (avg_s * CUS * NOMINAL_GHZ * 1e9), 4),
                    "frac_at_sustained_clock": round(f64 / (avg_s * CUS * sclk * 1e9), 4) if sclk else None,
                    "sustained_clock_ghz": sclk,
                    "valu_wave_instr_per_launch": int(dom["valu"]),
                    "source": "profiles/counters.json (SQ_INSTS_VALU_{FMA,ADD,MUL,TRANS}_F64, SQ_INSTS_VALU)"}
    step = {"alg_bytes": int(alg_op * ops_per_step), "ms": round(ms_per_step, 4),
            "achieved": round(alg_op * ops_per_step / ms_per_step / 1e6, 1),
            "frac": round(alg_op * ops_per_step / ms_per_step / 1e6 / HBM_PEAK_GBS, 4)}
    if roofline is not None:
        roofline["whole_step"] = step
    mm = modmul_per_op(n, L, K, D)
    mm_total = sum(mm.values())
    mm_peak = CUS * 64 * NOMINAL_GHZ * 1e9 / FP64_PER_BUTTERFLY
    int_ops = {"modmul_per_op": mm_total, "breakdown": mm, "achieved_per_s": round(mm_total * value, 1),
               "peak_per_s": mm_peak,
               "peak_basis": "39.3 T FP64 lane-instructions/s (78.6 TFLOP/s FMA) / 8 per butterfly "
                             "(one exact FP64 modular product + add + sub)",
               "frac": round(mm_total * value / mm_peak, 4)}
    if rec:
        # HBM bytes the NTT kernels move per launch (counters) over their launch time: the row passes,
        # each a single kernel under its timing name
        fr = {}
        for tname, kname in (("ntt_inv_rows", "k_inv_rows"), ("ntt_fwd_rows", "k_fwd_rows"), ("mult_rows", "k_mult_rows")):
            if tname not in prof:
                continue
            kc = [v for k, v in rec["kernels"].items() if kname in k and "fetch_bytes" in v and "write_bytes" in v]
            if len(kc) != 1:
                continue
            cnt_t, tot_t = prof[tname]
            nbytes = kc[0]["fetch_bytes"] + kc[0]["write_bytes"]
            gbs = nbytes / (tot_t / cnt_t / 1e3) / 1e9
            fr[tname] = {"hbm_bytes_per_launch": int(nbytes), "avg_launch_us": round(tot_t / cnt_t * 1e3, 2),
                         "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        if fr:
            ntt["kernels_measured_traffic"] = fr
    if rec and rec.get("per_step"):
        f64_op = rec["per_step"]["valu_f64"] / B
        int_ops["fp64_wave_instr_per_op"] = int(f64_op)
        int_ops["fp64_ceiling_ops_s"] = {
            "at_2.4GHz": round(CUS * NOMINAL_GHZ * 1e9 / f64_op, 1),
            "at_sustained_clock": round(CUS * sclk * 1e9 / f64_op, 1) if sclk else None,
            "sustained_clock_ghz": sclk}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(p, args.cpu_threads, host)

    line = {
        "metric": "ct×ct mult+relin ops/s at N=2^16 L=24, batched vs per-op GPU; bit-exact",
        "value": round(value, 2), "unit": "ops/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 residues (FP64-exact modular arithmetic)", "data": "synthetic (seeded uniform residues)",
        "config": {"workload": f"{args.mode} mult+relinearize, N=2^16, L=24 primes < 2^50, K=2 aux, D={D} digits"
                               + (f", {B} ciphertext pairs per step sharing one relin key" if B > 1 else ""),
                   "mode": args.mode, "batch": B, "keyswitch_method": args.method,
                   "per_op_ops_s": round(per_op_value, 2), "per_op_ops_s_by_method": per_op_by,
                   "batched_ops_s_by_method": batched_by,
                   "batched_vs_per_op": round(value / per_op_value, 3),
                   "bit_exact_vs_per_op": bit_exact,
                   "batched_with_rescale_ops_s": round(value_rescale, 2) if value_rescale else None},
        "roofline": roofline, "int_ops": int_ops, "ntt": ntt, "cpu_baseline": cpu,
        "tuning": {"per_op": brief(tune1), "batched": brief(report)},
        "kernels_ms_per_step": {k: round(v, 4) for k, v in kern_ms.items()},
    }
    print(json.dumps(line), flush=True)


def profile_kernels(G, fn, steps):
    """Per-kernel timing: `steps` eager runs of the measured work with HIP events around each launch,
    on the stream each kernel is launched on; {name: (launches, total ms)}."""
    G.profile_enable(True)
    for _ in range(steps):
        fn()
    prof = G.profile_report()
    G.profile_enable(False)
    return prof


def ntt_entry(G, ctx, p, n, L, K, iters=20):
    """NttPlan::forward + inverse (ntt.rs:78-89) of every limb of one polynomial over Q || P at this ring:
    a round trip (checked exact) timed with HIP events; 16 N bytes per limb transform."""
    rs = np.random.default_rng(9)
    x = uniform(p.primes + p.aux_primes, n, rs, 1)[0]
    buf = G.DeviceBuffer.from_numpy(x)
    ctx.ntt_forward(buf, L, K)
    ctx.ntt_inverse(buf, L, K)
    exact = bool(np.array_equal(buf.numpy().reshape(x.shape), x))
    stream = G.Stream()
    t_f = time_it(G, lambda: ctx.ntt_forward(buf, L, K, stream.ptr), iters, stream) / iters
    t_i = time_it(G, lambda: ctx.ntt_inverse(buf, L, K, stream.ptr), iters, stream) / iters
    G.profile_enable(True)
    ctx.ntt_forward(buf, L, K)
    ctx.ntt_inverse(buf, L, K)
    prof = G.profile_report()
    G.profile_enable(False)
    limbs = L + K
    nb = limbs * 16 * n
    res = {"limbs": limbs, "ring": f"N=2^{n.bit_length() - 1}", "round_trip_exact": exact,
           "forward_us": round(t_f * 1e3, 2), "inverse_us": round(t_i * 1e3, 2),
           "alg_bytes_per_direction": nb,
           "achieved_gbs": round(2 * nb / ((t_f + t_i) * 1e6), 1),
           "frac": round(2 * nb / ((t_f + t_i) * 1e6) / HBM_PEAK_GBS, 4),
           "kernels_us": {k: round(v[1] * 1e3, 2) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])}}
    log(f"[bench] NTT fwd+inv of {limbs} limbs: {res['forward_us']} + {res['inverse_us']} us, "
        f"{res['achieved_gbs']} GB/s algorithmic (round trip exact: {exact})")
    return res


def serve_main(args, G, PR):
    """Serving loop: Poisson arrivals of Paterson-Stockmeyer requests (degree 7..63 capped by the level
    budget, coefficients U[-1,1]) on fresh ciphertexts at N in {2^15, 2^16} with 12..24 starting
    limbs, keys and ciphertexts resident on the device, admission by working-set bytes."""
    rs = np.random.default_rng(0)
    rings = {}
    t0 = time.perf_counter()
    plans = []
    for _ in range(args.requests):
        n = int(rs.choice([1 << 15, 1 << 16]))
        limbs = int(rs.integers(12, 25))
        deg = int(rs.integers(7, 64))
        while True:  # cap the degree by the level budget, keeping 2 limbs for the result
            try:     # (at 1 limb Q ~ 2^50 equals the working scale and the result cannot decode)
                G.ps_required_levels(limbs - 2, deg + 1)
                break
            except G.RckError:
                deg -= 1
        plans.append((n, limbs, rs.uniform(-1, 1, deg + 1)))
    need = {}
    for n, limbs, coeffs in plans:
        need.setdefault(n, set()).update(G.ps_required_levels(limbs - 1, len(coeffs)))
    for n in sorted(need):
        p = PR.config_batched(n=n, limbs=24, aux=2)
        ctx = G.Context(p)
        ctx.set_keyswitch_method(KS_METHOD[args.method])
        rng = G.Rng(n)
        s = ctx.sample_sparse_ternary(rng, 64, 24)
        pk = ctx.keygen_public(rng, s, 24, sigma=3.2)
        keys = {lv: ctx.keygen_relin(rng, s, lv, sigma=3.2) for lv in sorted(need[n])}
        rings[n] = (p, ctx, rng, s, pk, keys)
        log(f"[bench] ring N={n}: {len(keys)} relinearization keys ({min(keys)}..{max(keys)}) "
            f"generated on the device in {time.perf_counter() - t0:.1f} s")
    # client side: encode + encrypt every request before the stream starts
    reqs = []
    for i, (n, limbs, coeffs) in enumerate(plans):
        p, ctx, rng, s, pk, keys = rings[n]
        x = rs.uniform(-1, 1, n // 2)
        ct = ctx.encrypt(rng, ctx.encode(x, limbs, 40), pk, 24, limbs, sigma=3.2)
        reqs.append({"n": n, "limbs": limbs, "coeffs": coeffs, "x": x, "ct": ct,
                     "out": G.DeviceBuffer((2, limbs, n))})
    log(f"[bench] {len(reqs)} requests encrypted ({time.perf_counter() - t0:.1f} s); "
        f"degrees {min(len(r['coeffs']) - 1 for r in reqs)}..{max(len(r['coeffs']) - 1 for r in reqs)}")
    budget = int(args.serve_budget_gb * 1e9)
    srv = G.Server(budget)
    for n, (p, ctx, rng, s, pk, keys) in rings.items():
        srv.register_ring(ctx, keys)
    # warm-up: one request per ring through the server (first-launch costs), not counted
    for n, (p, ctx, rng, s, pk, keys) in rings.items():
        r = next(q for q in reqs if q["n"] == n)
        rid = srv.submit(ctx, r["ct"], r["limbs"], 40, r["coeffs"], r["out"])
        srv.wait(rid, 120)
    srv.close()
    srv = G.Server(budget, lanes=args.lanes)
    for n, (p, ctx, rng, s, pk, keys) in rings.items():
        srv.register_ring(ctx, keys)
    gaps = rs.exponential(1.0 / args.qps, len(reqs))
    start = time.perf_counter()
    due = start
    ids = []
    for r, gap in zip(reqs, gaps):
        due += gap
        delay = due - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(rings[r["n"]][1], r["ct"], r["limbs"], 40, r["coeffs"], r["out"]))
    srv.drain(600)
    wall = time.perf_counter() - start
    st = srv.stats()
    log(f"[bench] served {st['completed']} requests in {wall:.2f} s: p50 {st['p50_ms']:.1f} ms, "
        f"p99 {st['p99_ms']:.1f} ms, {st['throughput_rps']:.1f} req/s, {st['mult_launches']} batched launches "
        f"for {st['mults']} ct x ct mults, peak working set {st['peak_bytes'] / 1e9:.1f} GB")
    # decode error on a sample, against FP64 evaluation
    errs = []
    for r, rid in list(zip(reqs, ids))[:8]:
        w = srv.wait(rid, 1.0)
        p, ctx, rng, s, pk, keys = rings[r["n"]]
        dec = ctx.decrypt(r["out"], w["limbs"], s)
        got = ctx.decode(dec, w["limbs"], w["scale_bits"]).real
        errs.append(float(np.max(np.abs(got - np.polyval(r["coeffs"][::-1], r["x"])))))
    srv.close()
    log(f"[bench] decode error vs FP64 over {len(errs)} requests: max {max(errs):.2e}")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_ps(args.cpu_threads)
    line = {
        "metric": "Paterson-Stockmeyer serving: requests/s with p50/p99 latency (continuous batching)",
        "value": round(st["throughput_rps"], 2), "unit": "requests/s", "n_gpus": 1, "steps": len(reqs), "warmup": 1,
        "ms_per_step": round(st["span_ms"] / max(1, st["completed"]), 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64 residues (FP64-exact modular arithmetic)",
        "data": "synthetic (seeded Poisson arrivals, uniform coefficients and inputs)",
        "config": {"workload": "PS degree 7..63 (capped by level budget), N in {2^15, 2^16}, start limbs 12..24, "
                               "primes just above 2^50, scale 2^40",
                   "qps": args.qps, "requests": len(reqs), "budget_bytes": budget,
                   "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2),
                   "queue_p99_ms": round(st["queue_p99_ms"], 2), "mults": st["mults"],
                   "mult_launches": st["mult_launches"], "batch_hist": st["batch_hist"],
                   "peak_bytes": st["peak_bytes"], "decode_err_max": max(errs)},
        "roofline": None, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def mixed_main(args, G, PR):
    """Batched form of a mixed workload: ciphertexts at N=2^15 at independent levels of 4..20 limbs,
    40% ct x ct mult+relin, 40% rotation by +-2^k (k < 14) with its Galois key, 20% ct x pt; one
    captured hipGraph (autotuned launch shapes) against the same ops run one by one."""
    n = 1 << 15
    p = PR.config_batched(n=n, limbs=20, aux=2)
    ctx = G.Context(p)
    ctx.set_keyswitch_method(KS_METHOD[args.method])
    rs = np.random.default_rng(0)
    rng = G.Rng(3)
    t0 = time.perf_counter()
    s = ctx.sample_ternary(rng, 20)
    kinds = rs.permutation(["mr"] * int(0.4 * args.cts) + ["rot"] * int(0.4 * args.cts) +
                           ["pt"] * (args.cts - 2 * int(0.4 * args.cts)))
    ops = []
    for kind in kinds:
        limbs = int(rs.integers(4, 21))
        step = int(rs.choice([1 << k for k in range(14)])) * int(rs.choice([-1, 1])) if kind == "rot" else 0
        ops.append((kind, limbs, step))
    keys = {}
    for kind, limbs, step in ops:
        if kind == "pt":
            continue
        kk = (limbs, step)
        if kk not in keys:
            keys[kk] = (ctx.keygen_rotation(rng, s, step, limbs - 1) if step else ctx.keygen_relin(rng, s, limbs - 1))
    log(f"[bench] mixed: {len(ops)} ops, {len(keys)} keys generated on the device in {time.perf_counter() - t0:.1f} s")
    bufs = []
    bt = G.Batch(ctx)
    for kind, limbs, step in ops:
        a = G.DeviceBuffer.from_numpy(uniform(p.primes[:limbs], n, rs, 2))
        out = G.DeviceBuffer((2, limbs, n))
        if kind == "mr":
            b = G.DeviceBuffer.from_numpy(uniform(p.primes[:limbs], n, rs, 2))
            bt.add(G.OP_MULT_RELIN, limbs, a, b, keys[(limbs, 0)], out)
        elif kind == "rot":
            b = None
            bt.add(G.OP_ROTATE, limbs, a, None, keys[(limbs, step)], out)
        else:
            b = G.DeviceBuffer.from_numpy(uniform(p.primes[:limbs], n, rs, 1)[0])
            bt.add(G.OP_MULT_PLAIN, limbs, a, b, None, out)
        bufs.append((kind, limbs, step, a, b, out))
    if args.tune_table and os.path.exists(args.tune_table):
        with open(args.tune_table) as f:
            log(f"[bench] tuning table {args.tune_table}: {G.tune_table_import(f.read())} entries loaded")
    t = time.perf_counter()
    bt.compile(tune=not args.no_tune)
    rep = bt.report()
    log(f"[bench] compiled {len(rep['groups'])} groups in {time.perf_counter() - t:.1f} s")
    if args.tune_table:
        with open(args.tune_table, "w") as f:
            f.write(G.tune_table_export())
    stream = G.Stream()
    for _ in range(args.warmup):
        bt.run(stream)
    ms = time_it(G, lambda: bt.run(stream), args.steps, stream) / args.steps
    batched = [o[5].numpy() for o in bufs]
    quad = {}
    refs = []

    def per_op():
        for kind, limbs, step, a, b, out in bufs:
            if kind == "mr":
                if limbs not in quad:
                    quad[limbs] = G.DeviceBuffer((3, limbs, n))
                ctx.mult(a, b, limbs, 40, 40, quad[limbs])
                ctx.relinearize(quad[limbs], limbs, 80, keys[(limbs, 0)], out)
            elif kind == "rot":
                ctx.rotate(a, limbs, 40, keys[(limbs, step)], out)
            else:
                ctx.mult_plain(a, b, limbs, 40, out)

    for _ in range(args.warmup):
        per_op()
    ms_op = time_it(G, per_op, max(1, args.steps // 2)) / max(1, args.steps // 2)
    same = all(np.array_equal(o[5].numpy(), batched[i]) for i, o in enumerate(bufs))
    value = len(ops) * 1000.0 / ms
    log(f"[bench] mixed batch: {ms:.3f} ms per {len(ops)} ops = {value:.1f} ops/s; per-op {ms_op:.3f} ms "
        f"({ms_op / ms:.2f}x); batched == per-op: {same}")
    G.profile_enable(True)
    bt.run_eager()
    prof = G.profile_report()
    G.profile_enable(False)
    line = {
        "metric": "mixed batch ops/s (mult+relin / rotate / ct x pt at uneven levels), batched vs per-op; bit-exact",
        "value": round(value, 2), "unit": "ops/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 residues (FP64-exact modular arithmetic)", "data": "synthetic (seeded uniform residues)",
        "config": {"workload": f"{len(ops)} ciphertexts at N=2^15, 4..20 limbs (primes > 2^50, K=2), 40% mult+relin, "
                               "40% rotate +-2^k, 20% ct x pt, one hipGraph",
                   "groups": len(rep["groups"]), "keys": len(keys), "per_op_ms": round(ms_op, 4),
                   "batched_vs_per_op": round(ms_op / ms, 3), "bit_exact_vs_per_op": bool(same)},
        "roofline": None, "cpu_baseline": None,
        "kernels_ms_per_step": {k: round(v[1], 4) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])},
    }
    print(json.dumps(line), flush=True)


def ntt_main(args, G, PR):
    """NttPlan::forward + inverse (ntt.rs:78-89) of all 26 limbs of the headline chain (24 Q + 2 aux
    primes below 2^50) at N=2^16 through rck_ntt_forward / rck_ntt_inverse, nothing else: one step is
    one forward and one inverse transform of every limb (16 N bytes per limb and direction)."""
    p = PR.config_per_op_gpu()
    n, L, K = p.poly_degree, len(p.primes), len(p.aux_primes)
    ctx = G.Context(p)
    rs = np.random.default_rng(9)
    x = uniform(p.primes + p.aux_primes, n, rs, 1)[0]
    buf = G.DeviceBuffer.from_numpy(x)
    stream = G.Stream()

    def step():
        ctx.ntt_forward(buf, L, K, stream.ptr)
        ctx.ntt_inverse(buf, L, K, stream.ptr)

    for _ in range(args.warmup):
        step()
    ms = time_it(G, step, args.steps, stream) / args.steps
    exact = bool(np.array_equal(buf.numpy().reshape(x.shape), x))
    nb = 2 * (L + K) * 16 * n
    gbs = nb / ms / 1e6
    log(f"[bench] NTT round trip of {L + K} limbs at N=2^16: {ms * 1e3:.1f} us, {gbs:.1f} GB/s algorithmic, "
        f"exact: {exact}")
    line = {
        "metric": "NTT forward+inverse of every limb (rck_ntt_forward/inverse), algorithmic GB/s",
        "value": round(gbs, 1), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 residues (FP64-exact modular arithmetic)", "data": "synthetic (seeded uniform residues)",
        "config": {"workload": f"cyclic NTT round trip, N=2^16, {L} + {K} limbs (primes < 2^50)",
                   "round_trip_exact": exact, "alg_bytes_per_step": nb},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def wide_main(args, G, PR):
    """The 64-bit integer arithmetic at N=2^16: 20 primes below 2^60 (q = 1 mod 2^17) and 2 aux primes,
    so D = 40 digits as at the headline workload; per-op mult+relinearize, rotation, and the crate's
    NTT round trip of every limb, each against the same op on the FP64-exact kernels at the headline
    workload (24 primes below 2^50)."""
    n, L, K = 1 << 16, 20, 2
    qs = PR.primes_below(60, n, L + K)
    p = PR.RnsCkksParams.with_primes(n, qs[:L], qs[L:], 40)
    ctx = G.Context(p)
    assert ctx.arithmetic == G.ARITH_INT64
    D = ctx.num_decomp_digits(L)
    rs = np.random.default_rng(0)
    P = L * n
    lhs = G.DeviceBuffer.from_numpy(uniform(p.primes, n, rs, 2))
    rhs = G.DeviceBuffer.from_numpy(uniform(p.primes, n, rs, 2))
    key = np.stack([uniform(p.primes + p.aux_primes, n, rs, 2) for _ in range(D)])
    ksk = ctx.ksk_create(key, L - 1, K)
    gel = G.galois_element_for_rotation(n, 1)
    rk = ctx.ksk_create(key, L - 1, K, gel)
    del key
    quad, out = G.DeviceBuffer((3, L, n)), G.DeviceBuffer((2, L, n))
    stream = G.Stream()

    def mr():
        ctx.mult(lhs, rhs, L, 40, 40, quad)
        ctx.relinearize(quad, L, 80, ksk, out)

    for _ in range(args.warmup):
        mr()
    ms_mr = time_it(G, mr, args.steps) / args.steps
    ms_rot = time_it(G, lambda: ctx.rotate(lhs, L, 40, rk, out), args.steps) / args.steps
    ntt = ntt_entry(G, ctx, p, n, L, K)
    G.profile_enable(True)
    mr()
    prof = G.profile_report()
    G.profile_enable(False)
    value = 1000.0 / ms_mr
    log(f"[bench] 64-bit integer path, N=2^16, {L} primes < 2^60 + {K} aux, D={D}: mult+relin {ms_mr:.3f} ms "
        f"({value:.1f} ops/s), rotate {ms_rot:.3f} ms")
    for k, (cnt, tot) in sorted(prof.items(), key=lambda kv: -kv[1][1]):
        log(f"[bench]   {k:16s} launches={cnt:5d} total={tot:9.3f} ms")
    line = {
        "metric": "per-op ct x ct mult+relin ops/s with 60-bit moduli (64-bit integer arithmetic), N=2^16",
        "value": round(value, 2), "unit": "ops/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_mr, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 residues (64-bit integer Montgomery arithmetic)", "data": "synthetic (seeded uniform residues)",
        "config": {"workload": f"per-op mult+relinearize, N=2^16, L={L} primes < 2^60, K={K} aux, D={D} digits",
                   "rotate_ms": round(ms_rot, 4)},
        "roofline": None, "cpu_baseline": None, "ntt": ntt,
        "kernels_ms_per_op": {k: round(v[1], 4) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])},
    }
    print(json.dumps(line), flush=True)


def bsgs_main(args, G, PR):
    """Batched encrypted BSGS diagonal matvec: N=2^14 (8192 slots), a dense dim x dim matrix U[-1,1]
    (baby step = giant step = sqrt(dim)), 12 limbs, a batch of encrypted vectors U[-1,1] repeated to
    fill the slots; decode error against FP64."""
    n, L, dim, B = 1 << 14, 12, args.dim, args.vectors
    p = PR.config_batched(n=n, limbs=L, aux=2)
    ctx = G.Context(p)
    ctx.set_keyswitch_method(KS_METHOD[args.method])
    rs = np.random.default_rng(0)
    rng = G.Rng(4)
    t0 = time.perf_counter()
    s = ctx.sample_sparse_ternary(rng, 64, L)
    pk = ctx.keygen_public(rng, s, L, sigma=3.2)
    steps = G.bsgs_rotation_steps(dim)
    keys = {st: ctx.keygen_rotation(rng, s, st, L - 1, sigma=3.2) for st in steps}
    log(f"[bench] bsgs: {len(keys)} rotation keys generated on the device in {time.perf_counter() - t0:.1f} s")
    matrix = rs.uniform(-1, 1, (dim, dim))
    t = time.perf_counter()
    diag = G.Diag(ctx, matrix, L - 1, 40)
    log(f"[bench] {dim} diagonals encoded (baby step {diag.baby_step}, {diag.giant_steps} giant steps) in "
        f"{time.perf_counter() - t:.1f} s")
    vecs = rs.uniform(-1, 1, (B, dim))
    cts = G.DeviceBuffer((B, 2, L, n))
    CT = 2 * L * n * 8
    for v in range(B):
        ct = ctx.encrypt(rng, ctx.encode(np.tile(vecs[v], n // 2 // dim), L, 40), pk, L, L, sigma=3.2)
        G.lib().rck_memcpy_dtod(cts.ptr + v * CT, ct.ptr, CT, None)
    out = G.DeviceBuffer((B, 2, L - 1, n))
    stream = G.Stream()
    for _ in range(args.warmup):
        sc = ctx.bsgs_matvec(cts, B, L, 40, diag, keys, out, stream.ptr)
    ms = time_it(G, lambda: ctx.bsgs_matvec(cts, B, L, 40, diag, keys, out, stream.ptr), args.steps, stream)
    ms /= args.steps
    one = G.DeviceBuffer((1, 2, L - 1, n))
    ctx.bsgs_matvec(cts, 1, L, 40, diag, keys, one, stream.ptr)
    ms1 = time_it(G, lambda: ctx.bsgs_matvec(cts, 1, L, 40, diag, keys, one, stream.ptr), 2, stream) / 2
    P = 2 * (L - 1) * n * 8
    errs = []
    for v in range(min(B, 4)):
        dec = ctx.decrypt(out.ptr + v * P, L - 1, s)
        got = ctx.decode(dec, L - 1, sc).real
        errs.append(float(np.max(np.abs(got[:dim] - matrix @ vecs[v]))))
    G.profile_enable(True)
    ctx.bsgs_matvec(cts, B, L, 40, diag, keys, out, stream.ptr)
    prof = G.profile_report()
    G.profile_enable(False)
    value = B * 1000.0 / ms
    log(f"[bench] bsgs: {ms:.2f} ms per batch of {B} = {value:.1f} matvecs/s; one vector alone {ms1:.2f} ms "
        f"({ms1 * B / ms:.2f}x); decode error max {max(errs):.2e}")
    cpu = None if args.no_cpu_baseline else cpu_baseline_bsgs(args.cpu_threads, dim)
    line = {
        "metric": "encrypted BSGS diagonal matvec throughput (matvecs/s), batched",
        "value": round(value, 3), "unit": "matvecs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 residues (FP64-exact modular arithmetic)", "data": "synthetic (seeded U[-1,1] matrix and vectors)",
        "config": {"workload": f"N=2^14, {dim}x{dim} matrix, baby/giant step {diag.baby_step}, 12 limbs (primes > 2^50,"
                               f" K=2), {B} vectors per batch, {len(keys)} rotation keys",
                   "one_vector_ms": round(ms1, 3), "batched_vs_one_by_one": round(ms1 * B / ms, 3),
                   "decode_err_max": max(errs)},
        "roofline": None, "cpu_baseline": cpu,
        "kernels_ms_per_step": {k: round(v[1], 4) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])},
    }
    print(json.dumps(line), flush=True)


def cpu_baseline_bsgs(threads, dim):
    """The oracle on a bounded sample of one matvec at N=2^14, 12 limbs: 2 rotations and bs plaintext
    products (one giant step), extrapolated to the (bs - 1) + (gs - 1) rotations and dim products."""
    from oracle import oracle as O
    from rnsckks import params as PR
    O.set_threads(threads)
    n, L = 1 << 14, 12
    p = PR.config_batched(n=n, limbs=L, aux=2)
    rs = np.random.default_rng(5)
    q = O.Basis(n, p.primes, p.roots)
    e = O.Basis(n, p.primes + p.aux_primes, p.roots + p.aux_roots)
    D = O.num_decomp_digits(q.total_modulus_bits())
    key = np.stack([uniform(p.primes + p.aux_primes, n, rs, 2) for _ in range(D)])
    ct = uniform(p.primes, n, rs, 2)
    pt = uniform(p.primes, n, rs, 1)[0]
    bs = 1
    while bs * bs < dim:
        bs += 1
    gs = -(-dim // bs)
    log(f"[bench] cpu baseline: 2 rotations + {bs} ct x pt products (N=2^14, 12 limbs) on {threads} threads")
    t = time.perf_counter()
    for _ in range(2):
        O.rotate(q, e, ct, O.galois_element_for_rotation(n, 1), key)
    t_rot = (time.perf_counter() - t) / 2
    t = time.perf_counter()
    for _ in range(bs):
        O.multiply_plain(q, ct, pt)
    t_pt = (time.perf_counter() - t) / bs
    per = (bs - 1 + gs - 1) * t_rot + dim * t_pt
    log(f"[bench] cpu baseline: {t_rot:.2f} s/rotation, {t_pt * 1e3:.0f} ms/product -> {per:.0f} s per matvec")
    return {"value": round(1.0 / per, 6), "unit": "matvecs/s", "cores": threads, "kind": "port",
            "sample": f"2 rotations + {bs} ct x pt products at N=2^14, 12 limbs (oracle, OpenMP), extrapolated to "
                      f"{bs - 1 + gs - 1} rotations + {dim} products per matvec"}


def cpu_baseline_ps(threads):
    """The oracle's Paterson-Stockmeyer restatement on one degree-7 request at N=2^15, 12 limbs
    (uniform ciphertext and keys: the arithmetic is the same for any residues)."""
    from oracle import oracle as O
    from oracle import ps as OPS
    from rnsckks import params as PR
    O.set_threads(threads)
    n, L = 1 << 15, 12
    p = PR.config_batched(n=n, limbs=L, aux=2)
    rs = np.random.default_rng(3)
    coeffs = list(rs.uniform(-1, 1, 8))
    keys = {}
    for lv in OPS.required_levels(L - 1, len(coeffs)):
        D = O.num_decomp_digits(O.Basis(n, p.primes[: lv + 1], p.roots[: lv + 1]).total_modulus_bits())
        keys[lv] = np.stack([uniform(p.primes[: lv + 1] + p.aux_primes, n, rs, 2) for _ in range(D)])
    R = OPS.Ring(n, p.primes, p.roots, p.aux_primes, p.aux_roots, keys)
    ct = uniform(p.primes, n, rs, 2)
    log(f"[bench] cpu baseline: 1 degree-7 PS request (N=2^15, 12 limbs) on {threads} threads")
    t = time.perf_counter()
    OPS.ps_eval(R, (ct, 40), coeffs)
    dt = time.perf_counter() - t
    log(f"[bench] cpu baseline {dt:.2f} s/request")
    return {"value": round(1.0 / dt, 5), "unit": "requests/s", "cores": threads, "kind": "port",
            "sample": "1 degree-7 Paterson-Stockmeyer request at N=2^15, 12 limbs (oracle, OpenMP)"}


def cpu_baseline(p, threads, host=None):
    """The oracle (C restatement of the crate's CPU path: u128 % arithmetic, 3 NTTs per
    multiply_ntt, coefficient-form keys, per-digit mixed-radix lift, OpenMP over limbs and
    coefficients) timed on mult + relinearize of the same workload: ops are repeated until 10 s of
    CPU time have passed (at most 3)."""
    from oracle import oracle as O
    n, L, K = p.poly_degree, len(p.primes), len(p.aux_primes)
    O.set_threads(threads)
    qb = O.Basis(n, p.primes, p.roots)
    eb = O.Basis(n, p.primes + p.aux_primes, p.roots + p.aux_roots)
    rs = np.random.default_rng(1)
    D = O.num_decomp_digits(qb.total_modulus_bits())
    lhs, rhs = uniform(p.primes, n, rs, 2), uniform(p.primes, n, rs, 2)
    key = np.stack([uniform(p.primes + p.aux_primes, n, rs, 2) for _ in range(D)])
    log(f"[bench] cpu baseline: mult+relinearize on {threads} threads ({(host or {}).get('model')})")
    t = time.perf_counter()
    ops = 0
    while ops < 3 and (ops == 0 or time.perf_counter() - t < 10.0):
        quad = O.multiply_ciphertexts(qb, lhs, rhs)
        O.relinearize(qb, eb, quad, key)
        ops += 1
    dt = (time.perf_counter() - t) / ops
    log(f"[bench] cpu baseline {dt:.2f} s/op over {ops} ops")
    out = {"value": round(1.0 / dt, 5), "unit": "ops/s", "cores": threads, "kind": "port",
           "sample": f"{ops} mult+relinearize at N=2^16, L=24, K=2, D=40 (oracle, OpenMP over {threads} threads)"}
    if host:
        out.update({"cpu_model": host["model"], "cores_available": host["affinity"], "cores_visible": host["visible"]})
    return out


if __name__ == "__main__":
    main()
