"""The CPU oracle built with AddressSanitizer and UndefinedBehaviorSanitizer (errors fatal) and
driven through every function by oracle/sanitize_check.c, which also checks the exact invariants
listed in its header.  CPU only; skipped when gcc or its sanitizer runtimes are missing."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-fopenmp", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def test_oracle_under_address_and_ub_sanitizers(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    exe = str(tmp_path / "sanitize_check")
    build = subprocess.run([gcc, *FLAGS, "-o", exe, os.path.join(ORACLE, "sanitize_check.c"),
                            os.path.join(ORACLE, "ckks_oracle.c"), "-lm"], capture_output=True, text=True)
    if build.returncode != 0 and ("asan" in build.stderr or "ubsan" in build.stderr):
        pytest.skip("sanitizer runtimes not available: " + build.stderr.strip()[-200:])
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0", OMP_NUM_THREADS="2")
    run = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "sanitize_check ok" in run.stdout
    assert "runtime error" not in run.stderr and "AddressSanitizer" not in run.stderr, run.stderr
