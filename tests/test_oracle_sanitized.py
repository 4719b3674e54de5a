"""CPU: the oracle built with AddressSanitizer and UndefinedBehaviorSanitizer (host code only) runs
every entry point the tests use -- harmonics, translators, trees with a depth cap, lists, the FMM
in the reference's and the corrected flags, treecode, BEM, direct sums -- without a report
(oracle/sanitize_check.cpp; the sanitizer runtime is linked statically, the binary goes to a
temporary directory)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_runs_clean_under_asan_and_ubsan(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ is not installed")
    exe = str(tmp_path / "sanitize_check")
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fopenmp", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-static-libasan",
           os.path.join(ROOT, "oracle", "sanitize_check.cpp"), "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, OMP_NUM_THREADS="4", ASAN_OPTIONS="detect_leaks=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "every entry point ran" in r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
