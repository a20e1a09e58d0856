"""The oracle's restatement of the reference driver (IC, loop, potential,
forces, FFT of any length, C_m output) built with AddressSanitizer and
UndefinedBehaviorSanitizer (oracle/Makefile target oracle_diskdisk_asan) and
run at small grids with several OpenMP threads: no invalid access, leak or
undefined behaviour.  (ThreadSanitizer is not used: GCC's OpenMP runtime is
not instrumented, so TSan cannot see the barriers that end each parallel
loop and reports every later write of a shared variable; every parallel
loop of the oracle writes disjoint rows or columns.)"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "oracle", "oracle_diskdisk_asan")


@pytest.mark.parametrize("args", [
    ("32", "48", "0.02", "0.5", "1", "1e-3", "4"),     # radix-2/3 transforms
    ("35", "40", "0.02", "0.3", "1", "1e-3", "3"),     # factor 5 and 7 (2Nu = 70)
    ("17", "46", "0.02", "0.3", "1", "1e-3", "2"),     # primes 17 and 23
    ("24", "24", "0.02", "0.3", "0", "1e-3", "4"),     # no self-gravity
    ("3", "4", "0.02", "0.2", "1", "1e-3", "1"),       # smallest grid
])
def test_oracle_driver_under_sanitizers(tmp_path, args):
    assert os.path.exists(EXE), "build first: make -C oracle"
    env = dict(os.environ, OMP_NUM_THREADS=args[-1], ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([EXE, *args], cwd=tmp_path, capture_output=True, text=True, env=env,
                       timeout=120)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
    last = r.stdout.strip().splitlines()[-1]
    assert last.startswith("final t=") and "steps=" in last
    assert os.path.getsize(os.path.join(tmp_path, "diskfft_modes.dat")) > 0
