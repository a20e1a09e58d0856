"""The library's host-only code (heavy-disks_amd/csrc/dfk_model.cpp: disk
models, equilibrium and Toomre-Q solve, mass, growth extractor, random
streams) under AddressSanitizer and UndefinedBehaviorSanitizer: the harness
heavy-disks_amd/csrc/model_check.cpp calls every entry point over small and
odd grids, unreachable Q targets and short, flat, non-finite or non-growing
extractor inputs (Makefile target bin/model_check_asan)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "heavy-disks_amd", "bin", "model_check_asan")


def test_host_model_code_under_sanitizers():
    assert os.path.exists(EXE), "build first: make -C heavy-disks_amd"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([EXE], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
    assert r.stdout.strip() == "model_check: ok"
