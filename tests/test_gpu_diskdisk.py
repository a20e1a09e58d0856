"""The diskdisk drop-in: same command line and output as the reference driver
(diskmain.cpp:16-115), checked against the oracle's restatement of that
driver (oracle/oracle_diskdisk) run with the same arguments."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPU_BIN = os.path.join(ROOT, "heavy-disks_amd", "bin", "diskdisk")
ORC_BIN = os.path.join(ROOT, "oracle", "oracle_diskdisk")


def _run(exe, args, cwd):
    r = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines(), np.loadtxt(os.path.join(cwd, "diskfft_modes.dat"), ndmin=2)


@pytest.mark.parametrize("args", [["64", "64", "0.02", "1.0", "1", "1e-3", "4"],
                                  ["48", "36", "0.01", "0.6", "0", "2e-3"],
                                  ["96", "80", "0.005", "0.8"],
                                  ["64", "34", "0.02", "0.3"],          # 34 = 2 x 17
                                  ["128", "128", "0.02", "0.5"],        # three-pass rows and columns
                                  ["192", "256", "0.01", "0.2"],        # three-pass, 2Nu = 384
                                  ["1024", "1024", "0.005", "0.0004"]])
def test_diskdisk_matches_reference_driver(tmp_path, args):
    print("case diskdisk", " ".join(args), flush=True)
    a, b = tmp_path / "gpu", tmp_path / "orc"
    a.mkdir()
    b.mkdir()
    out_g, mod_g = _run(GPU_BIN, args, str(a))
    out_o, mod_o = _run(ORC_BIN, args, str(b))
    # identical line structure and header
    assert len(out_g) == len(out_o)
    assert open(a / "diskfft_modes.dat").readline() == "# t  |C0|  |C2|  arg(C2)\n"
    assert mod_g.shape == mod_o.shape
    # same output times and step count; values to the printed precision
    assert np.array_equal(mod_g[:, 0], mod_o[:, 0])
    assert np.max(np.abs(mod_g[:, 1:3] - mod_o[:, 1:3]) / np.abs(mod_o[:, 1:3])) < 2e-6
    # arg(C2), printed %.6e: to that precision, modulo 2 pi
    dph = np.angle(np.exp(1j * (mod_g[:, 3] - mod_o[:, 3])))
    assert np.max(np.abs(dph) / np.maximum(np.abs(mod_o[:, 3]), 1e-3)) < 2e-6
    assert out_g[-1].split("steps=")[1] == out_o[-1].split("steps=")[1]
    for lg, lo in zip(out_g[:-1], out_o[:-1]):
        assert lg.split("|C0|")[0] == lo.split("|C0|")[0]      # "t=..." fields


def test_diskdisk_defaults_and_summary(tmp_path):
    print("case diskdisk defaults", flush=True)
    r = subprocess.run([GPU_BIN, "--steps", "30", "--summary", "s.txt"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    s = dict(line.split("=", 1) for line in open(tmp_path / "s.txt").read().split())
    assert s["Nu"] == "128" and s["Nphi"] == "128" and s["steps"] == "30"
    assert r.stdout.splitlines()[0].startswith("t=  0.000  |C0|=1.00000  |C2|=5.000000e-04")
