"""Host-side logic around the kernels (no GPU): the autotuner's candidate
filter and cache application, the rocprof traffic calibration, bench.py's
byte counts and the ensemble draw."""
import csv
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_candidates_respect_lds_budget():
    import autotune
    for Nu, Nphi in [(128, 128), (512, 1024), (37, 52)]:
        tiles, prs, cbs = autotune.candidates(Nu, Nphi)
        assert tiles and prs and cbs
        for ti, tj in tiles:
            assert 8 * (3 * (ti + 4) * (tj + 4) + 3 * (ti + 1) * (tj + 4)) <= autotune.LDS_MAX
            assert ti <= Nu and tj <= Nphi
        assert all((2 * p + 1) * Nphi * 16 <= autotune.LDS_MAX and 2 * p <= Nu for p in prs)
        assert all((2 * c + 1) * 2 * Nu * 16 <= autotune.LDS_MAX for c in cbs)


class _FakeEnsemble:
    """The calls autotune.apply makes, recorded."""

    def __init__(self, shapes):
        self.n, self._shapes, self.calls = len(shapes), shapes, []

    def shape(self, d):
        return self._shapes[d]

    def set_update_kernel(self, k):
        self.calls.append(("kind", k))

    def set_shapes(self, ti, tj, pr, cb, d):
        self.calls.append(("shapes", d, ti, tj, pr, cb))

    def set_fft_options(self, rows, cols, wave, d):
        self.calls.append(("fft", d, rows, cols, wave))

    def set_threads(self, *t):
        self.calls.append(("threads",) + t)

    def set_row_transform(self, r):
        self.calls.append(("rows", r))

    def set_fused_rows(self, t):
        self.calls.append(("fused", t))


def test_apply_votes_kind_and_threads_by_zones(tmp_path):
    import autotune
    cache = {
        "128x128:throughput:gfx950": {"kind": 1, "ti": 16, "tj": 60, "pr": 4, "cb": 4,
                                      "threads": [128, 128, 128], "fft": [1, 0, 1],
                                      "rows_inplace": 1},
        "512x512:throughput:gfx950": {"kind": 0, "ti": 8, "tj": 32, "pr": 2, "cb": 1,
                                      "threads": [256, 256, 256]},
    }
    path = tmp_path / "cache.json"
    path.write_text(json.dumps(cache))
    ens = _FakeEnsemble([(128, 128), (512, 512), (512, 512), (96, 96)])
    n = autotune.apply(ens, path=str(path))
    assert n == 3                                    # 96x96 has no entry
    assert ("kind", 0) in ens.calls                  # 2 x 512^2 outweigh 128^2
    assert ("rows", 0) in ens.calls and ("rows", 1) not in ens.calls
    assert ("threads", 256, 256, 256) in ens.calls
    assert ("shapes", 0, 16, 60, 4, 4) in ens.calls and ("fft", 0, 1, 0, 1) in ens.calls
    assert ("fft", 1, 1, 1, 0) in ens.calls          # defaults when the entry has none
    assert not any(c[0] == "shapes" and c[1] == 3 for c in ens.calls)


def test_apply_fused_rows_vote(tmp_path):
    """Latency entries tuned for fused row kernels: the handle runs them when
    the disks tuned for them hold most zones, and those disks then get their
    fused row-block height; otherwise every disk keeps its tile."""
    import autotune
    fz = {"threads": 512, "ti": 2, "step_ms": 0.018, "step4_ms": 0.022}
    cache = {
        "64x64:latency:gfx950": {"kind": 0, "ti": 8, "tj": 16, "pr": 1, "cb": 1,
                                 "threads": [256, 128, 256], "fused": fz},
        "96x96:latency:gfx950": {"kind": 0, "ti": 4, "tj": 32, "pr": 1, "cb": 1,
                                 "threads": [256, 128, 256], "fused": {}},
    }
    path = tmp_path / "cache.json"
    path.write_text(json.dumps(cache))
    ens = _FakeEnsemble([(64, 64), (64, 64), (64, 64), (96, 96)])   # 12288 vs 9216 zones
    autotune.apply(ens, path=str(path), regime="latency")
    assert ("fused", 512) in ens.calls
    assert ("shapes", 0, 2, 16, 1, 1) in ens.calls and ("shapes", 3, 4, 32, 1, 1) in ens.calls
    ens = _FakeEnsemble([(64, 64), (96, 96), (96, 96)])
    autotune.apply(ens, path=str(path), regime="latency")
    assert not any(c[0] == "fused" for c in ens.calls)
    assert ("shapes", 0, 8, 16, 1, 1) in ens.calls


def _counter_csv(d, rows):
    os.makedirs(d)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, ["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"], r)))


def test_traffic_calibration(tmp_path):
    import prof_summary as ps
    cf, cw = str(tmp_path / "cf"), str(tmp_path / "cw")
    _counter_csv(cf, [("void copy_kernel<double>(double const*)", 1024, "FETCH_SIZE", 131072.0)])
    _counter_csv(cw, [("void copy_kernel<double>(double const*)", 1024, "WRITE_SIZE", 262144.0)])
    kf, kw = str(tmp_path / "kf"), str(tmp_path / "kw")
    # a small per-disk call and the loop's launch over all disks
    _counter_csv(kf, [("k_cols_conv(int)", 100, "FETCH_SIZE", 10.0),
                      ("k_cols_conv(int)", 9000, "FETCH_SIZE", 500.0)])
    _counter_csv(kw, [("k_cols_conv(int)", 100, "WRITE_SIZE", 20.0),
                      ("k_cols_conv(int)", 9000, "WRITE_SIZE", 600.0)])
    out = str(tmp_path / "t.json")
    ps.traffic(out, "w", kf, kw, cf, cw, 256, largest=True)
    t = json.load(open(out))
    assert t["fetch_factor"] == 2.0 and t["write_factor"] == 1.0
    k = t["kernels"]["k_cols_conv"]
    assert k["fetch_kib"] == 500.0 and k["traffic_bytes"] == 1024.0 * (2.0 * 500.0 + 600.0)
    ps.traffic(out, "w", kf, kw, cf, cw, 256)
    assert json.load(open(out))["kernels"]["k_cols_conv"]["fetch_kib"] == 255.0


def test_bench_byte_counts():
    import bench
    Nu, Nphi = 256, 256
    kb = bench.kernel_bytes(Nu, Nphi)
    N, S = Nu * Nphi, Nu * (Nphi // 2 + 1)
    # the step's algorithmic bytes are the update's state in/out plus the
    # spectra passes and K-hat; per-kernel counts add Phi and the S packing
    assert bench.b_alg(Nu, Nphi) == 48 * N + 80 * S
    assert kb["update"] == 56 * N and kb["rows_r2c"] == kb["rows_c2r"] == 8 * N + 16 * S
    assert kb["cols_conv"] == 32 * S + 8 * (Nu + 1) * (Nphi // 2 + 1)
    assert 87 < bench.b_alg(512, 1024) / (512 * 1024) < 89


def test_ensemble_draw_is_the_configured_distribution():
    import ensemble_cfg
    a = ensemble_cfg.draw(64, seed=2026)
    b = ensemble_cfg.draw(64, seed=2026)
    assert a == b and len({d["seed"] for d in a}) == 64
    for d in a:
        assert (d["Nu"], d["Nphi"]) in ensemble_cfg.GRIDS
        assert 0.3 <= d["mass_ratio"] <= 1.5 and 1.0 <= d["Qmin"] <= 2.0
        assert 0.05 ** 2 <= d["eps"] <= 0.2 ** 2 and -1.75 <= d["slope"] <= -1.25
    zones = sum(d["Nu"] * d["Nphi"] for d in a)
    assert zones == 7700480
    assert ensemble_cfg.draw(64, seed=2027) != a


def test_surrogate_fit_error_metrics():
    import surrogate_fit
    true = np.array([[1.0, 10.0, 0.1], [2.0, 20.0, 0.2]])
    pred = true * np.array([1.02, 0.99, 1.05])
    e = surrogate_fit.errors(pred, true, np.array([1.0, 10.0, 0.1]))
    assert abs(e["gamma"]["median_rel_err"] - 0.02) < 1e-12
    assert e["omega_p"]["frac_rel_le_3pct"] == 1.0 and e["sat"]["frac_rel_le_3pct"] == 0.0
