"""Checks on the recorded long runs (tests/longrun/run_growth.py on the MI355X
box; GPU loop and CPU oracle from the same initial state).  No GPU needed:
these assert what the committed record shows."""
import json
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REC = os.path.join(HERE, "longrun", "results_lka98.json")

# published by the reference for the same setups (diskfft/README.md:27,37; README.md:42)
PUBLISHED_GAMMA = {"lka128_e0.04": 0.36, "lka128_e0.02": 0.53, "lka128_e0.01": 0.64,
                   "lka128_e0.005": 0.69, "lka128_e0.0025": 0.757, "lka256": 0.717}


@pytest.fixture(scope="module")
def rec():
    if not os.path.exists(REC):
        pytest.skip("no recorded long runs")
    return {r["case"]: r for r in json.load(open(REC))}


def test_gpu_matches_oracle_within_one_percent(rec):
    for name, r in rec.items():
        assert r["steps_gpu"] == r["steps_oracle"], name
        for q in ("gamma", "omega_p"):
            assert r["rel_diff"][q] < 0.01, (name, q)


def test_growth_close_to_published(rec):
    # different fit windows (the reference does not state its own); within 20%
    for name, g in PUBLISHED_GAMMA.items():
        assert abs(rec[name]["gpu"]["gamma"] - g) / g < 0.2, name
    assert abs(rec["lka128_e0.0025"]["gpu"]["omega_p"] - 1.49) / 1.49 < 0.02
    assert abs(rec["lka256"]["gpu"]["sat"] - 0.24) / 0.24 < 0.05


def test_mass_conserved(rec):
    for name, r in rec.items():
        assert abs(r["mass_drift_gpu"]) < 1e-9, name
