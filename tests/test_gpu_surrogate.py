"""The surrogate: its error bar on held-out hydro runs, the 10^5-point scan
with hydro fallback, and what it gains over the method without it.

Data: the table of 4352 batched hydro runs in tests/golden/surrogate_runs.npz
(parameters and extracted targets only): 256 x 256 power-law disks drawn from
the ensemble's parameter box (heavy-disks_amd/surrogate_data.py --tlim 8; the
73 runs of tests/golden/surrogate_extended_runs.txt, whose m = 2 growth had
not completed by t = 8, re-run to 20 outer orbits), exported by
surrogate_fit.py --export (its docstring gives the recipe).
The ensemble (5 members, BF16 weights, FP32 accumulation) is trained on the
GPU by the library's own kernels on runs 0..4095, with the configuration of
tests/longrun/surrogate_eval.json (6000 epochs, batch 256, lr 2e-3; the
training is deterministic), and judged on the 256 held-out runs.  Targets:
gamma and Omega_p of the first m = 2 growth and its saturation amplitude
sat_peak, the largest |C2|/|C0| within 2 time units of the growth window's
start (dfk_extract_growth_ex).

Statistic per output (relative error |p - h| / |h| over the held-out runs
whose |h| exceeds 5% of the training targets' 5-95% range, as
surrogate_fit.errors reports it):
  - with the hydro fallback where the members' spread exceeds 5% (those
    points take the hydro value, as the surrogate-with-fallback scan does): mean <= 3% for
    each output (recorded 0.24%, 0.06%, 2.1%);
  - surrogate alone: median <= 3% for each output (recorded 0.30%, 0.09%,
    2.7%);
  - the surrogate beats nearest-neighbour and linear interpolation of the
    same training runs (the screening one would do without a network).
The 10^5-point scan over the parameter box reports the fraction of points
that would fall back to the hydro (recorded 38%) and its throughput.
"""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "surrogate_runs.npz")
N_TRAIN = 4096


@pytest.fixture(scope="module")
def trained():
    import diskfft_amd
    if diskfft_amd.device_count() < 1:
        pytest.fail("no HIP device visible")
    import surrogate
    import surrogate_fit
    print("case surrogate: 5 members, 6000 epochs on 4096 runs", flush=True)
    X, Y, ok = surrogate_fit.load([TABLE])
    fin = ok & np.all(np.isfinite(Y), axis=1)
    idx = np.arange(len(X))
    tr, te = fin & (idx < N_TRAIN), fin & (idx >= N_TRAIN)
    assert tr.sum() >= 4000 and te.sum() >= 250
    s = surrogate.Surrogate(members=5, seed=1)
    t0 = time.perf_counter()
    hist = s.fit(X[tr], Y[tr], epochs=6000, batch=256, lr=2e-3)
    print(f"  trained in {time.perf_counter() - t0:.1f} s, loss {hist[0]:.3f} -> {hist[-1]:.5f}",
          flush=True)
    scale = np.quantile(Y[tr], 0.95, axis=0) - np.quantile(Y[tr], 0.05, axis=0)
    return surrogate, surrogate_fit, s, hist, X[tr], Y[tr], X[te], Y[te], scale


def test_surrogate_error_bar_on_held_out_runs(trained):
    sg, fit, s, hist, Xtr, Ytr, Xte, Yte, scale = trained
    assert hist[-1] < 0.01 * hist[0]
    mean, spread, fb = s.predict(Xte)
    e = fit.errors(mean, Yte, scale)
    e_fb = fit.errors(np.where(fb[:, None], Yte, mean), Yte, scale)
    for n in fit.NAMES:
        print(f"  {n}: median {e[n]['median_rel_err']:.4f} mean {e[n]['mean_rel_err']:.4f} "
              f"within 3% {e[n]['frac_rel_le_3pct']:.3f} | with fallback: mean "
              f"{e_fb[n]['mean_rel_err']:.4f} within 3% {e_fb[n]['frac_rel_le_3pct']:.3f}",
              flush=True)
        assert e[n]["median_rel_err"] <= 0.03
        assert e_fb[n]["mean_rel_err"] <= 0.03
        # the fallback only replaces predictions by the hydro's own values
        assert e_fb[n]["mean_rel_err"] <= e[n]["mean_rel_err"] + 1e-15
    print(f"  held-out fallback fraction {fb.mean():.3f}", flush=True)
    assert fb.mean() < 0.5


def test_surrogate_beats_interpolation(trained):
    sg, fit, s, hist, Xtr, Ytr, Xte, Yte, scale = trained
    mean, _, _ = s.predict(Xte)
    nn, lin = fit.baselines(Xtr, Ytr, Xte)

    def med(p):
        return np.median(np.abs(p - Yte) / np.abs(Yte), axis=0)

    m_mlp, m_nn, m_lin = med(mean), med(nn), med(lin)
    print("  median relative error (gamma, Omega_p, sat): mlp", m_mlp, "nearest", m_nn,
          "linear", m_lin, flush=True)
    assert np.all(m_mlp < m_nn) and np.all(m_mlp < m_lin)


def test_surrogate_scan_1e5_points(trained):
    """The 10^5-point scan of the parameter box: predictions inside the
    training targets' range where the members agree, and the fraction that
    would go to the hydro (member spread > 5%)."""
    sg, fit, s, hist, Xtr, Ytr, Xte, Yte, scale = trained
    rng = np.random.default_rng(2026)
    Xs = sg.LO + (sg.HI - sg.LO) * rng.random((100000, 4))
    s.predict(Xs[:1000])
    t0 = time.perf_counter()
    mean, spread, fb = s.predict(Xs)
    dt = time.perf_counter() - t0
    print(f"  scan: {len(Xs) / dt:.3e} points/s, fallback fraction {fb.mean():.3f}", flush=True)
    assert mean.shape == (100000, 3) and np.all(np.isfinite(mean))
    assert 0.0 < fb.mean() < 0.5
    agree = ~fb
    lo, hi = Ytr.min(0), Ytr.max(0)
    span = hi - lo
    inside = np.all((mean[agree] >= lo - 0.1 * span) & (mean[agree] <= hi + 0.1 * span), axis=1)
    assert inside.mean() > 0.99


def test_training_table_reproduces():
    """Four held-out runs of the table re-made now by the batched hydro (256 x
    256, t = 8, their own handle): the same parameters, and the targets the
    table holds, to round-off (1e-9 relative).  The table was made by this
    build's kernels: a build whose arithmetic differs in the last bits (a
    tuned transform variant, another form of the update) reproduces gamma
    and Omega_p, fitted in the linear growth, to ~1e-12, but the saturation
    peak comes after the growth turns nonlinear, which amplifies last-bit
    differences of the state (up to a few per cent in the runs that saturate
    chaotically); such a change regenerates the table (surrogate_fit.py)."""
    import diskfft_amd
    if diskfft_amd.device_count() < 1:
        pytest.fail("no HIP device visible")
    import surrogate_data
    print("case surrogate table: runs 4096..4099 re-made", flush=True)
    tab = np.load(TABLE)
    r = surrogate_data.run_shard(4096, 4, 256, 256, tlim=8.0, verbose=False)
    assert r["ok"].all() and np.array_equal(r["X"], tab["X"][4096:4100])
    assert not tab["tlim_extended"][4096:4100].any()
    rel = np.abs(r["Y"] - tab["Y"][4096:4100]) / np.abs(tab["Y"][4096:4100])
    print(f"  max relative difference from the table (gamma, Omega_p, sat): "
          f"{rel.max(axis=0)}", flush=True)
    assert rel.max() < 1e-9, (r["Y"], tab["Y"][4096:4100])
