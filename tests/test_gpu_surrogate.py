"""The surrogate against the method without it.

Data: the table of 4352 batched hydro runs in tests/golden/surrogate_runs.npz
(parameters and extracted targets only, written by
heavy-disks_amd/surrogate_data.py and surrogate_fit.py --export).  The
ensemble is trained on the GPU by the library's own kernels on runs 0..4095
and judged on the 256 held-out runs, next to nearest-neighbour and linear
interpolation of the same training runs: the screening one would do from the
same hydro runs without a network.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "surrogate_runs.npz")
N_TRAIN = 4096


@pytest.fixture(scope="module")
def sg():
    import diskfft_amd
    if diskfft_amd.device_count() < 1:
        pytest.fail("no HIP device visible")
    import surrogate
    return surrogate


def test_surrogate_beats_interpolation(sg):
    import surrogate_fit
    print("case surrogate vs interpolation", flush=True)
    X, Y, ok = surrogate_fit.load([TABLE])
    fin = ok & np.all(np.isfinite(Y), axis=1)
    idx = np.arange(len(X))
    tr, te = fin & (idx < N_TRAIN), fin & (idx >= N_TRAIN)
    assert te.sum() >= 250
    s = sg.Surrogate(members=5, seed=1)
    hist = s.fit(X[tr], Y[tr], epochs=300, batch=256, lr=2e-3)
    assert hist[-1] < 0.2 * hist[0]
    mean, spread, fb = s.predict(X[te])
    nn, lin = surrogate_fit.baselines(X[tr], Y[tr], X[te])

    def med(p):
        return np.median(np.abs(p - Y[te]) / np.abs(Y[te]), axis=0)

    m_mlp, m_nn, m_lin = med(mean), med(nn), med(lin)
    print("  median relative error (gamma, Omega_p, sat): mlp", m_mlp, "nearest", m_nn,
          "linear", m_lin, "fallback fraction", fb.mean(), flush=True)
    assert np.all(m_mlp < m_nn) and np.all(m_mlp < m_lin)
    # the hydro fallback where the members disagree can only remove error
    with_fb = np.where(fb[:, None], Y[te], mean)
    assert np.all(med(with_fb) <= m_mlp + 1e-15)
