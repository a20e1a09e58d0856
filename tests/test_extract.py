"""The gamma / Omega_p / saturation extractor (host code, no GPU).

No reference counterpart exists (the reference only writes the series), so
the extractor is pinned by synthetic series with known answers."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def dfk():
    import diskfft_amd
    return diskfft_amd


def _series(gamma, omega, a0=1e-4, sat=0.24, t1=20.0, dt=0.25, phase0=0.3):
    t = np.arange(0, t1 + 1e-9, dt)
    # exponential growth that levels off at `sat` (logistic in the amplitude)
    a = a0 * np.exp(gamma * t) / (1 + (a0 / sat) * (np.exp(gamma * t) - 1))
    ph = np.angle(np.exp(1j * (phase0 - 2 * omega * t)))      # wrapped to (-pi, pi]
    c0 = np.ones_like(t)
    return np.stack([t, c0, a, ph], axis=1)


@pytest.mark.parametrize("gamma,omega", [(0.72, 1.57), (0.36, 1.2), (1.1, 0.8)])
def test_recovers_growth_and_pattern_speed(dfk, gamma, omega):
    s = _series(gamma, omega, t1=max(20.0, 25.0 / gamma))
    g = dfk.extract_growth(s, lo_factor=3.0, hi_frac=0.01)
    assert g["grew"]
    # the logistic bend lowers the slope by < 1% at 1% of saturation
    assert abs(g["gamma"] - gamma) / gamma < 1e-2
    assert abs(g["omega_p"] - omega) / omega < 1e-12
    assert abs(g["sat"] - 0.24) < 5e-3


def test_exact_exponential_and_retrograde_phase(dfk):
    t = np.arange(0, 10, 0.25)
    a = 1e-3 * np.exp(0.5 * t)
    ph = np.angle(np.exp(1j * (+2 * 0.7 * t)))     # retrograde: Omega_p = -0.7
    s = np.stack([t, np.full_like(t, 2.0), 2.0 * a, ph], axis=1)
    g = dfk.extract_growth(s, lo_factor=3.0, hi_frac=0.9)
    assert abs(g["gamma"] - 0.5) < 1e-12
    assert abs(g["omega_p"] + 0.7) < 1e-12
    # |C2|/|C0| is what saturates
    assert abs(g["sat"] - np.mean(a[len(a) - len(a) // 4:])) < 1e-15


def test_no_growth_uses_whole_series(dfk):
    t = np.arange(0, 5, 0.25)
    a = 1e-3 * np.exp(-0.1 * t)
    s = np.stack([t, np.ones_like(t), a, np.zeros_like(t)], axis=1)
    g = dfk.extract_growth(s)
    assert not g["grew"] and g["n_fit"] == len(t)
    assert abs(g["gamma"] + 0.1) < 1e-12 and g["omega_p"] == 0.0


def test_too_short_raises(dfk):
    with pytest.raises(ValueError):
        dfk.extract_growth(np.zeros((3, 4)))


def test_window_cap_keeps_first_growth(dfk):
    """An edge mode grows e^{9t} from 2e-5 to a plateau; much later a second,
    larger peak appears.  With hi_factor the fit window stops inside the first
    growth (exact rate and pattern speed); without it the window stretches to
    0.3 of the late peak, across the plateau."""
    t = np.arange(0, 40, 0.01)
    la = np.minimum(np.log(2e-5) + 9.0 * t, np.log(2e-2))          # growth, then plateau
    la = np.where(t > 30, la + 2.5 * (t - 30) / 10.0, la)           # late rise to ~0.24
    ph = np.angle(np.exp(1j * (0.4 - 2 * 13.5 * t)))
    s = np.stack([t, np.ones_like(t), np.exp(la), ph], axis=1)
    g = dfk.extract_growth(s, lo_factor=10.0, hi_factor=100.0, hi_frac=0.3, sat_frac=0.5,
                           peak_window=2.0)
    assert g["grew"] and g["t_hi"] < 0.8
    # the first saturation: largest a within 2 time units of the window start
    w = (t >= g["t_lo"]) & (t <= g["t_lo"] + 2.0)
    assert g["sat_peak"] == np.max(np.exp(la)[w]) and abs(g["sat_peak"] - 2e-2) < 1e-15
    assert abs(g["gamma"] - 9.0) < 1e-9
    assert abs(g["omega_p"] - 13.5) < 1e-9
    assert abs(g["sat"] - np.mean(np.exp(la[len(t) - len(t) // 2:]))) < 1e-15
    g_old = dfk.extract_growth(s, lo_factor=10.0, hi_frac=0.3)
    assert g_old["t_hi"] > 30 and g_old["gamma"] < 1.0


def test_ex_defaults_reproduce_extractor(dfk):
    s = _series(0.72, 1.57)
    a = dfk.extract_growth(s, lo_factor=3.0, hi_frac=0.05)
    b = dfk.extract_growth(s, lo_factor=3.0, hi_frac=0.05, hi_factor=np.inf, sat_frac=0.25 + 1e-17)
    assert b.pop("sat_peak") == np.max(s[:, 2]) and a == b
    with pytest.raises(ValueError):
        dfk.extract_growth(s, sat_frac=0.0)
