"""Disk models and initial conditions (host code in libdiskfft, no GPU).

The LKA98 ring restates diskmain.cpp:28-63 and is checked bit for bit against
the oracle's restatement given the same self-gravity field.  The power-law
model has no reference counterpart: it is pinned by what it promises (mass
ratio, min Q over the disk body, noise rms, determinism, closed forms).
"""
import ctypes

import numpy as np
import pytest

from helpers import grid


@pytest.fixture(scope="module")
def dfk():
    import diskfft_amd
    return diskfft_amd


def _oracle_grav(orc, m, Nu, Nphi, eps):
    k = orc.Kalnajs(Nu, Nphi, m.Rin, m.Rout, eps, m.G)
    return lambda S: k.forces(k.potential(S))


def test_lka98_matches_oracle_ic(orc, dfk):
    Nu, Nphi, eps, amp = 48, 32, 0.02, 2e-3
    m = dfk.model_defaults(dfk.MODEL_LKA98, amp=amp)
    p, S, U, W, mD, _ = dfk.model_ic(m, Nu, Nphi, eps,
                                     potential_forces=_oracle_grav(orc, m, Nu, Nphi, eps))
    h = orc.Hydro(Nu, Nphi)
    k = orc.Kalnajs(Nu, Nphi, 0.25, 1.0, eps)
    orc.ic_lka98(h, k, 1, amp)
    for a, b in zip((S, U, W), h.state()):
        assert np.array_equal(a, b)
    assert mD == h.disk_mass()
    assert p["K"] == 0.25 and p["gmstar"] == 0.6 and p["Rin"] == 0.25


def test_powerlaw_mass_q_and_noise(orc, dfk):
    Nu, Nphi, eps = 128, 128, 0.01
    m = dfk.model_defaults(dfk.MODEL_POWERLAW)
    p, S, U, W, mD, Q = dfk.model_ic(m, Nu, Nphi, eps,
                                     potential_forces=_oracle_grav(orc, m, Nu, Nphi, eps))
    assert abs(mD - 1.0) < 1e-12                       # M_d / M* = 1, M* = GM*/G = 1
    R, phi, _ = grid(Nu, Nphi, 0.1, 1.0)
    body = (R >= 0.2) & (R <= 0.8)
    assert abs(Q[body].min() - 1.3) < 1e-9 and abs(m.Qmin_got - 1.3) < 1e-9
    assert p["K"] == m.K > 0
    sig0 = dfk.model_sigma0(m, Nu, Nphi)
    d = S / sig0[:, None] - 1.0
    assert 0.8e-3 < d.std() < 1.2e-3                    # 1e-3 rms seed
    # only m = 1..8 are seeded
    spec = np.abs(np.fft.rfft(d, axis=1)).mean(axis=0)
    assert spec[0] < 1e-12 and spec[9:].max() < 1e-12 and spec[1:9].min() > 1e-5
    assert np.all(U == 0) and np.all(W[:, 0] > 0)
    assert np.all(W == W[:, :1])


def test_powerlaw_deterministic_and_seeded(orc, dfk):
    Nu, Nphi, eps = 32, 16, 0.01
    a = dfk.model_ic(dfk.model_defaults(dfk.MODEL_POWERLAW), Nu, Nphi, eps, selfgrav=0)
    b = dfk.model_ic(dfk.model_defaults(dfk.MODEL_POWERLAW), Nu, Nphi, eps, selfgrav=0)
    c = dfk.model_ic(dfk.model_defaults(dfk.MODEL_POWERLAW, seed=43), Nu, Nphi, eps, selfgrav=0)
    assert np.array_equal(a[1], b[1]) and not np.array_equal(a[1], c[1])


def test_sigma0_shape_closed_form(dfk):
    Nu, Nphi = 64, 8
    m = dfk.model_defaults(dfk.MODEL_POWERLAW, slope=-1.25)
    s0 = dfk.model_sigma0(m, Nu, Nphi)
    R, _, _ = grid(Nu, Nphi, 0.1, 1.0)
    shape = R ** -1.25 * 0.5 * (1 + np.tanh((R - 0.15) / 0.02)) * 0.5 * (1 - np.tanh((R - 0.9) / 0.02))
    ratio = s0 / shape
    assert np.max(np.abs(ratio / ratio[0] - 1)) < 1e-13
    m2 = dfk.model_defaults(dfk.MODEL_LKA98)
    s2 = dfk.model_sigma0(m2, Nu, Nphi)
    R2, _, _ = grid(Nu, Nphi)
    assert np.max(np.abs(s2 - 0.372 * np.exp(-(R2 - 0.45) ** 2 / 0.05))) < 1e-15


def test_splitmix_and_normal(dfk):
    L = dfk._model_lib()
    st = ctypes.c_ulonglong(1234567)
    # SplitMix64 published test vector for seed 1234567
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    got = [L.dfk_splitmix64(ctypes.byref(st)) for _ in range(5)]
    assert got == want
    st = ctypes.c_ulonglong(7)
    x = np.array([L.dfk_normal(ctypes.byref(st)) for _ in range(20000)])
    assert abs(x.mean()) < 0.03 and abs(x.std() - 1) < 0.03


def test_unreachable_q_reports_best(orc, dfk):
    Nu, Nphi, eps = 64, 16, 0.01
    m = dfk.model_defaults(dfk.MODEL_POWERLAW, Qmin=50.0)
    dfk.model_ic(m, Nu, Nphi, eps, potential_forces=_oracle_grav(orc, m, Nu, Nphi, eps))
    assert 0 < m.Qmin_got < 50.0 and m.K > 0


@pytest.mark.parametrize("w,whole_ok,body_ok", [(0.02, False, True), (0.05, False, False),
                                                (0.1, False, False), (0.2, True, True)])
def test_taper_width_and_q_reachability(orc, dfk, w, whole_ok, body_ok):
    """The taper study's first finding (tests/longrun/taper_study.py; 128 x 128,
    M_d = M*, eps 0.01): min Toomre Q = 1.3 over the whole grid is reachable by
    some K only for the smoothest taper, w = 0.2 (where Sigma at both grid
    edges is still 38% of the power law); over the disk body [0.2, 0.8] only
    for the sharp default w = 0.02 (pressure at sharp tapers drives kappa^2
    negative before min Q reaches 1.3, at wider ones the tapers reach into the
    body).  Recomputed here with the host model code and the oracle's gravity."""
    Nu, Nphi, eps = 128, 128, 0.01
    got = {}
    for name, q_in, q_out in (("whole", 0.0, 10.0), ("body", 0.2, 0.8)):
        m = dfk.model_defaults(dfk.MODEL_POWERLAW, taper_w=w, q_in=q_in, q_out=q_out, Qmin=1e9)
        dfk.model_ic(m, Nu, Nphi, eps, potential_forces=_oracle_grav(orc, m, Nu, Nphi, eps))
        got[name] = m.Qmin_got
    print(f"w={w}: largest min Q whole grid {got['whole']:.3f}, body {got['body']:.3f}")
    assert (got["whole"] >= 1.3) == whole_ok
    assert (got["body"] >= 1.3) == body_ok
    assert dfk.model_defaults(dfk.MODEL_POWERLAW).taper_w == 0.02
