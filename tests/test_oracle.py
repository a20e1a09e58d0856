"""Pins for the CPU oracle (no GPU needed).

Each test checks the oracle against something other than its own formula:
the reference's own hydro.hpp compiled from source (oracle/_ref), the
reference's committed-here golden vectors, direct DFT / numpy.fft, the
test_kalnajs.cpp known answer, and closed forms.
"""
import os

import numpy as np
import pytest

from helpers import grid, relerr, rough_state

HERE = os.path.dirname(os.path.abspath(__file__))


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 16, 20, 30, 32, 48, 64, 77,
                               96, 128, 192, 256, 384, 512, 768, 1024])
def test_fft_matches_direct_dft_and_numpy(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = orc.fft(x, -1)
    k = np.arange(n)
    if n <= 256:
        W = np.exp(-2j * np.pi * np.outer(k, k) / n)
        assert relerr(X, W @ x) < 1e-13
    assert relerr(X, np.fft.fft(x)) < 1e-13
    xi = orc.fft(X, +1)
    assert relerr(xi, n * x) < 1e-13


def test_fft_matches_reference_radix2(orc):
    if not orc.RefHydro.available():
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    ref = orc.RefHydro(4, 4)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((16, 32)) + 1j * rng.standard_normal((16, 32))
    Xr = ref.fft2d(x, -1)
    # the oracle's 1-D transform applied along both axes
    Xo = np.array([orc.fft(r, -1) for r in x])
    Xo = np.array([orc.fft(c, -1) for c in Xo.T]).T
    assert relerr(Xo, Xr) < 1e-13


@pytest.mark.parametrize("Nu,Nphi", [(4, 6), (8, 8), (6, 10), (16, 24), (12, 9), (32, 32)])
def test_r2c_c2r_2d_match_numpy(orc, Nu, Nphi):
    k = orc.Kalnajs(Nu, Nphi)
    rng = np.random.default_rng(Nu * 100 + Nphi)
    x = rng.standard_normal((2 * Nu, Nphi))
    X = k.r2c2d(x)
    assert relerr(X, np.fft.rfft2(x)) < 1e-13
    # c2r unnormalised, DC/Nyquist imaginary parts ignored (FFTW convention)
    Y = X + 1j * rng.standard_normal(X.shape) * 0.0
    y = k.c2r2d(Y)
    assert relerr(y, x * (2 * Nu * Nphi)) < 1e-13
    Z = rng.standard_normal(X.shape) + 1j * rng.standard_normal(X.shape)
    z = k.c2r2d(Z)
    zn = np.fft.irfft2(Z, s=(2 * Nu, Nphi), norm="forward")
    assert relerr(z, zn) < 1e-13


# ------------------------------------------------------------------ Kalnajs
def _ring(R, phi):
    S0 = 0.372 * np.exp(-(R - 0.45) ** 2 / 0.05)
    return S0[:, None] * (1.0 + 0.3 * np.cos(2.0 * phi)[None, :])


def test_kalnajs_known_answer_32x32(orc):
    """test_kalnajs.cpp:10-42: FFT potential vs direct sum, rel < 1e-10."""
    Nu = Nphi = 32
    k = orc.Kalnajs(Nu, Nphi, 0.25, 1.0, 0.02)
    R, phi, _ = grid(Nu, Nphi)
    Sig = _ring(R, phi)
    pf = k.potential(Sig)
    pd = k.potential_direct(Sig)
    rel = np.max(np.abs(pf - pd)) / np.max(np.abs(pd))
    assert rel < 1e-10
    assert rel < 1e-13  # reference README records 2.4e-15


@pytest.mark.parametrize("Nu,Nphi,eps", [(16, 24, 0.02), (12, 20, 0.005), (20, 30, 0.04),
                                         (9, 14, 0.01), (24, 16, 0.0025)])
def test_kalnajs_fft_vs_direct_random(orc, Nu, Nphi, eps):
    k = orc.Kalnajs(Nu, Nphi, 0.1, 1.0, eps)
    rng = np.random.default_rng(7)
    Sig = rng.random((Nu, Nphi))
    pf = k.potential(Sig)
    pd = k.potential_direct(Sig)
    assert relerr(pf, pd) < 1e-12


def test_direct_sum_brute_force_tiny(orc):
    """The direct sum itself against an independent numpy evaluation of the
    Kalnajs convolution integral (kalnajs.hpp:4-9)."""
    Nu, Nphi, eps = 5, 6, 0.03
    k = orc.Kalnajs(Nu, Nphi, 0.25, 1.0, eps, 1.0)
    R, phi, du = grid(Nu, Nphi)
    dphi = 2 * np.pi / Nphi
    rng = np.random.default_rng(2)
    Sig = rng.random((Nu, Nphi))
    u = np.log(R)
    out = np.zeros((Nu, Nphi))
    for i in range(Nu):
        for j in range(Nphi):
            Du = u[i] - u[:, None]
            Dp = phi[j] - phi[None, :]
            Kk = 1 / np.sqrt(2 * np.cosh(Du) - 2 * np.cos(Dp) + eps)
            out[i, j] = -du * dphi * np.sum(R[:, None] ** 1.5 * Sig * Kk) / np.sqrt(R[i])
    assert relerr(k.potential_direct(Sig), out) < 1e-13
    assert relerr(k.potential(Sig), out) < 1e-13


def test_khat_is_real_even(orc):
    k = orc.Kalnajs(32, 48, 0.1, 1.0, 0.01)
    K = k.khat()
    assert np.max(np.abs(K.imag)) / np.max(np.abs(K.real)) < 1e-13
    M = 64
    assert relerr(K[1:M // 2], K[M - 1:M // 2:-1]) < 1e-13


def test_forces_exact_on_linear_and_cosine(orc):
    Nu, Nphi = 10, 16
    k = orc.Kalnajs(Nu, Nphi, 0.25, 1.0, 0.02)
    R, phi, du = grid(Nu, Nphi)
    u = np.log(R)
    Phi = (0.3 + 1.7 * u)[:, None] * np.ones(Nphi)[None, :]
    gR, gp = k.forces(Phi)
    assert relerr(gR, (-1.7 / R)[:, None] * np.ones((1, Nphi))) < 1e-12
    assert np.max(np.abs(gp)) == 0.0
    m = 3
    dphi = 2 * np.pi / Nphi
    Phi = np.ones(Nu)[:, None] * np.cos(m * phi)[None, :]
    gR, gp = k.forces(Phi)
    exact = np.sin(m * phi) * np.sin(m * dphi) / dphi     # -(centred diff of cos)
    assert relerr(gp, exact[None, :] / R[:, None]) < 1e-12
    assert np.max(np.abs(gR)) < 1e-15


# ------------------------------------------------------------------ hydro vs reference
needs_ref = pytest.mark.skipif(
    not os.path.exists(os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_hydro.so")),
    reason="oracle/_ref not built (reference checkout absent)")


@needs_ref
@pytest.mark.parametrize("Nu,Nphi", [(16, 24), (33, 20), (64, 64)])
def test_hydro_bitexact_vs_reference(orc, Nu, Nphi):
    Sig, vR, vp = rough_state(Nu, Nphi, seed=Nu + Nphi)
    rng = np.random.default_rng(5)
    gR = -0.5 + 0.1 * rng.standard_normal((Nu, Nphi))
    gp = 0.1 * rng.standard_normal((Nu, Nphi))
    h = orc.Hydro(Nu, Nphi)
    r = orc.RefHydro(Nu, Nphi)
    R, Rf, V = r.geom()
    assert np.array_equal(R, h.R) and np.array_equal(Rf, h.Rf) and np.array_equal(V, h.Vol)
    h.set_state(Sig, vR, vp)
    r.set_state(Sig, vR, vp)
    assert h.timestep(0.3) == r.timestep(0.3)
    dt = h.timestep(0.3)
    # stages one at a time
    for stage, fn in ((1, lambda: h.source(dt, gR, gp)), (2, lambda: h.transportR(dt)),
                      (3, lambda: h.transportPhi(dt))):
        fn()
        r.stage(stage, dt, gR, gp) if stage == 1 else r.stage(stage, dt)
        for a, b in zip(h.state(), r.state()):
            assert np.array_equal(a, b), f"stage {stage}"
    # full steps with and without gravity
    for it in range(5):
        dt = h.timestep(0.3)
        assert dt == r.timestep(0.3)
        g = (gR, gp) if it % 2 == 0 else (None, None)
        h.step(dt, *g)
        r.stage(0, dt, *g)
        for a, b in zip(h.state(), r.state()):
            assert np.array_equal(a, b)


@needs_ref
def test_hydro_nan_semantics_vs_reference(orc):
    """NaN zones: std::max keeps NaN, the CFL max ignores it (hydro.hpp:54-57,69)."""
    Nu, Nphi = 8, 8
    Sig, vR, vp = rough_state(Nu, Nphi, seed=3)
    Sig[3, 4] = np.nan
    h = orc.Hydro(Nu, Nphi)
    r = orc.RefHydro(Nu, Nphi)
    h.set_state(Sig, vR, vp)
    r.set_state(Sig, vR, vp)
    assert h.timestep(0.3) == r.timestep(0.3)
    h.step(1e-3)
    r.stage(0, 1e-3)
    for a, b in zip(h.state(), r.state()):
        assert np.array_equal(a, b, equal_nan=True)


def test_hydro_golden_vectors(orc):
    """Vectors produced by the reference hydro.hpp (tests/golden/make_golden.py)."""
    path = os.path.join(HERE, "golden", "hydro_ref.npz")
    if not os.path.exists(path):
        pytest.skip("golden vectors not generated")
    g = np.load(path)
    Nu, Nphi = int(g["Nu"]), int(g["Nphi"])
    h = orc.Hydro(Nu, Nphi)
    h.set_state(g["Sig0"], g["vR0"], g["vp0"])
    for s in range(int(g["nsteps"])):
        dt = h.timestep(0.3)
        assert dt == g["dt"][s]
        h.step(dt, g["gR"], g["gp"])
    for a, b in zip(h.state(), (g["Sig"], g["vR"], g["vp"])):
        assert np.array_equal(a, b)


# ------------------------------------------------------------------ closed forms
def test_mass_conservation(orc):
    Nu, Nphi = 32, 48
    Sig, vR, vp = rough_state(Nu, Nphi, seed=11)
    h = orc.Hydro(Nu, Nphi)
    h.set_state(Sig, vR, vp)
    m0 = np.sum(h.Sig * h.Vol[:, None])
    for _ in range(20):
        h.step(h.timestep(0.3))
    m1 = np.sum(h.Sig * h.Vol[:, None])
    assert abs(m1 - m0) / m0 < 1e-13


def test_uniform_state_invariant_under_transport(orc):
    Nu, Nphi = 12, 16
    h = orc.Hydro(Nu, Nphi)
    h.set_state(np.full((Nu, Nphi), 0.7), np.zeros((Nu, Nphi)), np.zeros((Nu, Nphi)))
    h.transportR(1e-2)
    h.transportPhi(1e-2)
    assert np.all(h.Sig == 0.7) and np.all(h.vR == 0) and np.all(h.vp == 0)


def test_source_kick_closed_form(orc):
    """Uniform Sigma, v = 0: dv_R = -dt GM/R^2 + dt g_R, dv_phi = dt g_phi."""
    Nu, Nphi, gm, dt = 12, 16, 0.6, 1e-3
    h = orc.Hydro(Nu, Nphi, gmstar=gm)
    R = h.R
    h.set_state(np.full((Nu, Nphi), 0.5), np.zeros((Nu, Nphi)), np.zeros((Nu, Nphi)))
    gR = np.full((Nu, Nphi), -0.25)
    gp = np.full((Nu, Nphi), 0.125)
    h.source(dt, gR, gp)
    assert relerr(h.vR, (dt * (-gm / R ** 2 - 0.25))[:, None] * np.ones((1, Nphi))) < 1e-14
    assert relerr(h.vp, np.full((Nu, Nphi), dt * 0.125)) < 1e-14


def test_courant_one_shift(orc):
    """v_phi = R dphi, dt = 1: the phi sweep shifts every row by one cell."""
    Nu, Nphi = 6, 16
    h = orc.Hydro(Nu, Nphi)
    R = h.R
    dphi = 2 * np.pi / Nphi
    rng = np.random.default_rng(4)
    Sig = 0.5 + rng.random((Nu, Nphi))
    vp = (R * dphi)[:, None] * np.ones((1, Nphi))
    h.set_state(Sig, np.zeros((Nu, Nphi)), vp)
    h.transportPhi(1.0)
    assert relerr(h.Sig, np.roll(Sig, 1, axis=1)) < 1e-14
    assert relerr(h.vp, vp) < 1e-14
    assert np.max(np.abs(h.vR)) == 0.0


def test_cm_closed_form(orc):
    Nu, Nphi, amp = 32, 32, 2e-3
    h = orc.Hydro(Nu, Nphi)
    orc.ic_lka98(h, None, 0, amp)
    mD = h.disk_mass()
    c0, _ = h.cm(0, mD)
    c2, ph2 = h.cm(2, mD)
    assert abs(c0 - 1.0) < 1e-14
    assert abs(c2 - amp / 2) / (amp / 2) < 1e-12
    assert abs(ph2) < 1e-10


def test_timestep_closed_form_uniform(orc):
    Nu, Nphi, K, cfl = 8, 8, 0.25, 0.3
    h = orc.Hydro(Nu, Nphi, K=K)
    s, u, w = 0.4, 0.1, 0.9
    h.set_state(np.full((Nu, Nphi), s), np.full((Nu, Nphi), -u), np.full((Nu, Nphi), w))
    R, Rf = h.R, h.Rf
    cs = np.sqrt(2 * K * s)
    rate = (u + cs) / (Rf[1:] - Rf[:-1]) + (w + cs) / (R * 2 * np.pi / Nphi)
    assert abs(h.timestep(cfl) - cfl / rate.max()) / (cfl / rate.max()) < 1e-14


# ------------------------------------------------------------------ loop semantics
def test_run_loop_matches_manual_loop(orc):
    """diskmain.cpp:92-110 output schedule and last-step clip, traced by hand."""
    Nu, Nphi, tlim = 16, 16, 0.07
    h1 = orc.Hydro(Nu, Nphi)
    k1 = orc.Kalnajs(Nu, Nphi, 0.25, 1.0, 0.02)
    orc.ic_lka98(h1, k1, 1, 1e-3)
    h2 = orc.Hydro(Nu, Nphi)
    h2.set_state(*h1.state())
    mD = h1.disk_mass()
    samples, tf, c2f, steps = orc.run(h1, k1, 1, tlim, dtout=0.02, cfl=0.3, mD=mD)
    t, tnext, n, out = 0.0, 0.0, 0, []
    while t < tlim:
        if t >= tnext - 1e-12:
            out.append(t)
            tnext += 0.02
        dt = h2.timestep(0.3)
        if t + dt > tlim:
            dt = tlim - t
        gR, gp = k1.forces(k1.potential(h2.Sig))
        h2.step(dt, gR, gp)
        t += dt
        n += 1
    assert steps == n and tf == t
    assert np.array_equal(samples[:, 0], np.array(out))
    for a, b in zip(h1.state(), h2.state()):
        assert np.array_equal(a, b)
    assert c2f == h2.cm(2, mD)[0]
