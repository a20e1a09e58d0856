"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Bar: Sigma and velocities after 10 steps within 1e-10 relative
(max-norm over the field, relative to the field's max).  The kernels use the
reference's formulas with reciprocal tables instead of some divisions and FMA
contraction, so agreement is round-off level; the tests assert 1e-10 for the
loop and tighter bounds (stated per test) for single operators.
"""
import numpy as np
import pytest

from helpers import grid, relerr, rough_state

pytestmark = pytest.mark.gpu

LOOP_TOL = 1e-10      # loop tolerance, Sigma/vR/vphi after 10 steps
OP_TOL = 1e-12        # single operator (one transform, one stage, one step)


def state_errs(got, ref):
    """Relative errors of (Sigma, vR, vphi): Sigma against its own maximum,
    each velocity component against the speed scale max(|vR|, |vphi|) of the
    reference state (vR starts at zero in these disks: after ten steps of a
    fine grid its own maximum is ~1e-7 of the speed, and a per-component
    scale would measure round-off of the speed against it)."""
    vscale = max(np.max(np.abs(ref[1])), np.max(np.abs(ref[2])))
    return [relerr(got[0], ref[0]),
            float(np.max(np.abs(got[1] - ref[1])) / vscale),
            float(np.max(np.abs(got[2] - ref[2])) / vscale)]


@pytest.fixture(scope="module")
def dfk():
    import diskfft_amd
    if diskfft_amd.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return diskfft_amd


def _pair(orc, dfk, Nu, Nphi, eps=0.02, Rin=0.25, Rout=1.0, selfgrav=1, **kw):
    p = dfk.disk_params(Nu, Nphi, Rin=Rin, Rout=Rout, eps=eps, selfgrav=selfgrav, **kw)
    e = dfk.Ensemble(p)
    h = orc.Hydro(Nu, Nphi, Rin, Rout, p["K"], p["gamma"], p["gmstar"])
    k = orc.Kalnajs(Nu, Nphi, Rin, Rout, eps, p["G"]) if selfgrav else None
    return e, h, k


# ------------------------------------------------------------------ set-up tables
@pytest.mark.parametrize("Nu,Nphi", [(32, 32), (24, 20), (12, 15), (64, 96)])
def test_geometry_and_khat(orc, dfk, Nu, Nphi):
    print(f"case geometry/khat {Nu}x{Nphi}", flush=True)
    e, h, k = _pair(orc, dfk, Nu, Nphi, eps=0.01)
    R, Rf, V = e.geometry()
    assert np.array_equal(R, h.R) and np.array_equal(Rf, h.Rf) and np.array_equal(V, h.Vol)
    Kg = e.khat()
    Ko = k.khat()
    assert relerr(Kg, Ko) < 1e-13
    assert np.max(np.abs(Kg.imag)) / np.max(np.abs(Kg.real)) < 1e-13


# ------------------------------------------------------------------ gravity
def test_potential_known_answer_32x32(orc, dfk):
    """test_kalnajs.cpp:10-42: FFT potential vs direct convolution < 1e-10."""
    print("case potential 32x32 known answer", flush=True)
    Nu = Nphi = 32
    e, h, k = _pair(orc, dfk, Nu, Nphi, eps=0.02)
    R, phi, _ = grid(Nu, Nphi)
    Sig = 0.372 * np.exp(-(R - 0.45) ** 2 / 0.05)[:, None] * (1 + 0.3 * np.cos(2 * phi))[None, :]
    pg = e.potential(Sig)
    pd = k.potential_direct(Sig)
    assert np.max(np.abs(pg - pd)) / np.max(np.abs(pd)) < 1e-10
    assert relerr(pg, k.potential(Sig)) < 1e-13


@pytest.mark.parametrize("Nu,Nphi,eps", [(16, 24, 0.02), (12, 15, 0.005), (64, 64, 0.01),
                                         (48, 80, 0.04), (128, 192, 0.0025), (256, 256, 0.01),
                                         (96, 28, 0.02)])
def test_potential_matches_oracle(orc, dfk, Nu, Nphi, eps):
    print(f"case potential {Nu}x{Nphi} eps={eps}", flush=True)
    e, h, k = _pair(orc, dfk, Nu, Nphi, eps=eps, Rin=0.1)
    rng = np.random.default_rng(Nu + Nphi)
    Sig = rng.random((Nu, Nphi))
    pg = e.potential(Sig)
    assert relerr(pg, k.potential(Sig)) < 1e-13
    if Nu * Nphi <= 4096:
        assert relerr(pg, k.potential_direct(Sig)) < 1e-12


@pytest.mark.parametrize("Nu,Nphi,pr,cb", [(256, 256, 8, 8), (512, 1024, 4, 2), (96, 80, 16, 16)])
def test_potential_large_batches(orc, dfk, Nu, Nphi, pr, cb):
    """Row / column batches whose LDS use exceeds 64 KiB (gfx950 allows 160 KiB
    per workgroup): pr = 8 rows of 256 use 68 KiB, cb = 8 columns of 512 use
    136 KiB; 512 x 1024 uses 144 KiB (rows) and 80 KiB (columns)."""
    print(f"case potential {Nu}x{Nphi} pr={pr} cb={cb}", flush=True)
    e, h, k = _pair(orc, dfk, Nu, Nphi, eps=0.01, Rin=0.1)
    e.set_shapes(0, 0, pr, cb)
    assert tuple(e.get_shapes(0))[2:] == (min(pr, (Nu + 1) // 2), min(cb, Nphi // 2 + 1))
    rng = np.random.default_rng(7)
    Sig = rng.random((Nu, Nphi))
    assert relerr(e.potential(Sig), k.potential(Sig)) < 1e-13


@pytest.mark.parametrize("Nu,Nphi,wave_ok", [(256, 256, 1), (35, 40, 0), (96, 80, 1), (33, 26, 0),
                                             (512, 1024, 1), (384, 512, 1), (3, 4, 1)])
@pytest.mark.parametrize("opts,cb,threads", [((0, 0, 0), 0, 0), ((1, 1, 1), 3, 128),
                                             ((1, 0, 1), 1, 64), ((0, 1, 1), 8, 256),
                                             ((1, 1, 2), 2, 128), ((1, 0, 2), 1, 256),
                                             ((0, 0, 2), 3, 64), ((1, 0, 2), 1, 192)])
def test_potential_fft_options(orc, dfk, Nu, Nphi, wave_ok, opts, cb, threads):
    """Transform options (dfk_set_fft_options): twiddles from HBM instead of
    LDS give the same bits; the in-place column transforms (kernels of their
    own, which may contract multiply-adds differently) agree to round-off:
    one wave per column with fewer, as many and more columns than waves per
    workgroup, and the whole workgroup on one column after another at every
    workgroup size; on the HBM-I/O stage path (radices 2/3/4/8) and the
    general one (35 x 40: 70 = 2 x 5 x 7; 33 x 26: 66 = 2 x 3 x 11), where
    the in-place modes are off."""
    print(f"case potential {Nu}x{Nphi} fft options {opts} cb {cb} threads {threads}", flush=True)
    e, h, k = _pair(orc, dfk, Nu, Nphi, eps=0.01, Rin=0.1)
    rng = np.random.default_rng(11)
    Sig = rng.random((Nu, Nphi))
    p_ref = e.potential(Sig)
    e.set_fft_options(*opts)
    if cb:
        e.set_shapes(0, 0, 0, cb)
        e.set_threads(256, threads, 256)
    assert e.get_fft_options(0) == (opts[0], opts[1], opts[2] * wave_ok)
    p_opt = e.potential(Sig)
    if opts[2] * wave_ok:
        assert relerr(p_opt, p_ref) < 1e-14
    else:
        assert np.array_equal(p_ref, p_opt)
    assert relerr(p_opt, k.potential(Sig)) < 1e-13


@pytest.mark.parametrize("Nu,Nphi,ok", [(64, 64, 1), (96, 80, 1), (128, 128, 1), (192, 192, 1),
                                        (256, 256, 1), (384, 512, 1), (512, 512, 1),
                                        (512, 1024, 1), (64, 30, 1), (35, 40, 0), (100, 64, 0)])
def test_potential_register_columns(orc, dfk, Nu, Nphi, ok):
    """The two-pass register column convolution (cols_wave = 3) at every
    transform length it is built for (2Nu = 128 .. 1024): round-off agreement
    with the cooperative transform and the oracle; other lengths keep the
    cooperative one (read back as 0), the same bits."""
    print(f"case potential {Nu}x{Nphi} register columns", flush=True)
    e, h, k = _pair(orc, dfk, Nu, Nphi, eps=0.01, Rin=0.1)
    rng = np.random.default_rng(Nu + 3 * Nphi)
    Sig = rng.random((Nu, Nphi))
    p_ref = e.potential(Sig)
    e.set_fft_options(1, 0, 3)
    assert e.get_fft_options(0)[2] == 3 * ok
    p_reg = e.potential(Sig)
    if ok:
        assert relerr(p_reg, p_ref) < 1e-14
    else:
        assert np.array_equal(p_reg, p_ref)
    assert relerr(p_reg, k.potential(Sig)) < 1e-13


@pytest.mark.parametrize("Nu,Nphi,ok", [(128, 128, 1), (192, 192, 1), (256, 256, 1), (256, 30, 1),
                                        (128, 1024, 1), (384, 512, 1), (512, 512, 1), (512, 34, 1),
                                        (64, 64, 0), (1024, 64, 0), (35, 40, 0)])
def test_potential_three_pass_columns(orc, dfk, Nu, Nphi, ok):
    """Column mode 4, the three-pass register convolution (one column per
    wavefront, 2Nu = 256, 384, 512, 768, 1024; two k1 groups per lane from
    768): round-off agreement with the
    cooperative transform and the oracle; other lengths keep the cooperative
    one (read back as 0), the same bits."""
    print(f"case potential {Nu}x{Nphi} three-pass columns", flush=True)
    e, h, k = _pair(orc, dfk, Nu, Nphi, eps=0.01, Rin=0.1)
    rng = np.random.default_rng(Nu + 5 * Nphi)
    Sig = rng.random((Nu, Nphi))
    p_ref = e.potential(Sig)
    e.set_fft_options(1, 0, 4)
    assert e.get_fft_options(0)[2] == 4 * ok
    p_3 = e.potential(Sig)
    if ok:
        assert relerr(p_3, p_ref) < 1e-14
    else:
        assert np.array_equal(p_3, p_ref)
    assert relerr(p_3, k.potential(Sig)) < 1e-13


@pytest.mark.parametrize("Nu,Nphi,mode", [(64, 128, 2), (37, 192, 2), (256, 256, 2), (33, 384, 2),
                                          (512, 512, 2), (96, 768, 2), (512, 1024, 2), (3, 128, 2),
                                          (5, 256, 2), (40, 96, 1), (35, 40, 0)])
def test_potential_register_rows(orc, dfk, Nu, Nphi, mode):
    """The two-pass register row transforms (dfk_set_row_transform 2) at every
    row length they are built for (Nphi = 128 .. 1024), with odd Nu (the last
    pair's second row past the disk) and disks of fewer rows than a
    workgroup's pairs: round-off agreement with the cooperative transforms and
    the oracle, alone and with the register column convolution; other Nphi
    fall back to the in-place rows, or to the cooperative ones where those do
    not apply (same bits as asking for them)."""
    print(f"case potential {Nu}x{Nphi} register rows", flush=True)
    e, h, k = _pair(orc, dfk, Nu, Nphi, eps=0.01, Rin=0.1)
    rng = np.random.default_rng(Nu + 5 * Nphi)
    Sig = rng.random((Nu, Nphi))
    p_ref = e.potential(Sig)
    e.set_row_transform(mode)
    p_fb = e.potential(Sig)
    e.set_row_transform(2)
    assert e.get_row_transform() == mode
    p_reg = e.potential(Sig)
    assert np.array_equal(p_reg, p_fb)
    if mode == 2:
        assert relerr(p_reg, p_ref) < 1e-14
    assert relerr(p_reg, k.potential(Sig)) < 1e-13
    e.set_fft_options(1, 0, 3)
    p_both = e.potential(Sig)
    assert relerr(p_both, k.potential(Sig)) < 1e-13


@pytest.mark.parametrize("Nu,Nphi,mode", [(64, 128, 3), (37, 192, 3), (256, 256, 3), (3, 128, 3),
                                          (5, 256, 3), (128, 384, 3), (33, 512, 3), (64, 1024, 2),
                                          (40, 96, 1)])
def test_potential_three_pass_rows(orc, dfk, Nu, Nphi, mode):
    """The three-pass register row transforms (dfk_set_row_transform 3, one
    row pair per wavefront) at Nphi = 128, 192, 256, 384, 512, with odd Nu and
    disks of one or two pairs: round-off agreement with the cooperative transforms and
    the oracle, alone and with the three-pass columns; other Nphi take the
    next variant their rows allow (same bits as asking for it)."""
    print(f"case potential {Nu}x{Nphi} three-pass rows", flush=True)
    e, h, k = _pair(orc, dfk, Nu, Nphi, eps=0.01, Rin=0.1)
    rng = np.random.default_rng(Nu + 7 * Nphi)
    Sig = rng.random((Nu, Nphi))
    p_ref = e.potential(Sig)
    e.set_row_transform(mode)
    p_fb = e.potential(Sig)
    e.set_row_transform(3)
    assert e.get_row_transform() == mode
    p_3 = e.potential(Sig)
    assert np.array_equal(p_3, p_fb)
    if mode == 3:
        assert relerr(p_3, p_ref) < 1e-14
    assert relerr(p_3, k.potential(Sig)) < 1e-13
    e.set_fft_options(1, 0, 4)
    assert relerr(e.potential(Sig), k.potential(Sig)) < 1e-13


@pytest.mark.parametrize("Nu,Nphi", [(16, 34), (20, 38), (17, 46), (12, 1031), (1024, 1024),
                                     (64, 4096), (2048, 64), (2048, 4096)])
def test_potential_any_size(orc, dfk, Nu, Nphi):
    """Transform lengths with prime factors above 13 (34 = 2 x 17, 38 = 2 x
    19, 2Nu = 34, 46 = 2 x 23, 1031 prime: the direct-DFT stage) and the
    largest grids (2Nu up to 4096, Nphi up to 4096, twiddles read through the
    caches) against the oracle, whose FFT takes any length."""
    print(f"case potential {Nu}x{Nphi} (any size)", flush=True)
    e, h, k = _pair(orc, dfk, Nu, Nphi, eps=0.01, Rin=0.1)
    rng = np.random.default_rng(Nu * 7 + Nphi)
    Sig = rng.random((Nu, Nphi))
    assert relerr(e.potential(Sig), k.potential(Sig)) < 1e-13
    Kg = e.khat()
    assert np.max(np.abs(Kg.imag)) / np.max(np.abs(Kg.real)) < 1e-13


@pytest.mark.parametrize("Nu,Nphi", [(16, 16), (35, 40), (128, 128), (33, 26)])
def test_forces_match_oracle(orc, dfk, Nu, Nphi):
    print(f"case forces {Nu}x{Nphi}", flush=True)
    e, h, k = _pair(orc, dfk, Nu, Nphi)
    rng = np.random.default_rng(1)
    Phi = rng.standard_normal((Nu, Nphi))
    gRg, gpg = e.forces(Phi)
    gRo, gpo = k.forces(Phi)
    assert relerr(gRg, gRo) < 1e-14 and relerr(gpg, gpo) < 1e-14


# ------------------------------------------------------------------ hydro operators
@pytest.mark.parametrize("Nu,Nphi", [(16, 24), (37, 52), (128, 128), (256, 256), (512, 1024)])
@pytest.mark.parametrize("seed", [5, 6, 7])
def test_timestep_matches_oracle(orc, dfk, Nu, Nphi, seed):
    """DiskHydro::timestep (hydro.hpp:48-61) is the same double: the rate of
    the maximum is formed with the reference's sqrt and IEEE divisions and the
    max is order-free.  Also for states with zones below the 1e-12 clamp."""
    print(f"case timestep {Nu}x{Nphi} seed {seed}", flush=True)
    e, h, _ = _pair(orc, dfk, Nu, Nphi, selfgrav=0)
    st = rough_state(Nu, Nphi, seed=seed)
    if seed == 7:
        st[0][1::3, ::5] = 1e-14
    e.set_state(*st)
    h.set_state(*st)
    for cfl in (0.3, 0.4):
        dg, do = e.timestep(cfl), h.timestep(cfl)
        assert dg == do, (dg, do)


def test_timestep_general_eos(orc, dfk):
    """gamma = 1.5 takes pow (hydro.hpp:44): the device pow and the host's may
    differ in the last bit, so dt agrees to round-off there."""
    print("case timestep general EOS", flush=True)
    Nu, Nphi = 64, 96
    e, h, _ = _pair(orc, dfk, Nu, Nphi, selfgrav=0, K=0.2, gamma=1.5)
    st = rough_state(Nu, Nphi, seed=8)
    e.set_state(*st)
    h.set_state(*st)
    dg, do = e.timestep(0.3), h.timestep(0.3)
    assert abs(dg - do) / do < 1e-15


@pytest.mark.parametrize("stage", [1, 2, 3, 4])
@pytest.mark.parametrize("Nu,Nphi", [(16, 24), (37, 52)])
def test_single_stage_matches_oracle(orc, dfk, stage, Nu, Nphi):
    print(f"case stage {stage} {Nu}x{Nphi}", flush=True)
    e, h, _ = _pair(orc, dfk, Nu, Nphi, selfgrav=0)
    st = rough_state(Nu, Nphi, seed=stage)
    if stage == 4:
        st[0][3, :5] = -1.0   # below the floor
    e.set_state(*st)
    h.set_state(*st)
    rng = np.random.default_rng(3)
    gR = -0.3 + 0.05 * rng.standard_normal((Nu, Nphi))
    gp = 0.05 * rng.standard_normal((Nu, Nphi))
    dt = 0.5 * h.timestep(0.3)
    e.stage(stage, dt, gR, gp)
    {1: lambda: h.source(dt, gR, gp), 2: lambda: h.transportR(dt),
     3: lambda: h.transportPhi(dt), 4: lambda: h.floor()}[stage]()
    for a, b in zip(e.get_state(), h.state()):
        assert relerr(a, b) < OP_TOL


@pytest.mark.parametrize("kind,shape,Nu,Nphi", [
    (0, (8, 64, 0, 0), 37, 52), (0, (4, 16, 0, 0), 37, 52), (0, (3, 7, 0, 0), 37, 52),
    (0, (16, 32, 0, 0), 37, 52), (0, (1, 5, 0, 0), 37, 52),
    (1, (8, 0, 0, 0), 37, 52), (1, (5, 0, 0, 0), 37, 52), (1, (64, 0, 0, 0), 37, 52),
    (1, (16, 0, 0, 0), 40, 130), (1, (3, 0, 0, 0), 21, 7), (1, (32, 0, 0, 0), 70, 256)])
def test_fused_step_matches_oracle(orc, dfk, kind, shape, Nu, Nphi):
    """DiskHydro::step with given g_R/g_phi through the fused update kernels
    (0 LDS tile, 1 register streaming), for tile / strip shapes that do and do
    not divide the grid, and column counts below, at and above one wave."""
    print(f"case fused step kind={kind} tiles={shape} grid {Nu}x{Nphi}", flush=True)
    e, h, _ = _pair(orc, dfk, Nu, Nphi, selfgrav=0)
    e.set_update_kernel(kind)
    e.set_shapes(*shape)
    st = rough_state(Nu, Nphi, seed=9)
    e.set_state(*st)
    h.set_state(*st)
    rng = np.random.default_rng(4)
    gR = -0.3 + 0.05 * rng.standard_normal((Nu, Nphi))
    gp = 0.05 * rng.standard_normal((Nu, Nphi))
    for it in range(3):
        dt = h.timestep(0.3)
        g = (gR, gp) if it != 1 else (None, None)
        e.step(dt, *g)
        h.step(dt, *g)
    for a, b in zip(e.get_state(), h.state()):
        assert relerr(a, b) < OP_TOL


# ------------------------------------------------------------------ the loop
def _lka98(orc, Nu, Nphi, eps, amp=1e-3, selfgrav=1):
    h = orc.Hydro(Nu, Nphi)
    k = orc.Kalnajs(Nu, Nphi, 0.25, 1.0, eps)
    orc.ic_lka98(h, k, selfgrav, amp)
    return h, k


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("Nu,Nphi,eps", [(64, 64, 0.02), (128, 128, 0.02), (96, 80, 0.01),
                                         (256, 256, 0.005), (48, 36, 0.02), (512, 1024, 0.01),
                                         (3, 4, 0.02), (64, 34, 0.02), (1024, 1024, 0.01),
                                         (2048, 128, 0.01), (2048, 4096, 0.01)])
def test_ten_step_loop_parity(orc, dfk, Nu, Nphi, eps, kind):
    """The loop gate: Sigma, vR, vphi after 10 steps <= 1e-10 relative, from
    the smallest grid to the largest the library takes (2048 x 4096)."""
    print(f"case 10-step loop {Nu}x{Nphi} eps={eps} update kernel {kind}", flush=True)
    h, k = _lka98(orc, Nu, Nphi, eps)
    e = dfk.Ensemble(dfk.disk_params(Nu, Nphi, eps=eps))
    e.set_update_kernel(kind)
    e.set_state(*h.state())
    mD = h.disk_mass()
    so, tf, c2f, steps = orc.run(h, k, 1, 1e9, dtout=0.25, cfl=0.3, max_steps=10, mD=mD)
    e.run_init(1e9, dtout=0.25, cfl=0.3, mD=mD, max_steps=10)
    e.advance(12)
    st = e.status()
    assert st["done"] and st["steps"] == 10 and steps == 10
    assert abs(st["t"] - tf) / tf < 1e-13
    errs = state_errs(e.get_state(), h.state())
    print("  rel errors Sigma, vR, vphi (velocities vs the speed scale):", errs, flush=True)
    assert max(errs) <= LOOP_TOL
    sg = e.samples()
    assert sg.shape == so.shape
    assert abs(sg[0, 2] - so[0, 2]) / so[0, 2] < 1e-12


@pytest.mark.parametrize("kind", [0, 1])
def test_general_eos_loop_and_mixed_ensemble(orc, dfk, kind):
    """gamma != 2 takes the EOS path with pow (hydro.hpp:44-45): 10-step loop
    parity with the oracle; a handle mixing gamma = 2 and 1.5 disks runs that
    path for both and still gives each disk the bits of its own run."""
    print(f"case general EOS, update kernel {kind}", flush=True)
    Nu, Nphi, eps = 64, 64, 0.02
    runs = []
    for gamma, K in [(1.5, 0.2), (2.0, 0.25)]:
        h = orc.Hydro(Nu, Nphi, K=K, gamma=gamma)
        k = orc.Kalnajs(Nu, Nphi, 0.25, 1.0, eps)
        orc.ic_lka98(h, k, 1, 1e-3)
        p = dfk.disk_params(Nu, Nphi, eps=eps, K=K, gamma=gamma)
        runs.append((h, k, p, h.state(), h.disk_mass()))
    h, k, p, ic, mD = runs[0]
    e = dfk.Ensemble(p)
    e.set_update_kernel(kind)
    e.set_state(*ic)
    orc.run(h, k, 1, 1e9, dtout=0.25, cfl=0.3, max_steps=10, mD=mD)
    e.run_init(1e9, dtout=0.25, cfl=0.3, mD=mD, max_steps=10)
    e.advance(12)
    errs = [relerr(a, b) for a, b in zip(e.get_state(), h.state())]
    print("  gamma 1.5 rel errors Sigma, vR, vphi:", errs, flush=True)
    assert e.status()["steps"] == 10 and max(errs) <= LOOP_TOL
    mixed = dfk.Ensemble([r[2] for r in runs])
    mixed.set_update_kernel(kind)
    for d, r in enumerate(runs):
        mixed.set_state(*r[3], d=d)
    mixed.run_init(1e9, mD=[r[4] for r in runs], max_steps=15)
    mixed.advance(16)
    for d, r in enumerate(runs):
        one = dfk.Ensemble(r[2])
        one.set_update_kernel(kind)
        one.set_shapes(*mixed.get_shapes(d))
        one.set_state(*r[3])
        one.run_init(1e9, mD=r[4], max_steps=15)
        one.advance(16)
        for a, b in zip(mixed.get_state(d), one.get_state()):
            assert np.array_equal(a, b)


def _powerlaw_ic(orc, dfk, Nu, Nphi, eps=0.01):
    """The benchmark's power-law disk (r in [0.1, 1]), its equilibrium from the
    oracle's gravity, so both sides start from the same arrays."""
    m = dfk.model_defaults(dfk.MODEL_POWERLAW)
    k = orc.Kalnajs(Nu, Nphi, m.Rin, m.Rout, eps, m.G)
    p, S, U, W, mD, _ = dfk.model_ic(m, Nu, Nphi, eps,
                                     potential_forces=lambda x: k.forces(k.potential(x)))
    h = orc.Hydro(Nu, Nphi, p["Rin"], p["Rout"], p["K"], p["gamma"], p["gmstar"])
    h.set_state(S, U, W)
    return p, (S, U, W), mD, h, k


@pytest.mark.parametrize("kind", [0, 1, "tuned"])
@pytest.mark.parametrize("Nu,Nphi", [(256, 256), (512, 1024)])
def test_powerlaw_ten_step_loop_parity(orc, dfk, Nu, Nphi, kind):
    """The loop gate on the benchmark's own disk: power-law model, eps 0.01,
    CFL 0.4, 256 x 256 (the single-disk benchmark) and 512 x 1024 (the
    bandwidth case), with each update kernel and with the tuned shapes the
    benchmark runs.  Sigma, vR, vphi after 10 steps within 1e-10; the first
    step's dt (same state on both sides) is the same double, so t after one
    step is equal."""
    import autotune
    print(f"case power-law 10-step loop {Nu}x{Nphi} kernel {kind}", flush=True)
    p, ic, mD, h, k = _powerlaw_ic(orc, dfk, Nu, Nphi)
    e = dfk.Ensemble(p)
    if kind == "tuned":
        assert autotune.apply(e, regime="latency") or autotune.apply(e, regime="throughput")
    else:
        e.set_update_kernel(kind)
    e.set_state(*ic)
    e.run_init(1e30, dtout=0.25, cfl=0.4, mD=mD, max_steps=1)
    e.advance(2)
    h1 = orc.Hydro(Nu, Nphi, p["Rin"], p["Rout"], p["K"], p["gamma"], p["gmstar"])
    h1.set_state(*ic)
    _, t1, _, _ = orc.run(h1, k, 1, 1e30, dtout=0.25, cfl=0.4, max_steps=1, mD=mD)
    assert e.status()["t"] == t1
    e.set_state(*ic)
    so, tf, c2f, steps = orc.run(h, k, 1, 1e30, dtout=0.25, cfl=0.4, max_steps=10, mD=mD)
    e.run_init(1e30, dtout=0.25, cfl=0.4, mD=mD, max_steps=10)
    e.advance(12)
    st = e.status()
    assert st["done"] and st["steps"] == 10 and steps == 10
    assert abs(st["t"] - tf) / tf < 1e-13
    errs = [relerr(a, b) for a, b in zip(e.get_state(), h.state())]
    print("  rel errors Sigma, vR, vphi:", errs, flush=True)
    assert max(errs) <= LOOP_TOL
    sg = e.samples()
    # |C2| of the seeded noise is a cancelling sum (~6e-6 of |C0|): compared
    # relative to |C0|
    assert sg.shape == so.shape and abs(sg[0, 2] - so[0, 2]) < 1e-12 * so[0, 1]


@pytest.mark.parametrize("Nu,Nphi", [(128, 128), (192, 192), (256, 256), (256, 96)])
def test_three_pass_columns_loop_parity(orc, dfk, Nu, Nphi):
    """The three-pass register column convolution (column mode 4) and row
    transforms (row mode 3, or the in-place rows for Nphi = 96, which has no
    register form) in the time loop of the power-law disk: 10 steps within
    1e-10 of the oracle, the state within 1e-12 of the same run with the
    cooperative transforms."""
    print(f"case power-law 10-step loop {Nu}x{Nphi} three-pass columns", flush=True)
    p, ic, mD, h, k = _powerlaw_ic(orc, dfk, Nu, Nphi)
    out = []
    for mode in (0, 4):
        e = dfk.Ensemble(p)
        e.set_fft_options(1, 0, mode)
        assert e.get_fft_options(0)[2] == mode
        if mode == 4:   # with the three-pass rows where Nphi has them (96: in place)
            e.set_row_transform(3)
            assert e.get_row_transform() == (3 if Nphi in (128, 192, 256) else 1)
        e.set_state(*ic)
        e.run_init(1e30, dtout=0.25, cfl=0.4, mD=mD, max_steps=10)
        e.advance(12)
        assert e.status()["steps"] == 10
        out.append(e.get_state())
        e.close()
    orc.run(h, k, 1, 1e30, dtout=0.25, cfl=0.4, max_steps=10, mD=mD)
    errs = [relerr(a, b) for a, b in zip(out[1], h.state())]
    print("  rel errors Sigma, vR, vphi:", errs, flush=True)
    assert max(errs) <= LOOP_TOL
    assert max(relerr(a, b) for a, b in zip(out[1], out[0])) < 1e-12


def test_loop_samples_and_clip(orc, dfk):
    """Output schedule and last-step clip of diskmain.cpp:92-110 over ~300 steps."""
    Nu, Nphi, eps, tlim = 64, 64, 0.02, 0.6
    print("case loop samples 64x64 tlim=0.6", flush=True)
    h, k = _lka98(orc, Nu, Nphi, eps)
    e = dfk.Ensemble(dfk.disk_params(Nu, Nphi, eps=eps))
    e.set_state(*h.state())
    mD = h.disk_mass()
    so, tf, c2f, steps = orc.run(h, k, 1, tlim, dtout=0.25, cfl=0.3, mD=mD)
    ((sg, st),) = e.run(tlim, dtout=0.25, cfl=0.3, mD=mD, chunk=100)
    assert st["steps"] == steps
    assert st["t"] == pytest.approx(tf, rel=1e-14)
    assert sg.shape == so.shape
    assert np.max(np.abs(sg[:, 0] - so[:, 0])) < 1e-12
    assert relerr(sg[:, 2], so[:, 2]) < 1e-9
    # arg C2 (the fourth column of diskfft_modes.dat), compared modulo 2 pi
    assert np.max(np.abs(np.angle(np.exp(1j * (sg[:, 3] - so[:, 3]))))) < 1e-9
    assert abs(st["c2_final"] - c2f) / c2f < 1e-9
    errs = [relerr(a, b) for a, b in zip(e.get_state(), h.state())]
    print(f"  {steps} steps, rel errors:", errs, flush=True)
    assert max(errs) < 1e-9


def test_mass_conservation_gpu(dfk):
    Nu, Nphi = 64, 96
    print("case mass conservation", flush=True)
    e = dfk.Ensemble(dfk.disk_params(Nu, Nphi, selfgrav=1))
    e.set_state(*rough_state(Nu, Nphi, seed=21))
    m0 = e.disk_mass()
    e.run_init(1e9, cfl=0.3, mD=m0, max_steps=50)
    e.advance(50)
    assert abs(e.disk_mass() - m0) / m0 < 1e-13


@pytest.mark.parametrize("kind", [0, 1])
def test_ensemble_matches_single_disks(orc, dfk, kind):
    """Disks of different shapes in one handle give the same bits as each disk
    run alone (identical kernels per disk)."""
    print(f"case ensemble vs single, update kernel {kind}", flush=True)
    shapes = [(32, 48, 0.02), (64, 64, 0.01), (48, 30, 0.04), (40, 130, 0.01)]
    params = [dfk.disk_params(a, b, eps=c) for a, b, c in shapes]
    ens = dfk.Ensemble(params)
    ens.set_update_kernel(kind)
    ics, mDs = [], []
    for d, (a, b, c) in enumerate(shapes):
        h, _ = _lka98(orc, a, b, c)
        ics.append(h.state())
        mDs.append(h.disk_mass())
        ens.set_state(*ics[-1], d=d)
    caps = [20, 35, 27, 31]
    ens.run_init(1e9, mD=mDs, max_steps=caps)
    ens.advance(40)
    for d, p in enumerate(params):
        one = dfk.Ensemble(p)
        one.set_update_kernel(kind)
        one.set_shapes(*ens.get_shapes(d))
        one.set_state(*ics[d])
        one.run_init(1e9, mD=mDs[d], max_steps=caps[d])
        one.advance(40)
        assert ens.status(d)["steps"] == caps[d]
        for a, b in zip(ens.get_state(d), one.get_state()):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("kind", [0, 1])
def test_mixed_selfgrav_ensemble(orc, dfk, kind):
    """A handle mixing self-gravitating disks and disks without self-gravity
    (one graph chain; every disk's loop control rides in the row-transform
    launch) gives each disk the bits of its own run, and a disk without
    self-gravity matches the oracle's 10 steps."""
    print(f"case mixed selfgrav ensemble, update kernel {kind}", flush=True)
    shapes = [(32, 48, 0.02, 1), (64, 64, 0.01, 0), (48, 30, 0.04, 1), (40, 130, 0.01, 0)]
    params = [dfk.disk_params(a, b, eps=c, selfgrav=g) for a, b, c, g in shapes]
    hs = [_lka98(orc, a, b, c, selfgrav=g) for a, b, c, g in shapes]
    ics = [h.state() for h, _ in hs]
    mDs = [h.disk_mass() for h, _ in hs]
    caps = [20, 10, 27, 31]
    ens = dfk.Ensemble(params)
    ens.set_update_kernel(kind)
    ens.set_chains(3)
    assert ens.get_chains() == 1
    for d in range(len(shapes)):
        ens.set_state(*ics[d], d=d)
    ens.run_init(1e9, dtout=0.01, mD=mDs, max_steps=caps, sample_cap=64)
    ens.advance(40)
    for d, p in enumerate(params):
        one = dfk.Ensemble(p)
        one.set_update_kernel(kind)
        one.set_shapes(*ens.get_shapes(d))
        one.set_state(*ics[d])
        one.run_init(1e9, dtout=0.01, mD=mDs[d], max_steps=caps[d], sample_cap=64)
        one.advance(40)
        assert ens.status(d) == one.status()
        assert ens.status(d)["steps"] == caps[d]
        for a, b in zip(ens.get_state(d), one.get_state()):
            assert np.array_equal(a, b)
        assert np.array_equal(ens.samples(d), one.samples())
        one.close()
    h, k = hs[1]
    orc.run(h, k, 0, 1e9, dtout=0.01, cfl=0.3, max_steps=10, mD=mDs[1])
    e = dfk.Ensemble(params[1])
    e.set_state(*ics[1])
    e.run_init(1e9, dtout=0.01, mD=mDs[1], max_steps=10)
    e.advance(12)
    errs = [relerr(a, b) for a, b in zip(e.get_state(), h.state())]
    assert max(errs) <= LOOP_TOL, errs


@pytest.mark.parametrize("Nu,Nphi,ok", [(256, 256, 1), (96, 80, 0), (512, 1024, 1), (3, 4, 1),
                                        (64, 192, 1), (35, 40, 0)])
@pytest.mark.parametrize("threads,pr,cb", [(256, 8, 4), (128, 3, 1), (64, 1, 2), (192, 16, 8)])
def test_potential_row_transform(orc, dfk, Nu, Nphi, ok, threads, pr, cb):
    """In-place row transforms (dfk_set_row_transform) with the in-place
    workgroup column convolution, at every workgroup size and with batches
    clamped to the in-place limit: round-off agreement with the cooperative
    transforms and the oracle; Nphi with a factor 5 keeps the cooperative
    rows."""
    print(f"case potential {Nu}x{Nphi} rows in place threads {threads} pr {pr} cb {cb}", flush=True)
    e, h, k = _pair(orc, dfk, Nu, Nphi, eps=0.01, Rin=0.1)
    rng = np.random.default_rng(5)
    Sig = rng.random((Nu, Nphi))
    p_ref = e.potential(Sig)
    e.set_row_transform(1)
    e.set_fft_options(1, 0, 2)
    e.set_threads(threads, threads, 256)
    e.set_shapes(0, 0, pr, cb)
    assert e.get_row_transform() == ok
    _, _, pr_, cb_ = e.get_shapes(0)
    nmax = max(8 * threads, 1024)
    if ok:
        assert pr_ * Nphi <= nmax
    if e.get_fft_options(0)[2] == 2:
        assert cb_ * 2 * Nu <= nmax
    p_opt = e.potential(Sig)
    assert relerr(p_opt, p_ref) < 1e-14
    assert relerr(p_opt, k.potential(Sig)) < 1e-13


@pytest.mark.parametrize("rows,cols", [(1, 2), (1, 3), (2, 3), (2, 0)])
@pytest.mark.parametrize("Nu,Nphi,eps", [(64, 96, 0.02), (256, 256, 0.01), (61, 384, 0.01)])
def test_loop_10_steps_inplace_transforms(orc, dfk, Nu, Nphi, eps, rows, cols):
    """The 10-step loop gate with in-place (1) or register (2; Nphi = 96 has
    none and keeps in place) row transforms and the cooperative (0), in-place
    workgroup (2) or two-pass register (3) column convolution."""
    print(f"case loop {Nu}x{Nphi} rows mode {rows}, columns mode {cols}", flush=True)
    h, k = _lka98(orc, Nu, Nphi, eps)
    e = dfk.Ensemble(dfk.disk_params(Nu, Nphi, eps=eps))
    e.set_row_transform(rows)
    assert e.get_row_transform() == (1 if Nphi == 96 else rows)
    e.set_fft_options(1, 0, cols)
    reg_lengths = (128, 192, 256, 384, 512, 768, 1024)   # both in-place modes cover these
    assert e.get_fft_options(0)[2] == (cols if cols == 0 or 2 * Nu in reg_lengths else 0)
    e.set_threads(128, 128, 256)
    e.set_state(*h.state())
    mD = h.disk_mass()
    orc.run(h, k, 1, 1e9, dtout=0.25, cfl=0.3, max_steps=10, mD=mD)
    e.run_init(1e9, dtout=0.25, cfl=0.3, mD=mD, max_steps=10)
    e.advance(12)
    errs = [relerr(a, b) for a, b in zip(e.get_state(), h.state())]
    print("  rel errors Sigma, vR, vphi:", errs, flush=True)
    assert e.status()["steps"] == 10 and max(errs) <= LOOP_TOL


def test_three_pass_rows_ensemble(orc, dfk):
    """A handle of disks of every three-pass row length (odd and even Nu;
    cooperative, two- and three-pass register columns), rows in mode 3: each
    disk the same bits as alone with the same options, and within the loop
    gate of the oracle."""
    print("case three-pass rows ensemble", flush=True)
    shapes = [(64, 128, 0.02), (37, 192, 0.01), (48, 256, 0.02), (128, 128, 0.01), (33, 256, 0.02)]
    cols = [3, 0, 4, 4, 3]
    params = [dfk.disk_params(a, b, eps=c) for a, b, c in shapes]
    ics, mDs, hs = [], [], []
    for a, b, c in shapes:
        h, k = _lka98(orc, a, b, c)
        ics.append(h.state())
        mDs.append(h.disk_mass())
        hs.append((h, k))

    def setup(e, ds):
        e.set_row_transform(3)
        for q, d in enumerate(ds):
            e.set_fft_options(1, 0, cols[d], d=q)
            e.set_state(*ics[d], d=q)

    ens = dfk.Ensemble(params)
    setup(ens, range(len(shapes)))
    assert ens.get_row_transform() == 3
    ens.run_init(1e9, cfl=0.3, mD=mDs, max_steps=10)
    ens.advance(11)
    for d, p in enumerate(params):
        one = dfk.Ensemble(p)
        setup(one, [d])
        one.run_init(1e9, cfl=0.3, mD=mDs[d], max_steps=10)
        one.advance(11)
        for a, b in zip(ens.get_state(d), one.get_state()):
            assert np.array_equal(a, b)
        one.close()
        h, k = hs[d]
        orc.run(h, k, 1, 1e9, cfl=0.3, max_steps=10, mD=mDs[d])
        assert max(state_errs(ens.get_state(d), h.state())) <= LOOP_TOL
    ens.close()


@pytest.mark.parametrize("big", [False, True])
def test_register_rows_ensemble(orc, dfk, big):
    """A handle of disks of every register row length up to 512 (odd and even
    Nu, register and cooperative columns), rows in register mode: each disk
    the same bits as alone with the same options, one of them within the loop
    gate of the oracle.  With a disk of Nphi = 768 in the handle (big), the
    rows run in the build for lengths up to 1024: round-off agreement."""
    print(f"case register rows ensemble, big {big}", flush=True)
    shapes = [(64, 128, 0.02), (37, 192, 0.01), (48, 256, 0.02), (33, 384, 0.01), (20, 512, 0.02)]
    cols = [3, 0, 3, 0, 3]
    if big:
        shapes.append((24, 768, 0.02))
        cols.append(3)
    params = [dfk.disk_params(a, b, eps=c) for a, b, c in shapes]
    ics, mDs, hs = [], [], []
    for a, b, c in shapes:
        h, k = _lka98(orc, a, b, c)
        ics.append(h.state())
        mDs.append(h.disk_mass())
        hs.append((h, k))

    def setup(e, ds):
        e.set_update_kernel(1)
        e.set_row_transform(2)
        for q, d in enumerate(ds):
            e.set_fft_options(1, 0, cols[d], d=q)
            e.set_state(*ics[d], d=q)

    ens = dfk.Ensemble(params)
    setup(ens, range(len(shapes)))
    assert ens.get_row_transform() == 2
    ens.run_init(1e9, cfl=0.3, mD=mDs, max_steps=10)
    ens.advance(12)
    for d, p in enumerate(params):
        one = dfk.Ensemble(p)
        setup(one, [d])
        one.run_init(1e9, cfl=0.3, mD=mDs[d], max_steps=10)
        one.advance(12)
        for a, b in zip(ens.get_state(d), one.get_state()):
            if big:
                assert relerr(a, b) < 1e-13
            else:
                assert np.array_equal(a, b)
        one.close()
    h, k = hs[1]
    orc.run(h, k, 1, 1e9, dtout=1e9, cfl=0.3, max_steps=10, mD=mDs[1])
    errs = [relerr(a, b) for a, b in zip(ens.get_state(1), h.state())]
    print("  rel errors Sigma, vR, vphi (37x192 vs oracle):", errs, flush=True)
    assert max(errs) <= LOOP_TOL
    ens.close()


@pytest.mark.parametrize("rows_ip", [0, 1])
def test_mixed_transform_modes_ensemble(orc, dfk, rows_ip):
    """One handle whose disks use different column-convolution modes
    (cooperative, wave-local, workgroup in place batched and single, the
    two- and three-pass register forms, which share a launch, at one
    transform length too) and batch sizes gives each disk the same bits as
    running it alone with the same options."""
    print(f"case mixed transform modes, rows in place {rows_ip}", flush=True)
    # every Nphi factors into 2/3/4/8 (in-place rows are a handle-wide choice);
    # 2 x 40 = 80 has a factor 5, so that disk keeps the cooperative columns
    shapes = [(32, 48, 0.02), (64, 64, 0.01), (48, 24, 0.04), (40, 128, 0.01), (24, 24, 0.03),
              (96, 48, 0.02), (64, 32, 0.01), (128, 64, 0.01), (192, 48, 0.02), (128, 32, 0.02)]
    opts = [(1, 1, 0, 2), (1, 0, 1, 3), (0, 0, 2, 2), (1, 0, 2, 1), (1, 1, 2, 4),
            (1, 0, 3, 1), (0, 0, 3, 2), (1, 0, 4, 1), (0, 0, 4, 1), (1, 0, 3, 1)]   # tw rows, tw cols, mode, cb
    params = [dfk.disk_params(a, b, eps=c) for a, b, c in shapes]
    ics, mDs = [], []
    for a, b, c in shapes:
        h, _ = _lka98(orc, a, b, c)
        ics.append(h.state())
        mDs.append(h.disk_mass())

    def setup(e, ds):
        e.set_update_kernel(1)
        e.set_row_transform(rows_ip)
        e.set_threads(128, 128, 256)
        for q, d in enumerate(ds):
            e.set_fft_options(*opts[d][:3], d=q)
            e.set_shapes(0, 0, 2, opts[d][3], d=q)
            e.set_state(*ics[d], d=q)

    ens = dfk.Ensemble(params)
    setup(ens, range(len(shapes)))
    ens.run_init(1e9, mD=mDs, max_steps=12)
    ens.advance(12)
    for d, p in enumerate(params):
        one = dfk.Ensemble(p)
        setup(one, [d])
        assert one.get_fft_options(0) == ens.get_fft_options(d)
        assert one.get_shapes(0) == ens.get_shapes(d)
        one.run_init(1e9, mD=mDs[d], max_steps=12)
        one.advance(12)
        for a, b in zip(ens.get_state(d), one.get_state()):
            assert np.array_equal(a, b)
        one.close()
    ens.close()


@pytest.mark.parametrize("chains", [2, 3, 5])
def test_graph_chains_same_bits(orc, dfk, chains):
    """Disks stepped as parallel graph chains (dfk_set_chains) give the same
    states, clocks and samples as one chain, with disks that stop at
    different steps."""
    print(f"case graph chains {chains}", flush=True)
    shapes = [(32, 48, 0.02), (64, 64, 0.01), (48, 30, 0.04), (40, 130, 0.01), (24, 24, 0.03)]
    params = [dfk.disk_params(a, b, eps=c) for a, b, c in shapes]
    ics, mDs = [], []
    for a, b, c in shapes:
        h, _ = _lka98(orc, a, b, c)
        ics.append(h.state())
        mDs.append(h.disk_mass())
    caps = [20, 35, 27, 31, 12]
    out = []
    for k in (1, chains):
        e = dfk.Ensemble(params)
        e.set_chains(k)
        for d in range(len(shapes)):
            e.set_state(*ics[d], d=d)
        e.set_graph_steps(8)
        e.run_init(1e9, dtout=0.002, mD=mDs, max_steps=caps, sample_cap=64)
        e.advance(40)
        out.append((e.get_chains(), [e.get_state(d) for d in range(len(shapes))],
                    [e.status(d) for d in range(len(shapes))],
                    [e.samples(d) for d in range(len(shapes))]))
        e.close()
    assert out[0][0] == 1 and out[1][0] == chains
    for d in range(len(shapes)):
        assert out[1][2][d]["steps"] == caps[d]
        assert out[0][2][d]["t"] == out[1][2][d]["t"]
        for a, b in zip(out[0][1][d], out[1][1][d]):
            assert np.array_equal(a, b)
    for a, b in zip(out[0][3], out[1][3]):
        assert len(a) > 2 and np.array_equal(a, b)


def test_diagnostics_leave_the_run_alone(orc, dfk):
    """dfk_disk_mass, dfk_modes and dfk_timestep in the middle of a run (and
    after it) change neither the disk's clock nor its later steps and
    samples: the run then ends on the bits of an uninterrupted one."""
    print("case diagnostics mid-run", flush=True)
    Nu, Nphi, eps = 64, 64, 0.02
    h, _ = _lka98(orc, Nu, Nphi, eps)
    ic, mD = h.state(), h.disk_mass()
    out = []
    for probe in (False, True):
        e = dfk.Ensemble(dfk.disk_params(Nu, Nphi, eps=eps))
        e.set_state(*ic)
        e.run_init(1e9, dtout=0.01, mD=mD, max_steps=30)
        e.advance(15)
        if probe:
            st = e.status()
            assert st["steps"] == 15 and st["t"] > 0
            m = e.disk_mass()
            c2, _ = e.modes(2, mD)
            dt = e.timestep(0.3)
            assert m > 0 and c2 > 0 and dt > 0
            assert e.status() == st
        e.advance(20)
        st = e.status()
        if probe:
            e.disk_mass()
            assert e.status() == st
        out.append((e.get_state(), st, e.samples()))
        e.close()
    assert out[0][1]["steps"] == 30 and out[0][1] == out[1][1]
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    assert len(out[0][2]) > 2 and np.array_equal(out[0][2], out[1][2])


def test_chains_and_graph_steps_change_mid_run(orc, dfk):
    """dfk_set_chains and dfk_set_graph_steps between two dfk_advance calls
    recapture the loop's graphs and the run goes on, on the bits of a run
    that never changed them."""
    print("case chains / graph steps mid-run", flush=True)
    shapes = [(32, 48, 0.02), (64, 64, 0.01), (48, 30, 0.04), (40, 130, 0.01), (24, 24, 0.03)]
    params = [dfk.disk_params(a, b, eps=c) for a, b, c in shapes]
    ics = [_lka98(orc, a, b, c)[0] for a, b, c in shapes]
    out = []
    for change in (False, True):
        e = dfk.Ensemble(params)
        for d, h in enumerate(ics):
            e.set_state(*h.state(), d=d)
        e.set_chains(1)
        e.set_graph_steps(7)
        e.run_init(1e9, dtout=0.002, mD=[h.disk_mass() for h in ics], max_steps=30, sample_cap=64)
        e.advance(13)
        if change:
            e.set_chains(3)
            e.set_graph_steps(4)
            assert e.get_chains() > 1
        e.advance(20)
        out.append([(e.get_state(d), e.status(d), e.samples(d)) for d in range(len(shapes))])
        e.close()
    for (sa, ta, pa), (sb, tb, pb) in zip(*out):
        assert ta == tb and ta["steps"] == 30
        for a, b in zip(sa, sb):
            assert np.array_equal(a, b)
        assert len(pa) > 2 and np.array_equal(pa, pb)


@pytest.mark.parametrize("mode", ["rows", "rows_3chains", "fused", "no_selfgrav"])
def test_advance_chunks_report_samples_and_done(orc, dfk, mode):
    """dfk_advance reads the loop's state after a run (the done flags it
    returns, the sample counts it drains) from the host copy that the
    control of each captured run's last iteration writes.  Runs split into
    advances of 0, 1, 2, 3 and 5 steps (graphs of 1 to 4 steps) report the
    same all_done, status, samples and state, step by step, as the run made
    of one advance; the control rides in the row r2c (one and three chains),
    in the column convolution (fused row kernels) or in a launch of its own
    (disks without self-gravity)."""
    print(f"case advance chunks / host control copy: {mode}", flush=True)
    shapes = [(32, 48, 0.02), (40, 32, 0.01), (24, 36, 0.03), (36, 64, 0.02)]
    sg = 0 if mode == "no_selfgrav" else 1
    params = [dfk.disk_params(a, b, eps=c, selfgrav=sg) for a, b, c in shapes]
    ics = [_lka98(orc, a, b, c, selfgrav=sg)[0] for a, b, c in shapes]
    caps = [17, 11, 17, 23]   # per-disk step caps: the disks finish at different advances

    def start():
        e = dfk.Ensemble(params)
        for d, h in enumerate(ics):
            e.set_state(*h.state(), d=d)
        if mode == "fused":
            e.set_fused_rows(256)
        e.set_chains(3 if mode == "rows_3chains" else 1)
        e.set_graph_steps(4)
        e.run_init(1e9, dtout=0.003, mD=[h.disk_mass() for h in ics], max_steps=caps,
                   sample_cap=64)
        return e

    one = start()
    assert not one.advance(0)
    assert one.advance(30)
    ref = [(one.get_state(d), one.status(d), one.samples(d)) for d in range(len(shapes))]
    one.close()
    e = start()
    if mode == "fused":
        assert e.get_fused_rows() > 0
    if mode == "rows_3chains":
        assert e.get_chains() > 1
    total, chunks = 0, [0, 1, 2, 3, 5, 0, 1, 3, 2, 5, 3, 5]
    for n in chunks:
        done = e.advance(n)
        total += n
        steps = [e.status(d)["steps"] for d in range(len(shapes))]
        assert steps == [min(total, c) for c in caps], (total, steps)
        # a disk is done at the control after its last step
        assert done == (total > max(caps)), (total, done)
    assert total > max(caps)
    for d, (s, t, p) in enumerate(ref):
        assert e.status(d) == t
        assert np.array_equal(e.samples(d), p) and len(p) > 2
        for a, b in zip(e.get_state(d), s):
            assert np.array_equal(a, b)
    e.close()


def test_nonfinite_flag_stops_disk(dfk):
    Nu, Nphi = 32, 32
    print("case nonfinite flag", flush=True)
    e = dfk.Ensemble([dfk.disk_params(Nu, Nphi), dfk.disk_params(Nu, Nphi)])
    st = rough_state(Nu, Nphi, seed=2)
    e.set_state(*st, d=1)
    bad = [x.copy() for x in st]
    bad[0][5, 5] = np.nan
    e.set_state(*bad, d=0)
    e.run_init(1e9, mD=1.0, max_steps=20, stop_on_nonfinite=1)
    e.advance(25)
    s0, s1 = e.status(0), e.status(1)
    assert s0["flags"] & 1 and s0["done"] and s0["steps"] == 0
    assert s1["flags"] == 0 and s1["steps"] == 20


def test_error_paths(dfk):
    print("case error paths", flush=True)
    with pytest.raises(dfk.DiskfftError) as ei:
        dfk.Ensemble(dfk.disk_params(2, 16))
    assert ei.value.code == dfk.DFK_EINVAL
    for shape in ((4096, 16), (16, 8192)):          # beyond 2048 x 4096
        with pytest.raises(dfk.DiskfftError) as ei:
            dfk.Ensemble(dfk.disk_params(*shape))
        assert ei.value.code == dfk.DFK_EINVAL
    e = dfk.Ensemble(dfk.disk_params(16, 16))
    with pytest.raises(dfk.DiskfftError) as ei:
        e.advance(1)
    assert ei.value.code == dfk.DFK_ESTATE


def test_streaming_update_size_limit(dfk):
    """The streaming update addresses a state buffer with 31-bit offsets: a
    handle of more than 89 478 485 zones (eleven 2048 x 4096 disks, 92 M
    zones) is refused with DFK_EUNSUPPORTED and keeps the LDS tile update;
    ten such disks (84 M zones) take it.  Disks without self-gravity: no
    spectra or K-hat to build."""
    print("case streaming update size limit (92 M and 84 M zones)", flush=True)
    big = [dfk.disk_params(2048, 4096, selfgrav=0) for _ in range(11)]
    e = dfk.Ensemble(big)
    with pytest.raises(dfk.DiskfftError) as ei:
        e.set_update_kernel(1)
    assert ei.value.code == dfk.DFK_EUNSUPPORTED
    assert e.get_update_kernel() == 0
    e.close()
    e = dfk.Ensemble(big[:10])
    e.set_update_kernel(1)
    assert e.get_update_kernel() == 1
    e.close()
