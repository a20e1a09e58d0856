"""GPU: the time loop with fused row kernels (dfk_set_fused_rows), through the
C ABI, against the oracle and against the four-launch loop.

Each iteration is then the column convolution (carrying the loop control)
and one kernel per block of ti full rows running the row c2r of the rows its
forces read, the update and the row r2c of the new Sigma.  Bar: the
1e-10 relative after 10 steps against the oracle (diskmain.cpp:92-110),
and the same against the four-launch loop on the same device: the formulas
are the same, but the two builds contract different products to FMAs, and
v_R (the smallest field, so the largest relative error in the max-norm)
carries the round-off of 10 steps (~5e-12 at 256 x 256).
"""
import os

import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu

LOOP_TOL = 1e-10
SAME_TOL = 1e-10


@pytest.fixture(scope="module")
def dfk():
    import diskfft_amd
    if diskfft_amd.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return diskfft_amd


def _lka98(orc, Nu, Nphi, eps, amp=1e-3):
    h = orc.Hydro(Nu, Nphi)
    k = orc.Kalnajs(Nu, Nphi, 0.25, 1.0, eps)
    orc.ic_lka98(h, k, 1, amp)
    return h, k


@pytest.mark.parametrize("threads", [256, 512, 1024])
@pytest.mark.parametrize("Nu,Nphi,eps,ti", [(64, 64, 0.02, 2), (64, 64, 0.02, 4), (64, 64, 0.02, 6),
                                            (256, 256, 0.005, 2), (48, 36, 0.02, 4),
                                            (35, 48, 0.01, 6), (3, 4, 0.02, 2), (3, 4, 0.02, 6),
                                            (128, 192, 0.01, 4), (99, 128, 0.01, 2)])
def test_fused_loop_parity(orc, dfk, Nu, Nphi, eps, ti, threads):
    """10 steps with fused row kernels: oracle and four-launch loop within
    1e-10, same clock and first sample."""
    print(f"case fused loop {Nu}x{Nphi} eps={eps} ti={ti} threads={threads}", flush=True)
    h, k = _lka98(orc, Nu, Nphi, eps)
    ic, mD = h.state(), h.disk_mass()
    so, tf, c2f, steps = orc.run(h, k, 1, 1e9, dtout=0.25, cfl=0.3, max_steps=10, mD=mD)
    runs = []
    for fused in (threads, 0):
        e = dfk.Ensemble(dfk.disk_params(Nu, Nphi, eps=eps))
        e.set_update_kernel(0)
        e.set_row_transform(1)
        e.set_shapes(ti, 16, 1, 1)
        e.set_fused_rows(fused)
        assert e.get_fused_rows() == fused
        e.set_state(*ic)
        e.run_init(1e9, dtout=0.25, cfl=0.3, mD=mD, max_steps=10)
        e.advance(12)
        st = e.status()
        assert st["done"] and st["steps"] == 10
        runs.append((e.get_state(), st, e.samples()))
        e.close()
    (fs, fst, fsm), (us, ust, usm) = runs
    assert abs(fst["t"] - tf) / tf < 1e-13
    errs = [relerr(a, b) for a, b in zip(fs, h.state())]
    same = [relerr(a, b) for a, b in zip(fs, us)]
    print("  rel errors vs oracle:", errs, " vs four launches:", same,
          " bit-identical:", all(np.array_equal(a, b) for a, b in zip(fs, us)), flush=True)
    assert max(errs) <= LOOP_TOL
    assert max(same) <= SAME_TOL
    assert fsm.shape == so.shape
    assert abs(fsm[0, 2] - so[0, 2]) / so[0, 2] < 1e-12


def test_fused_falls_back(dfk):
    """No fused rows where they cannot run: Nphi = 80 has a factor 5 (no
    in-place row transform); a 256-wide tile of 6 rows does not fit the LDS;
    the streaming update kernel is selected."""
    print("case fused fallback 96x80", flush=True)
    e = dfk.Ensemble(dfk.disk_params(96, 80, eps=0.01))
    e.set_fused_rows(512)
    assert e.get_fused_rows() == 0
    e.close()
    e = dfk.Ensemble(dfk.disk_params(256, 256, eps=0.01))
    e.set_update_kernel(0)
    e.set_shapes(6, 0, 0, 0)
    e.set_fused_rows(512)
    assert e.get_fused_rows() == 0
    e.set_shapes(2, 0, 0, 0)
    assert e.get_fused_rows() == 512
    e.close()
    e = dfk.Ensemble(dfk.disk_params(64, 64, eps=0.01))
    e.set_update_kernel(1)
    e.set_fused_rows(512)
    assert e.get_fused_rows() == 0   # fused rows ride on the LDS tile update
    e.set_update_kernel(0)
    assert e.get_fused_rows() == 512
    with pytest.raises(dfk.DiskfftError):
        e.set_fused_rows(128)
    e.close()


@pytest.mark.parametrize("chains", [1, 3])
def test_fused_ensemble_matches_single_disks(orc, dfk, chains):
    """Disks of different shapes and stopping steps in one fused handle give
    the bits of each disk run alone with the fused loop."""
    print(f"case fused ensemble vs single, chains {chains}", flush=True)
    shapes = [(32, 48, 0.02), (64, 64, 0.01), (48, 24, 0.04), (40, 128, 0.01), (24, 24, 0.03)]
    tis = [4, 2, 6, 4, 8]
    params = [dfk.disk_params(a, b, eps=c) for a, b, c in shapes]
    ics, mDs = [], []
    for a, b, c in shapes:
        h, _ = _lka98(orc, a, b, c)
        ics.append(h.state())
        mDs.append(h.disk_mass())
    caps = [20, 35, 27, 31, 12]

    def setup(e, ds):
        e.set_row_transform(1)
        e.set_fused_rows(512)
        for q, d in enumerate(ds):
            e.set_shapes(tis[d], 0, 1, 1, d=q)
            e.set_state(*ics[d], d=q)

    ens = dfk.Ensemble(params)
    setup(ens, range(len(shapes)))
    ens.set_chains(chains)
    assert ens.get_fused_rows() == 512
    ens.run_init(1e9, dtout=0.002, mD=mDs, max_steps=caps, sample_cap=64)
    ens.advance(40)
    for d, p in enumerate(params):
        one = dfk.Ensemble(p)
        setup(one, [d])
        one.run_init(1e9, dtout=0.002, mD=mDs[d], max_steps=caps[d], sample_cap=64)
        one.advance(40)
        assert ens.status(d)["steps"] == caps[d]
        assert ens.status(d)["t"] == one.status()["t"]
        for a, b in zip(ens.get_state(d), one.get_state()):
            assert np.array_equal(a, b)
        assert np.array_equal(ens.samples(d), one.samples())
        one.close()
    ens.close()


def test_fused_output_schedule(orc, dfk):
    """Output schedule and last-step clip (diskmain.cpp:92-110) over ~300
    steps with the control riding in the convolution launch."""
    Nu, Nphi, eps, tlim = 64, 64, 0.02, 0.6
    print("case fused loop samples 64x64 tlim=0.6", flush=True)
    h, k = _lka98(orc, Nu, Nphi, eps)
    e = dfk.Ensemble(dfk.disk_params(Nu, Nphi, eps=eps))
    e.set_row_transform(1)
    e.set_fused_rows(1024)
    e.set_state(*h.state())
    mD = h.disk_mass()
    so, tf, c2f, steps = orc.run(h, k, 1, tlim, dtout=0.25, cfl=0.3, mD=mD)
    ((sg, st),) = e.run(tlim, dtout=0.25, cfl=0.3, mD=mD, chunk=100)
    assert st["steps"] == steps
    assert st["t"] == pytest.approx(tf, rel=1e-14)
    assert sg.shape == so.shape
    assert np.max(np.abs(sg[:, 0] - so[:, 0])) < 1e-12
    assert relerr(sg[:, 2], so[:, 2]) < 1e-9
    assert abs(st["c2_final"] - c2f) / c2f < 1e-9
    errs = [relerr(a, b) for a, b in zip(e.get_state(), h.state())]
    print(f"  {steps} steps, rel errors:", errs, flush=True)
    assert max(errs) < 1e-9


def test_fused_snapshot_resume_and_profile(orc, dfk, tmp_path):
    """A fused run resumed from a snapshot, and one profiled in the middle
    (dfk_profile_step, which reruns the loop kernels), end on the bits of an
    uninterrupted run."""
    print("case fused snapshot resume + profile", flush=True)
    Nu, Nphi, eps = 64, 64, 0.02
    h, _ = _lka98(orc, Nu, Nphi, eps)
    st, mD = h.state(), h.disk_mass()
    p = dfk.disk_params(Nu, Nphi, eps=eps)

    def fused():
        e = dfk.Ensemble(p)
        e.set_row_transform(1)
        e.set_fused_rows(256)
        return e

    a = fused()
    a.set_state(*st)
    a.run_init(1e9, dtout=0.01, mD=mD, max_steps=40)
    a.advance(45)
    sa = a.samples()
    b = fused()
    b.set_state(*st)
    b.run_init(1e9, dtout=0.01, mD=mD, max_steps=40)
    b.advance(21)
    s1 = b.samples()
    path = os.path.join(tmp_path, "disk.snap")
    b.snapshot_save(path)
    ms = b.profile_step(5)
    assert ms[1] == 0.0 and ms[3] == 0.0 and ms[2] > 0 and ms[4] > 0
    b.advance(25)
    assert b.status()["steps"] == 40 and b.status()["t"] == a.status()["t"]
    for x, y in zip(a.get_state(), b.get_state()):
        assert np.array_equal(x, y)
    c = fused()
    c.run_init(1e9, dtout=0.01, mD=mD, max_steps=40)
    c.snapshot_load(path)
    c.advance(25)
    s2 = c.samples()
    assert c.status()["steps"] == 40 and c.status()["t"] == a.status()["t"]
    for x, y in zip(a.get_state(), c.get_state()):
        assert np.array_equal(x, y)
    assert np.array_equal(np.concatenate([s1, s2]), sa)
