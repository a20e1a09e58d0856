"""GPU loop over many steps: snapshot/resume, the m = 2 growth rate and
the reference's published physics, and the benchmark configurations run
through against the oracle.

The bar: gamma and Omega_p within 1% of the FP64 CPU path from the same
state (achieved: 1e-12 relative).  The reference's own 128 x 128 setup
(diskfft/README.md:27-30: LKA98 ring, tlim 9, amp 1e-3) runs in a couple of
seconds on both sides; its published numbers (diskfft/README.md:27,37;
README.md:42) are checked at the agreement the library reaches with them
(their fit window is not published, so a few percent).
"""
import os

import numpy as np
import pytest

from helpers import relerr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dfk():
    import diskfft_amd
    if diskfft_amd.device_count() < 1:
        pytest.fail("no HIP device visible")
    return diskfft_amd


def _lka(orc, Nu, Nphi, eps, amp=1e-3):
    h = orc.Hydro(Nu, Nphi)
    k = orc.Kalnajs(Nu, Nphi, 0.25, 1.0, eps)
    orc.ic_lka98(h, k, 1, amp)
    return h, k


@pytest.mark.parametrize("kind", [0, 1])
def test_snapshot_resume_is_bitwise(orc, dfk, tmp_path, kind):
    print(f"case snapshot resume kind {kind}", flush=True)
    Nu, Nphi, eps = 64, 64, 0.02
    h, _ = _lka(orc, Nu, Nphi, eps)
    st, mD = h.state(), h.disk_mass()
    p = dfk.disk_params(Nu, Nphi, eps=eps)
    a = dfk.Ensemble(p)
    a.set_update_kernel(kind)
    a.set_state(*st)
    a.run_init(1e9, dtout=0.01, mD=mD, max_steps=40)
    a.advance(45)
    sa = a.samples()
    b = dfk.Ensemble(p)
    b.set_update_kernel(kind)
    b.set_state(*st)
    b.run_init(1e9, dtout=0.01, mD=mD, max_steps=40)
    b.advance(20)
    s1 = b.samples()
    path = os.path.join(tmp_path, "disk.snap")
    b.snapshot_save(path)
    b.close()
    c = dfk.Ensemble(p)
    c.set_update_kernel(kind)
    c.run_init(1e9, dtout=0.01, mD=mD, max_steps=40)
    c.snapshot_load(path)
    c.advance(25)
    s2 = c.samples()
    assert c.status()["steps"] == 40 and c.status()["t"] == a.status()["t"]
    for x, y in zip(a.get_state(), c.get_state()):
        assert np.array_equal(x, y)
    assert np.array_equal(np.concatenate([s1, s2]), sa)
    with pytest.raises(dfk.DiskfftError):
        d = dfk.Ensemble(dfk.disk_params(Nu, Nphi, eps=0.03))
        d.run_init(1.0, mD=mD)
        d.snapshot_load(path)


def test_growth_rate_matches_oracle(orc, dfk):
    """LKA98 ring, 128 x 128, eps 0.0025, tlim 9: gamma and Omega_p of the m=2
    mode from the GPU loop within 1% of the oracle's."""
    print("case growth rate 128x128 eps=0.0025 tlim=9", flush=True)
    Nu = Nphi = 128
    eps, tlim = 0.0025, 9.0
    h, k = _lka(orc, Nu, Nphi, eps)
    mD = h.disk_mass()
    e = dfk.Ensemble(dfk.disk_params(Nu, Nphi, eps=eps))
    e.set_state(*h.state())
    ((sg, stg),) = e.run(tlim, dtout=0.25, cfl=0.3, mD=mD)
    so, tf, c2f, steps = orc.run(h, k, 1, tlim, dtout=0.25, cfl=0.3, mD=mD)
    gg = dfk.extract_growth(sg)
    go = dfk.extract_growth(so)
    print("  gpu", gg, "\n  oracle", go, f"\n  steps {stg['steps']} vs {steps}", flush=True)
    assert gg["grew"] and go["grew"]
    assert abs(gg["gamma"] - go["gamma"]) / abs(go["gamma"]) < 0.01
    assert abs(gg["omega_p"] - go["omega_p"]) / abs(go["omega_p"]) < 0.01
    # reference result for this setup (diskfft/README.md:27,30): gamma 0.757,
    # Omega_p 1.49 (achieved: 1.05% and 0.3%; Omega_p rounds to the published 1.49)
    assert abs(gg["gamma"] - 0.757) / 0.757 < 0.015
    assert abs(gg["omega_p"] - 1.49) / 1.49 < 0.005


def test_softening_sweep_published(dfk, orc):
    """The reference's softening sweep at 128 x 128 (diskfft/README.md:27):
    gamma 0.53 / 0.64 / 0.69 / 0.757 for eps 0.02 / 0.01 / 0.005 / 0.0025
    (achieved within 1.5%, 0.6%, 3.2%, 1.05%), rising as eps falls.  eps 0.04
    (published 0.36) does not grow tenfold by tlim 9, so its fit spans the
    whole run; it is only required to grow slower than eps 0.02."""
    print("case softening sweep 128x128 tlim=9", flush=True)
    pub = {0.04: 0.36, 0.02: 0.53, 0.01: 0.64, 0.005: 0.69, 0.0025: 0.757}
    got = {}
    for eps in sorted(pub, reverse=True):
        h, _ = _lka(orc, 128, 128, eps)
        e = dfk.Ensemble(dfk.disk_params(128, 128, eps=eps))
        e.set_state(*h.state())
        ((sg, st),) = e.run(9.0, dtout=0.25, cfl=0.3, mD=h.disk_mass())
        g = dfk.extract_growth(sg)
        got[eps] = g["gamma"]
        print(f"  eps {eps}: gamma {g['gamma']:.4f} (published {pub[eps]}), grew {g['grew']}",
              flush=True)
        if eps < 0.04:
            assert g["grew"] and abs(g["gamma"] - pub[eps]) / pub[eps] < 0.04
        e.close()
    g = [got[x] for x in sorted(pub, reverse=True)]
    assert all(a < b for a, b in zip(g, g[1:]))


def test_lka98_256_published_and_oracle(dfk, orc):
    """`diskdisk 256 256 0.005 16 1 1e-3` (diskfft/README.md:36-49): gamma 0.717
    and mass drift 0.00 (:37), saturated |C2|/|C0| 0.24 and linear pattern
    speed 1.46 falling on saturation (README.md:42).  Achieved: gamma 0.692
    (3.5%; the reference's fit window is not published), saturation (mean of
    the final quarter) 0.236 (1.5%), Omega_p 1.453 (0.5%).  The same run on the
    oracle from the same state: gamma and Omega_p within 1e-9, step count
    equal."""
    print("case LKA98 256x256 eps 0.005 tlim 16 (published + oracle)", flush=True)
    import oracle as orc_fast
    h = orc_fast.Hydro(256, 256, fast=True)
    k = orc_fast.Kalnajs(256, 256, 0.25, 1.0, 0.005, fast=True)
    orc_fast.ic_lka98(h, k, 1, 1e-3)
    mD = h.disk_mass()
    e = dfk.Ensemble(dfk.disk_params(256, 256, eps=0.005))
    e.set_state(*h.state())
    e.set_graph_steps(500)
    ((sg, stg),) = e.run(16.0, dtout=0.25, cfl=0.3, mD=mD, chunk=20000)
    gg = dfk.extract_growth(sg)
    q = int(0.75 * len(sg))
    om_sat = -0.5 * np.polyfit(sg[q:, 0], np.unwrap(sg[:, 3])[q:], 1)[0]
    print(f"  gpu: gamma {gg['gamma']:.4f} Omega_p {gg['omega_p']:.4f} sat {gg['sat']:.4f} "
          f"saturated Omega_p {om_sat:.4f} steps {stg['steps']}", flush=True)
    assert gg["grew"] and stg["t"] == 16.0
    assert abs(gg["gamma"] - 0.717) / 0.717 < 0.04
    assert abs(gg["sat"] - 0.24) / 0.24 < 0.03
    assert abs(gg["omega_p"] - 1.46) / 1.46 < 0.01 and om_sat < gg["omega_p"]
    # mass drift |C0| - 1 (published 0.00; achieved 1.3e-10, the Sigma floor
    # adding mass in the ring's empty outskirts)
    assert np.max(np.abs(sg[:, 1] - 1.0)) < 1e-9
    orc_fast.set_threads(max(1, min(16, os.cpu_count() or 1)))
    so, tf, c2f, steps = orc_fast.run_chunked(h, k, 1, 16.0, dtout=0.25, cfl=0.3, mD=mD)
    go = dfk.extract_growth(so)
    print(f"  oracle: gamma {go['gamma']:.4f} Omega_p {go['omega_p']:.4f} steps {steps}",
          flush=True)
    assert steps == stg["steps"]
    assert abs(gg["gamma"] - go["gamma"]) / go["gamma"] < 1e-9
    assert abs(gg["omega_p"] - go["omega_p"]) / go["omega_p"] < 1e-9


def _powerlaw_case(dfk, orc, Nu, Nphi, eps=0.01):
    """The benchmark's power-law model, its equilibrium from the oracle's
    gravity (as tests/longrun/run_growth.py builds it)."""
    import oracle as orc_fast
    m = dfk.model_defaults(dfk.MODEL_POWERLAW)
    k = orc_fast.Kalnajs(Nu, Nphi, m.Rin, m.Rout, eps, m.G, fast=True)
    p, S, U, W, mD, _ = dfk.model_ic(m, Nu, Nphi, eps,
                                     potential_forces=lambda x: k.forces(k.potential(x)))
    return p, (S, U, W), mD, k


@pytest.mark.parametrize("shapes", ["default", "tuned"])
def test_cfg1_powerlaw_128_2000_steps(dfk, orc, shapes):
    """Configuration 1 at its own parameters: power-law disk 128 x 128,
    r in [0.1, 1], eps 0.01, 2000 steps at CFL 0.4, on the GPU and the
    oracle from the same state: the state after 2000 steps within 1e-10 (Sigma
    against its maximum, the velocities against the speed scale, as in the
    ten-step gate; each field against its own maximum within 1e-9: v_R, which
    starts at zero, is the smallest field), the clock to 1e-12, the |C2|/|C0|
    series and its fit to 1e-9.  With the library's default shapes (achieved
    7e-12 per field) and with the tuned latency shapes, the three-pass register
    transforms (achieved 1e-11 for Sigma and v_phi, 1.2e-10 for v_R against
    its own maximum: 2000 steps of a growing mode amplify the round-off of
    transforms that associate differently from the oracle's)."""
    print(f"case cfg1: power-law 128x128, 2000 steps at CFL 0.4, {shapes} shapes", flush=True)
    import autotune
    import oracle as orc_fast
    p, ic, mD, k = _powerlaw_case(dfk, orc, 128, 128)
    e = dfk.Ensemble(p)
    if shapes == "tuned":
        assert autotune.apply(e, regime="latency")
        assert e.get_row_transform() == 3 and e.get_fft_options(0)[2] == 4
    e.set_state(*ic)
    e.set_graph_steps(500)
    ((sg, stg),) = e.run(1e9, dtout=0.25, cfl=0.4, mD=mD, max_steps=2000)
    h = orc_fast.Hydro(128, 128, p["Rin"], p["Rout"], p["K"], p["gamma"], p["gmstar"], fast=True)
    h.set_state(*ic)
    orc_fast.set_threads(max(1, min(16, os.cpu_count() or 1)))
    so, tf, c2f, steps = orc_fast.run(h, k, 1, 1e9, dtout=0.25, cfl=0.4, max_steps=2000, mD=mD)
    got, ref = e.get_state(), h.state()
    errs = [relerr(a, b) for a, b in zip(got, ref)]
    vscale = max(np.max(np.abs(ref[1])), np.max(np.abs(ref[2])))
    gate = [errs[0]] + [float(np.max(np.abs(got[f] - ref[f])) / vscale) for f in (1, 2)]
    print(f"  t gpu {stg['t']:.6f} oracle {tf:.6f}; state rel errors {errs}, "
          f"velocities against the speed scale {gate[1:]}", flush=True)
    assert stg["steps"] == steps == 2000
    assert abs(stg["t"] - tf) / tf < 1e-12
    assert max(gate) <= 1e-10
    assert max(errs) <= 1e-9
    assert sg.shape == so.shape and relerr(sg[:, 2], so[:, 2]) < 1e-9
    gg, go = dfk.extract_growth(sg), dfk.extract_growth(so)
    assert abs(gg["gamma"] - go["gamma"]) / abs(go["gamma"]) < 1e-9
    assert abs(gg["omega_p"] - go["omega_p"]) / abs(go["omega_p"]) < 1e-9


@pytest.mark.parametrize("shapes", ["default", "tuned"])
def test_cfg2_powerlaw_256_twenty_orbits(dfk, orc, shapes):
    """Configuration 2: the power-law disk at 256 x 256 through about 20 outer
    orbits (t = 89, CFL 0.4, samples every 0.01) on the GPU, against the
    oracle's run of the same initial state (tests/longrun/oracle_pl256.npz,
    written by tests/longrun/run_growth.py pl256; 784k steps, 32 CPU-minutes).
    gamma and Omega_p of the growth window within 1e-9 (achieved 1e-12); the
    saturated amplitude (peak |C2|/|C0| after the window) within 1%; the step
    count within 0.5% (the late, turbulent state amplifies round-off, so the
    two runs' steps differ by ~0.06%).  About 20 s on the GPU.  With the
    library's default shapes and with the tuned latency shapes the benchmark
    runs (the three-pass register transforms)."""
    print(f"case cfg2: power-law 256x256 to t = 89, {shapes} shapes", flush=True)
    import autotune
    z = np.load(os.path.join(ROOT, "tests", "longrun", "oracle_pl256.npz"))
    p, ic, mD, k = _powerlaw_case(dfk, orc, 256, 256)
    e = dfk.Ensemble(p)
    if shapes == "tuned":
        assert autotune.apply(e, regime="latency")
        assert e.get_row_transform() == 3 and e.get_fft_options(0)[2] == 4
    e.set_state(*ic)
    e.set_graph_steps(500)
    ((sg, stg),) = e.run(89.0, dtout=0.01, cfl=0.4, mD=mD, chunk=20000)
    so = z["samples"]
    ex = dict(lo_factor=10.0, hi_factor=100.0, sat_frac=0.5)
    gg, go = dfk.extract_growth(sg, **ex), dfk.extract_growth(so, **ex)
    print(f"  gpu {gg}\n  oracle {go}\n  steps {stg['steps']} / {int(z['steps'])}", flush=True)
    assert stg["t"] == 89.0 and float(z["t"]) == 89.0
    assert sg.shape == so.shape
    assert gg["grew"] and go["grew"]
    assert abs(gg["gamma"] - go["gamma"]) / go["gamma"] < 1e-9
    assert abs(gg["omega_p"] - go["omega_p"]) / go["omega_p"] < 1e-9
    assert abs(gg["sat_peak"] - go["sat_peak"]) / go["sat_peak"] < 0.01
    assert abs(stg["steps"] - int(z["steps"])) / int(z["steps"]) < 0.005
    # the growth window's samples themselves
    w = sg[:, 0] <= gg["t_hi"]
    assert relerr(sg[w, 2], so[w, 2]) < 1e-9


def test_snapshot_load_mid_run_multi_disk(orc, dfk, tmp_path):
    """A snapshot saved at an odd step of a single-disk run and loaded at an
    odd step into disk 1 of a two-disk handle whose dfk_run_init gave other
    loop parameters: the disk continues the saved run (its tlim, dtout, cfl,
    mD and step cap) and ends on the bits and samples of an uninterrupted
    run; the other disk is left alone."""
    print("case snapshot load mid-run into a two-disk handle", flush=True)
    (Nu, Nphi, eps), (Nu2, Nphi2, eps2) = (48, 36, 0.01), (64, 64, 0.02)
    h, _ = _lka(orc, Nu, Nphi, eps)
    st, mD = h.state(), h.disk_mass()
    h2, _ = _lka(orc, Nu2, Nphi2, eps2)
    p, p2 = dfk.disk_params(Nu, Nphi, eps=eps), dfk.disk_params(Nu2, Nphi2, eps=eps2)
    ref = dfk.Ensemble(p)
    ref.set_state(*st)
    ref.run_init(1e9, dtout=0.01, cfl=0.3, mD=mD, max_steps=40)
    ref.advance(45)
    s_ref = ref.samples()
    b = dfk.Ensemble(p)
    b.set_state(*st)
    b.run_init(1e9, dtout=0.01, cfl=0.3, mD=mD, max_steps=40)
    b.advance(17)
    t_saved = b.status()["t"]
    path = os.path.join(tmp_path, "disk.snap")
    b.snapshot_save(path)
    b.close()
    c = dfk.Ensemble([p2, p])
    c.set_state(*h2.state(), d=0)
    c.set_state(*st, d=1)
    c.run_init([1e9, 5.0], dtout=[0.01, 0.05], cfl=[0.3, 0.2],
               mD=[h2.disk_mass(), 2.0 * mD], max_steps=[30, 100])
    c.advance(13)
    c.snapshot_load(path, d=1)
    c.advance(40)
    assert c.status(1)["steps"] == 40 and c.status(1)["t"] == ref.status()["t"]
    for x, y in zip(ref.get_state(), c.get_state(1)):
        assert np.array_equal(x, y)
    s_c = c.samples(1)
    # the saved run had not yet taken its sample at t_saved (a sample is
    # taken at the top of the next iteration): the resumed run takes it
    tail = s_ref[s_ref[:, 0] >= t_saved]
    assert len(s_c) > 2 and np.array_equal(s_c, tail)
    assert c.status(0)["steps"] == 30
