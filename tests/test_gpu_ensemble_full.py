"""GPU: the 64-disk ensemble of the benchmark at full size (7.7 M zones,
grids 128^2 .. 512^2 drawn with seed 2026, tuned throughput shapes), checked
through properties that do not depend on its size:

  - every disk's mass sum(Sigma dA) is conserved to round-off (radial boundary
    fluxes are zero and phi is periodic, hydro.hpp:103-106, 143-179; no zone
    reaches the floor);
  - every zone stays finite and no disk raises a flag;
  - one disk of each grid gives the same bits run alone with the same
    options (the kernels are the same per disk);
  - one disk of every grid matches the oracle's run of the same steps from
    the same initial state within 1e-10 relative.
"""
import os
import sys

import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STEPS = 20


@pytest.fixture(scope="module")
def dfk():
    import diskfft_amd
    if diskfft_amd.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return diskfft_amd


@pytest.fixture(scope="module")
def ensemble(dfk):
    sys.path.insert(0, os.path.join(ROOT, "heavy-disks_amd"))
    import autotune
    import ensemble_cfg
    disks = ensemble_cfg.draw(64, seed=2026)
    print(f"case full ensemble: 64 disks, {sum(d['Nu'] * d['Nphi'] for d in disks)} zones",
          flush=True)
    ens, ics = ensemble_cfg.build(dfk, disks)
    assert autotune.apply(ens, regime="throughput") == 64
    m0 = [ens.disk_mass(d) for d in range(64)]
    ens.set_graph_steps(10)
    ens.run_init(1e30, cfl=0.4, mD=[ic[3] for ic in ics], max_steps=STEPS, sample_cap=64)
    ens.advance(STEPS + 2)
    yield dfk, disks, ics, ens, m0
    ens.close()


def test_full_ensemble_conserves_mass_and_stays_finite(ensemble):
    dfk, disks, ics, ens, m0 = ensemble
    worst = 0.0
    for d in range(64):
        st = ens.status(d)
        assert st["done"] and st["steps"] == STEPS and st["t"] > 0 and not st["flags"], (d, st)
        assert all(np.all(np.isfinite(x)) for x in ens.get_state(d)), d
        rel = abs(ens.disk_mass(d) - m0[d]) / m0[d]
        worst = max(worst, rel)
    print(f"  worst relative mass change over {STEPS} steps: {worst:.2e}", flush=True)
    assert worst < 1e-12


def test_full_ensemble_disks_match_single_runs(ensemble):
    dfk, disks, ics, ens, m0 = ensemble
    picked = {}
    for d, x in enumerate(disks):
        picked.setdefault((x["Nu"], x["Nphi"]), d)
    for (Nu, Nphi), d in sorted(picked.items()):
        print(f"  disk {d} ({Nu}x{Nphi}) alone", flush=True)
        one = dfk.Ensemble(ics[d][4])
        one.set_update_kernel(ens.get_update_kernel())
        one.set_row_transform(ens.get_row_transform())
        one.set_threads(*ens.get_threads())
        one.set_shapes(*ens.get_shapes(d))
        one.set_fft_options(*ens.get_fft_options(d))
        assert one.get_shapes(0) == ens.get_shapes(d)
        one.set_state(*ics[d][:3])
        one.run_init(1e30, cfl=0.4, mD=ics[d][3], max_steps=STEPS, sample_cap=64)
        one.advance(STEPS + 2)
        assert one.status()["t"] == ens.status(d)["t"]
        for a, b in zip(one.get_state(), ens.get_state(d)):
            assert np.array_equal(a, b)
        one.close()


def test_full_ensemble_disks_match_oracle(orc, ensemble):
    """One disk of every grid of the ensemble (128^2, 192^2, 256^2, 384 x 512,
    512^2): its state after the ensemble's own STEPS steps within 1e-10 of the
    oracle's STEPS steps from the same initial state, same clock."""
    dfk, disks, ics, ens, m0 = ensemble
    picked = {}
    for d, x in enumerate(disks):
        picked.setdefault((x["Nu"], x["Nphi"]), d)
    assert len(picked) == 5
    for (Nu, Nphi), d in sorted(picked.items()):
        p = ics[d][4]
        h = orc.Hydro(p["Nu"], p["Nphi"], p["Rin"], p["Rout"], p["K"], p["gamma"], p["gmstar"])
        k = orc.Kalnajs(p["Nu"], p["Nphi"], p["Rin"], p["Rout"], p["eps"], p["G"])
        h.set_state(*ics[d][:3])
        _, tf, _, steps = orc.run(h, k, 1, 1e30, cfl=0.4, max_steps=STEPS, mD=ics[d][3])
        errs = [relerr(a, b) for a, b in zip(ens.get_state(d), h.state())]
        st = ens.status(d)
        print(f"  disk {d} ({Nu}x{Nphi}) {STEPS} steps vs oracle:", errs, flush=True)
        assert steps == STEPS and st["steps"] == STEPS
        assert abs(st["t"] - tf) / tf < 1e-13
        assert max(errs) <= 1e-10
