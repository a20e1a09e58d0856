"""Writes tests/golden/hydro_ref.npz from the reference's own hydro.hpp.

The reference hydro (diskfft/src/hydro.hpp) is compiled from its source in the
reference checkout by oracle/Makefile (target `ref`, output oracle/_ref/).
This script drives it through that harness on a seeded 24x20 state with a
fixed self-gravity field for 6 steps (DiskHydro::timestep + DiskHydro::step,
hydro.hpp:48-70) and stores inputs and outputs as a fixture, so the check
travels to machines without the reference.

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle  # noqa: E402
from helpers import rough_state  # noqa: E402


def main():
    Nu, Nphi, nsteps = 24, 20, 6
    Sig, vR, vp = rough_state(Nu, Nphi, seed=2024)
    rng = np.random.default_rng(99)
    gR = -0.4 + 0.05 * rng.standard_normal((Nu, Nphi))
    gp = 0.05 * rng.standard_normal((Nu, Nphi))
    r = oracle.RefHydro(Nu, Nphi)
    r.set_state(Sig, vR, vp)
    dts = []
    for _ in range(nsteps):
        dt = r.timestep(0.3)
        dts.append(dt)
        r.stage(0, dt, gR, gp)
    S1, R1, P1 = r.state()
    np.savez_compressed(os.path.join(HERE, "hydro_ref.npz"), Nu=Nu, Nphi=Nphi, nsteps=nsteps,
                        Sig0=Sig, vR0=vR, vp0=vp, gR=gR, gp=gp, dt=np.array(dts),
                        Sig=S1, vR=R1, vp=P1)
    print("wrote hydro_ref.npz", Nu, Nphi, nsteps)


if __name__ == "__main__":
    main()
