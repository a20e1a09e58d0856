"""Shared seeded inputs for the parity tests."""
import numpy as np


def grid(Nu, Nphi, Rin=0.25, Rout=1.0):
    """hydro.hpp:31-38 geometry, in numpy (used only to shape test inputs)."""
    umin, umax = np.log(Rin), np.log(Rout)
    du = (umax - umin) / Nu
    R = np.exp(umin + (np.arange(Nu) + 0.5) * du)
    phi = (np.arange(Nphi) + 0.5) * 2 * np.pi / Nphi
    return R, phi, du


def rough_state(Nu, Nphi, seed=1, Rin=0.25, Rout=1.0, noise=0.05, vnoise=0.05):
    """A ring disk with random perturbations in every field: every branch of
    the MUSCL reconstruction (both upwind directions, both slope signs,
    flat regions) is exercised."""
    rng = np.random.default_rng(seed)
    R, phi, _ = grid(Nu, Nphi, Rin, Rout)
    S0 = 0.372 * np.exp(-(R - 0.45) ** 2 / 0.05)
    Sig = S0[:, None] * (1 + noise * rng.standard_normal((Nu, Nphi)))
    Sig = np.maximum(Sig, 1e-3)
    vR = vnoise * rng.standard_normal((Nu, Nphi))
    vp = np.sqrt(0.6 / R)[:, None] * (1 + vnoise * rng.standard_normal((Nu, Nphi)))
    return Sig, vR, vp


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    if den == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / den)
