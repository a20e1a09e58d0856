"""Taper width of the power-law disk model: a CPU-only study (host model code
of libdiskfft + the CPU oracle; no GPU needed).

    python tests/longrun/taper_study.py --out tests/longrun/taper_study.json [--nu 128] [--tlim 89]

For each tanh taper width w (Sigma ~ r^slope x tapers at 0.15 / 0.9, M_d = M*):

  1. Toomre Q.  K is scanned geometrically (as dfk_model_state does) and the
     largest min-Q reachable is recorded over the whole grid [Rin, Rout] and
     over the disk body [0.2, 0.8]; whether min Q = 1.3 (--qmin) is reachable
     on each (for M_d / M* = --mass-ratio, default 1).
  2. Dynamics.  The oracle runs the model (K from Q_min = 1.3 over the whole
     grid where reachable, else over the body) for `tlim` time units (20 outer
     orbits: t = 89) at CFL 0.4, eps 0.01, sampling |C2|/|C0| and arg C2 every
     0.05 and the m = 2 Fourier amplitude of Sigma/Sigma_0 per radius every
     0.5 time units.
  3. Mode.  gamma and Omega_p of the first growth (the extractor's windowed
     fit, 10x .. 100x the initial amplitude); the corotation radius R_c where
     Omega(R) = v_phi / R equals Omega_p; the radius where the relative m = 2
     amplitude peaks at the end of that growth.  An inner-edge mode has R_c
     and the amplitude peak at the inner taper; a global spiral has them in
     the disk body.
"""
import argparse
import ctypes
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "heavy-disks_amd"), os.path.join(ROOT, "oracle")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import diskfft_amd as dfk  # noqa: E402
import oracle  # noqa: E402


def q_scan(w, Nu, Nphi, eps, k, q_in, q_out, mass_ratio):
    """Largest min Q over [q_in, q_out] reachable by any K (gravity of the
    axisymmetric disk from the oracle's potential)."""
    m = dfk.model_defaults(dfk.MODEL_POWERLAW, taper_w=w, q_in=q_in, q_out=q_out, Qmin=1e9,
                           mass_ratio=mass_ratio)
    dfk.model_ic(m, Nu, Nphi, eps, potential_forces=lambda x: k.forces(k.potential(x)))
    return float(m.Qmin_got), float(m.K)


def m2_profile(Sig, sig0):
    """|m = 2 Fourier coefficient of Sigma| / (Nphi/2 * Sigma_0) per radius."""
    F = np.fft.rfft(Sig, axis=1)
    return np.abs(F[:, 2]) / (0.5 * Sig.shape[1] * np.maximum(sig0, 1e-300))


def run_width(w, Nu, Nphi, eps, tlim, whole, qmin=1.3, mass_ratio=1.0):
    k = oracle.Kalnajs(Nu, Nphi, 0.1, 1.0, eps, 1.0, fast=True)
    out = {"taper_w": w, "Nu": Nu, "Nphi": Nphi, "eps": eps, "Qmin": qmin, "mass_ratio": mass_ratio}
    out["q_whole_max"], out["K_whole_best"] = q_scan(w, Nu, Nphi, eps, k, 0.0, 10.0, mass_ratio)
    out["q_body_max"], out["K_body_best"] = q_scan(w, Nu, Nphi, eps, k, 0.2, 0.8, mass_ratio)
    out["q13_whole_reachable"] = out["q_whole_max"] >= qmin
    out["q13_body_reachable"] = out["q_body_max"] >= qmin
    use_whole = whole and out["q13_whole_reachable"]
    m = dfk.model_defaults(dfk.MODEL_POWERLAW, taper_w=w, Qmin=qmin, mass_ratio=mass_ratio,
                           q_in=0.0 if use_whole else 0.2, q_out=10.0 if use_whole else 0.8)
    p, S, U, W, mD, Q = dfk.model_ic(m, Nu, Nphi, eps,
                                     potential_forces=lambda x: k.forces(k.potential(x)))
    out["q_window"] = "whole grid" if use_whole else "[0.2, 0.8]"
    out["K"], out["Qmin_got"] = float(m.K), float(m.Qmin_got)
    R = np.exp(np.log(0.1) + (np.arange(Nu) + 0.5) * np.log(10.0) / Nu)
    out["Q_profile"] = Q.tolist()
    omega = W[:, 0] / R
    out["Omega_profile"] = omega.tolist()
    sig0 = S.mean(axis=1)
    h = oracle.Hydro(Nu, Nphi, p["Rin"], p["Rout"], p["K"], p["gamma"], p["gmstar"], fast=True)
    h.set_state(S, U, W)
    t, tn, st, c2f = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_long(0), ctypes.c_double()
    cap = 1 << 16
    buf = (oracle.Sample * cap)()
    samples, profiles, ptimes = [], [], []
    t0 = time.perf_counter()
    T = 0.0
    while T < tlim - 1e-12:
        T = min(tlim, T + 0.5)
        n = h.L.orc_run_from(h.h, k.k, 1, T, 0.05, 0.4, 0, mD, buf, cap, ctypes.byref(t),
                             ctypes.byref(tn), ctypes.byref(st), ctypes.byref(c2f))
        samples.extend((buf[i].t, buf[i].c0, buf[i].c2, buf[i].ph2) for i in range(min(n, cap)))
        profiles.append(m2_profile(h.Sig.copy(), sig0))
        ptimes.append(t.value)
        if int(T * 2) % 20 == 0:
            print(f"  w={w}: t = {t.value:.1f} after {st.value} steps "
                  f"({time.perf_counter() - t0:.0f} s)", flush=True)
    s = np.array(samples)
    out["steps"] = int(st.value)
    out["wall_s"] = time.perf_counter() - t0
    out["threads"] = oracle.lib(True).orc_get_threads()
    g = dfk.extract_growth(s, lo_factor=10.0, hi_factor=100.0, sat_frac=0.25, peak_window=2.0)
    g_all = dfk.extract_growth(s)
    out["growth_windowed"] = g
    out["growth_default"] = g_all
    # corotation of the first growth and where its amplitude sits at its end
    op = g["omega_p"]
    cr = None
    if op > 0:
        sgn = np.sign(omega - op)
        idx = np.where(np.diff(sgn) != 0)[0]
        cr = float(R[idx[0]]) if len(idx) else None
    out["corotation_R"] = cr
    prof = np.array(profiles)
    kk = int(np.searchsorted(np.array(ptimes), g["t_hi"]))
    kk = min(kk, len(prof) - 1)
    body = (R > 0.12) & (R < 0.95)
    pk = prof[kk] * body
    out["m2_peak_R_end_of_growth"] = float(R[int(np.argmax(pk))])
    out["m2_profile_end_of_growth"] = prof[kk].tolist()
    out["m2_profile_t"] = float(ptimes[kk])
    a = s[:, 2] / s[:, 1]
    out["amp_series"] = {"t": s[::10, 0].tolist(), "c2c0": a[::10].tolist()}
    out["amp_max"] = float(a.max())
    out["amp_final_quarter_mean"] = float(a[int(0.75 * len(a)):].mean())
    out["amp_final_quarter_std"] = float(a[int(0.75 * len(a)):].std())
    print(f"w={w}: Qmax whole {out['q_whole_max']:.3f} body {out['q_body_max']:.3f}; "
          f"K {m.K:.4g} ({out['q_window']}); gamma {g['gamma']:.3f} Omega_p {op:.3f} "
          f"R_c {cr} m2 peak R {out['m2_peak_R_end_of_growth']:.3f}; |C2|/|C0| max "
          f"{out['amp_max']:.3f} late {out['amp_final_quarter_mean']:.3f}; {out['steps']} steps "
          f"{out['wall_s']:.0f} s", flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--widths", default="0.02,0.05,0.1")
    ap.add_argument("--nu", type=int, default=128)
    ap.add_argument("--nphi", type=int, default=128)
    ap.add_argument("--eps", type=float, default=0.01)
    ap.add_argument("--tlim", type=float, default=89.0)
    ap.add_argument("--body-q", action="store_true",
                    help="always take Q_min over the body [0.2, 0.8]")
    ap.add_argument("--qmin", type=float, default=1.3)
    ap.add_argument("--mass-ratio", type=float, default=1.0)
    args = ap.parse_args()
    oracle.set_threads(max(1, min(16, os.cpu_count() or 1)))
    res = []
    for w in [float(x) for x in args.widths.split(",")]:
        res.append(run_width(w, args.nu, args.nphi, args.eps, args.tlim, not args.body_q,
                             args.qmin, args.mass_ratio))
        with open(args.out, "w") as f:
            json.dump(res, f)


if __name__ == "__main__":
    main()
