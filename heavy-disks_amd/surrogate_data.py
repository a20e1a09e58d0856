"""Training data for the surrogate: batched hydro runs of the power-law disk.

Each run draws (M_d/M*, Q_min, softening s, slope) from the ensemble
distribution (ensemble_cfg.draw).  The 1e-3 seed noise uses one fixed
realization (noise_seed, default 42) for every run, so that the targets are a
function of the four parameters: with per-run seeds the inner-edge growth
rate scatters by about +-10% at fixed parameters, which no surrogate of the
parameters can learn.  Each run builds its equilibrium with this library's
GPU gravity, and advances ~20 outer orbits (t = 20 * 2 pi / sqrt(G (M* + M_d))
at R = 1) in one batched handle with every other run of the shard.  The
extractor (dfk_extract_growth) turns each run's (t, |C0|, |C2|, arg C2)
series into the targets (gamma, Omega_p, saturated |C2|/|C0|).

All runs use one grid (default 256 x 256, the committed table's) so a shard
is a uniform batch; the grid is a numerical choice, not a surrogate input.
In this model the m = 2 amplitude first grows from the 1e-3 noise by a fast
mode of the sharp inner edge (gamma ~ 10, Omega_p ~ 14, within t < 2) and
then wanders nonlinearly.
The targets are therefore: gamma and Omega_p of that first linear growth,
fitted while |C2|/|C0| rises from 10x to 100x its initial value
(dfk_extract_growth_ex, hi_factor = 100), from samples every dtout = 0.01;
sat = the amplitude at which that growth saturates, the largest |C2|/|C0|
within 2 time units of the window start (sat_peak).  The saturation of the
extractor's default definition, the mean |C2|/|C0| over the final quarter of
the run (sat_late), is stored too; it follows the turbulent late state and
scatters by ~20% between neighbouring parameter points, so it is not a
surrogate target.  The stored series keep the first 600 samples and every
25th after.

    python heavy-disks_amd/surrogate_data.py --first 0 --count 544 --tlim 8 --out shard0.npz

(surrogate_fit.py's docstring gives the recipe of the committed table.)
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import autotune  # noqa: E402
import diskfft_amd as dfk  # noqa: E402
import ensemble_cfg  # noqa: E402

ORBITS = 20


def tlim_for(d):
    return ORBITS * 2.0 * np.pi / np.sqrt(1.0 + d["mass_ratio"])


def run_shard(first, count, Nu=256, Nphi=256, seed=2026, total=4352, chunk=2000,
              max_wall=None, verbose=True, dtout=0.01, lo_factor=10.0, hi_factor=100.0,
              sat_frac=0.25, peak_window=2.0, keep_first=600, keep_every=25, noise_seed=42,
              tlim=None, indices=None):
    """Runs first .. first+count-1 of the draw on one Nu x Nphi handle, each to
    its 20 outer orbits (or to `tlim`: gamma, Omega_p and sat_peak are fixed
    by t ~ 6, sat_late needs the whole run).  indices: the runs of the draw to
    make instead of first .. first+count-1."""
    alld = ensemble_cfg.draw(total, seed=seed, grids=[(Nu, Nphi)])
    disks = [alld[i] for i in indices] if indices is not None else alld[first:first + count]
    if noise_seed is not None:
        for d in disks:
            d["seed"] = noise_seed
    t0 = time.perf_counter()
    # one handle; gravity of the axisymmetric disks first (K does not enter it)
    params = [dfk.disk_params(Nu, Nphi, Rin=0.1, Rout=1.0, K=1.0, gmstar=1.0, eps=d["eps"])
              for d in disks]
    ens = dfk.Ensemble(params)
    mDs, Ks, Qgot = [], [], []
    for k, d in enumerate(disks):
        m = ensemble_cfg.model_for(dfk, d)
        sig0 = dfk.model_sigma0(m, Nu, Nphi)
        Sax = np.repeat(sig0[:, None], Nphi, axis=1)
        gR, _ = ens.forces(ens.potential(Sax, d=k), d=k)
        gRdisk = np.cumsum(gR, axis=1)[:, -1] / Nphi
        S, U, W, _ = dfk.model_state(m, Nu, Nphi, gRdisk)
        ens.set_eos(m.K, 2.0, d=k)
        ens.set_state(S, U, W, d=k)
        mDs.append(dfk.model_mass(m, S))
        Ks.append(m.K)
        Qgot.append(m.Qmin_got)
    autotune.apply(ens, regime="throughput")
    t_setup = time.perf_counter() - t0
    tl = [tlim_for(d) if tlim is None else tlim for d in disks]
    ens.set_graph_steps(200)
    ens.run_init(tl, dtout=dtout, cfl=0.4, mD=mDs, stop_on_nonfinite=1, sample_cap=512)
    series = [[] for _ in disks]
    t1 = time.perf_counter()
    done = False
    while not done:
        done = ens.advance(chunk)
        for k in range(len(disks)):
            s = ens.samples(k)
            if len(s):
                series[k].append(s)
        if verbose:
            st = [ens.status(k) for k in range(0, len(disks), max(1, len(disks) // 4))]
            print(f"  shard {first}: {time.perf_counter() - t1:.0f}s, t = "
                  f"{[round(x['t'], 1) for x in st]}", flush=True)
        if max_wall and time.perf_counter() - t1 > max_wall:
            break
    wall = time.perf_counter() - t1
    steps = np.array([ens.status(k)["steps"] for k in range(len(disks))])
    flags = np.array([ens.status(k)["flags"] for k in range(len(disks))])
    tend = np.array([ens.status(k)["t"] for k in range(len(disks))])
    X, Y, late, ok = [], [], [], []
    nmax = max(sum(len(x) for x in s) for s in series)
    nkeep = keep_first + max(0, nmax - keep_first + keep_every - 1) // keep_every
    ser = np.full((len(disks), nkeep, 4), np.nan, dtype=np.float32)
    for k, d in enumerate(disks):
        s = np.concatenate(series[k]) if series[k] else np.zeros((0, 4))
        sk = np.concatenate([s[:keep_first], s[keep_first::keep_every]])
        ser[k, :len(sk)] = sk
        X.append([d["mass_ratio"], d["Qmin"], np.sqrt(d["eps"]), d["slope"]])
        good = len(s) >= 8 and flags[k] == 0 and tend[k] >= tl[k] * (1 - 1e-9)
        g = (dfk.extract_growth(s, lo_factor=lo_factor, hi_factor=hi_factor, sat_frac=sat_frac,
                                peak_window=peak_window)
             if len(s) >= 4 else dict(gamma=np.nan, omega_p=np.nan, sat=np.nan, sat_peak=np.nan))
        Y.append([g["gamma"], g["omega_p"], g["sat_peak"]])
        late.append(g["sat"])
        ok.append(good)
    zones = float(Nu * Nphi) * steps.sum()
    ens.close()
    return dict(X=np.array(X), Y=np.array(Y), sat_late=np.array(late), ok=np.array(ok),
                K=np.array(Ks),
                Qgot=np.array(Qgot), steps=steps, flags=flags, tlim=np.array(tl),
                series=ser, wall=wall, setup=t_setup, zone_updates=zones,
                first=first, grid=np.array([Nu, Nphi]),
                index=np.array(indices if indices is not None else range(first, first + count)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--first", type=int, default=0)
    ap.add_argument("--count", type=int, default=1024)
    ap.add_argument("--total", type=int, default=4352)
    ap.add_argument("--nu", type=int, default=256)
    ap.add_argument("--nphi", type=int, default=256)
    ap.add_argument("--max-wall", type=float, default=None)
    ap.add_argument("--noise-seed", type=int, default=42,
                    help="noise realization of every run (negative: each run's own seed)")
    ap.add_argument("--tlim", type=float, default=None,
                    help="end every run at this time instead of after 20 outer orbits")
    ap.add_argument("--indices", default=None,
                    help="file with the draw indices of the runs to make (one per line)")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    idx = None
    if a.indices:
        idx = [int(x) for x in open(a.indices).read().split()]
    r = run_shard(a.first, a.count, a.nu, a.nphi, total=a.total, max_wall=a.max_wall,
                  noise_seed=None if a.noise_seed < 0 else a.noise_seed, tlim=a.tlim,
                  indices=idx)
    np.savez_compressed(a.out, **r)
    print(f"shard {a.first}+{a.count}: {r['ok'].sum()} good runs, {r['zone_updates']:.3e} "
          f"zone-updates in {r['wall']:.1f}s ({r['zone_updates'] / r['wall']:.3e}/s), setup "
          f"{r['setup']:.1f}s", flush=True)


if __name__ == "__main__":
    main()
