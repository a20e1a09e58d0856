"""Resolution of the surrogate's training runs: the targets (gamma, Omega_p of
the first m = 2 growth, its saturation peak sat_peak) of held-out parameter
points of the surrogate's draw, recomputed by batched GPU runs on several
square grids (the first is the training grid) and by the CPU oracle on the
training grid from the same parameters.  The runs stop at --tlim (the
training runs' own end, t = 8: the growth of 98% of the draws ends by t = 6).

    python tests/longrun/surrogate_convergence.py --grids 256,384,512 --tlim 8 \
        --out tests/longrun/surrogate_convergence.json
"""
import argparse
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
import ensemble_cfg  # noqa: E402
import oracle  # noqa: E402
import surrogate_data  # noqa: E402

EX = dict(lo_factor=10.0, hi_factor=100.0, sat_frac=0.5, peak_window=2.0)
TABLE = os.path.join(ROOT, "tests", "golden", "surrogate_runs.npz")


def oracle_targets(k, Nu=96, Nphi=96, tlim=8.0, seed=2026, total=4352):
    """Targets of draw k on the oracle, its equilibrium from the oracle's own
    gravity (as the GPU's runs take theirs from the GPU's)."""
    d = ensemble_cfg.draw(total, seed=seed, grids=[(Nu, Nphi)])[k]
    d["seed"] = 42
    m = ensemble_cfg.model_for(dfk, d)
    kal = oracle.Kalnajs(Nu, Nphi, 0.1, 1.0, d["eps"], 1.0, fast=True)
    p, S, U, W, mD, _ = dfk.model_ic(m, Nu, Nphi, d["eps"],
                                     potential_forces=lambda x: kal.forces(kal.potential(x)))
    h = oracle.Hydro(Nu, Nphi, p["Rin"], p["Rout"], p["K"], p["gamma"], p["gmstar"], fast=True)
    h.set_state(S, U, W)
    so, tf, _, steps = oracle.run_chunked(h, kal, 1, tlim, dtout=0.01, cfl=0.4, mD=mD)
    g = dfk.extract_growth(so, **EX)
    return [g["gamma"], g["omega_p"], g["sat_peak"]], int(steps)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--first", type=int, default=4096)
    ap.add_argument("--count", type=int, default=16)
    ap.add_argument("--oracle-count", type=int, default=4)
    ap.add_argument("--tlim", type=float, default=8.0)
    ap.add_argument("--grids", default="256,384,512")
    ap.add_argument("--table", default=TABLE, help="the training table the first grid made")
    args = ap.parse_args()
    grids = [int(x) for x in args.grids.split(",")]
    g0 = str(grids[0])
    oracle.set_threads(max(1, min(16, os.cpu_count() or 1)))
    res = {"first": args.first, "count": args.count, "tlim": args.tlim, "grids": {}}
    for Nu in grids:
        t0 = time.perf_counter()
        r = surrogate_data.run_shard(args.first, args.count, Nu, Nu, tlim=args.tlim, verbose=False)
        res["grids"][str(Nu)] = {"Y": r["Y"].tolist(), "X": r["X"].tolist(),
                                 "steps": r["steps"].tolist(), "wall": time.perf_counter() - t0}
        print(f"{Nu}^2: {time.perf_counter() - t0:.1f} s", flush=True)
    Y0 = np.array(res["grids"][g0]["Y"])
    tab = np.load(args.table)
    ty = tab["Y"][args.first:args.first + args.count]
    res["table_Y"] = ty.tolist()
    res["rel_vs_table"] = np.max(np.abs(Y0 - ty) / np.abs(ty), axis=0).tolist()
    for Nu in grids[1:]:
        Y = np.array(res["grids"][str(Nu)]["Y"])
        rel = np.abs(Y0 - Y) / np.abs(Y)
        res["grids"][str(Nu)]["rel_first_vs_this"] = {"median": np.median(rel, 0).tolist(),
                                                     "max": rel.max(0).tolist()}
    orc = []
    for k in range(args.oracle_count):
        y, steps = oracle_targets(args.first + k, grids[0], grids[0], tlim=args.tlim)
        orc.append({"k": args.first + k, "Y": y, "steps": steps,
                    "gpu_Y": res["grids"][g0]["Y"][k], "gpu_steps": res["grids"][g0]["steps"][k]})
        print(f"oracle run {args.first + k}: {y} vs gpu {res['grids'][g0]['Y'][k]}", flush=True)
    res["oracle_first_grid"] = orc
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)
    for Nu in grids[1:]:
        print(Nu, res["grids"][str(Nu)]["rel_first_vs_this"])
    print(f"{g0}^2 vs table", res["rel_vs_table"])


if __name__ == "__main__":
    main()
