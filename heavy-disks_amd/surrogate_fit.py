"""Train the surrogate ensemble on hydro shards, judge it on held-out hydro runs,
scan 10^5 parameter points, and compare with surrogate-free baselines.

    python heavy-disks_amd/surrogate_fit.py shard*.npz [--extend rerun.npz ...] \
        --out surrogate_eval.json [--export tests/golden/surrogate_runs.npz]

The committed table: eight shards of 544 runs at 256 x 256 to t = 8
(surrogate_data.py --first F --count 544 --tlim 8, F = 0, 544, ..., 3808),
and the 73 runs listed in tests/golden/surrogate_extended_runs.txt, which
had not grown and saturated by t = 8, re-run to 20 outer orbits
(surrogate_data.py --indices tests/golden/surrogate_extended_runs.txt) and
merged with --extend.

Runs 0 .. n_train-1 of the shards train the ensemble, runs n_train .. end are
held out.  Error of a prediction p of a hydro value h for output k:
    range-normalised  |p - h| / (P95_k - P5_k of the training targets)
    relative          |p - h| / |h|            (reported where |h| > 5% of the range)
With the fallback, held-out points whose member spread exceeds 5% take the
hydro value (error 0) and count toward the fallback fraction.  Baselines:
nearest neighbour and linear interpolation of the training runs in the
normalised parameter space.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import surrogate  # noqa: E402

NAMES = ["gamma", "omega_p", "sat"]


def load(paths):
    """(X, Y, ok) of shard files in run order, or of one exported table."""
    Xs, Ys, oks = [], [], []
    for p in sorted(paths, key=lambda q: int(np.load(q)["first"])):
        d = np.load(p)
        Xs.append(d["X"])
        Ys.append(d["Y"])
        oks.append(d["ok"])
    return np.concatenate(Xs), np.concatenate(Ys), np.concatenate(oks)


def table(paths, extend=()):
    """(X, Y, ok, tlim_extended, grid) of the shards, the rows of the runs in
    the `extend` files (made with surrogate_data.py --indices) replaced by
    those runs and marked."""
    X, Y, ok = load(paths)
    ext = np.zeros(len(X), np.int8)
    d0 = np.load(sorted(paths)[0])
    grid = d0["grid"] if "grid" in d0.files else np.array([0, 0])
    for p in extend:
        d = np.load(p)
        i = d["index"]
        X[i], Y[i], ok[i] = d["X"], d["Y"], d["ok"]
        ext[i] = 1
    return X, Y, ok, ext, grid


def export(paths, out, extend=()):
    """The runs' parameters and targets only (a few hundred KB): the table the
    surrogate tests train on without regenerating the hydro runs."""
    X, Y, ok, ext, grid = table(paths, extend)
    np.savez_compressed(out, X=X, Y=Y, ok=ok, first=0, grid=grid, tlim_extended=ext)


def errors(pred, true, scale):
    e = np.abs(pred - true) / scale
    rel_mask = np.abs(true) > 0.05 * scale
    rel = np.where(rel_mask, np.abs(pred - true) / np.maximum(np.abs(true), 1e-300), np.nan)
    out = {}
    for k, n in enumerate(NAMES):
        r = rel[:, k][rel_mask[:, k]]
        out[n] = dict(mean_range_err=float(e[:, k].mean()), p90_range_err=float(np.quantile(e[:, k], 0.9)),
                      median_rel_err=float(np.median(r)) if len(r) else None,
                      mean_rel_err=float(r.mean()) if len(r) else None,
                      frac_rel_le_3pct=float(np.mean(r <= 0.03)) if len(r) else None)
    out["all_mean_range_err"] = float(e.mean())
    return out


def baselines(Xtr, Ytr, Xte):
    from scipy.interpolate import LinearNDInterpolator
    from scipy.spatial import cKDTree
    a, b = surrogate.Surrogate.norm_x(Xtr), surrogate.Surrogate.norm_x(Xte)
    nn = Ytr[cKDTree(a).query(b)[1]]
    lin = LinearNDInterpolator(a, Ytr)(b)
    lin = np.where(np.isnan(lin), nn, lin)
    return nn, lin


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("shards", nargs="+")
    ap.add_argument("--out", required=True)
    ap.add_argument("--n-train", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--members", type=int, default=5)
    ap.add_argument("--scan", type=int, default=100000)
    ap.add_argument("--extend", nargs="*", default=[],
                    help="re-run files whose rows replace those of the same runs")
    ap.add_argument("--export", default=None, help="also write the (X, Y, ok) table here")
    a = ap.parse_args()
    if a.export:
        export(a.shards, a.export, a.extend)
    X, Y, ok, _, _ = table(a.shards, a.extend)
    tr = np.arange(len(X)) < a.n_train
    te = ~tr
    tr &= ok & np.all(np.isfinite(Y), axis=1)
    te &= ok & np.all(np.isfinite(Y), axis=1)
    Xtr, Ytr, Xte, Yte = X[tr], Y[tr], X[te], Y[te]
    scale = np.quantile(Ytr, 0.95, axis=0) - np.quantile(Ytr, 0.05, axis=0)
    s = surrogate.Surrogate(members=a.members, seed=1)
    t0 = time.perf_counter()
    hist = s.fit(Xtr, Ytr, epochs=a.epochs, batch=a.batch, lr=a.lr)
    t_train = time.perf_counter() - t0
    mean, spread, fb = s.predict(Xte)
    with_fb = np.where(fb[:, None], Yte, mean)
    nn, lin = baselines(Xtr, Ytr, Xte)
    rng = np.random.default_rng(2026)
    Xs = surrogate.LO + (surrogate.HI - surrogate.LO) * rng.random((a.scan, 4))
    s.predict(Xs[:1000])
    t0 = time.perf_counter()
    ms, ss, fbs = s.predict(Xs)
    t_scan = time.perf_counter() - t0
    res = dict(n_train=int(tr.sum()), n_test=int(te.sum()), n_bad=int((~ok).sum()),
               epochs=a.epochs, batch=a.batch, lr=a.lr, members=a.members,
               train_seconds=t_train, loss_first=float(hist[0]), loss_last=float(hist[-1]),
               loss_hist=hist[:: max(1, len(hist) // 50)].tolist(), target_scale=scale.tolist(),
               surrogate=errors(mean, Yte, scale),
               surrogate_with_fallback=errors(with_fb, Yte, scale),
               fallback_fraction_test=float(fb.mean()),
               nearest_neighbour=errors(nn, Yte, scale), linear_interp=errors(lin, Yte, scale),
               scan=dict(points=a.scan, seconds=t_scan, points_per_s=a.scan / t_scan,
                         fallback_fraction=float(fbs.mean()),
                         pred_mean=ms.mean(0).tolist(), pred_min=ms.min(0).tolist(),
                         pred_max=ms.max(0).tolist()))
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps({k: res[k] for k in ("n_train", "n_test", "loss_first", "loss_last",
                                          "fallback_fraction_test", "train_seconds")}))
    for k in ("surrogate", "surrogate_with_fallback", "nearest_neighbour", "linear_interp"):
        print(k, {n: round(res[k][n]["mean_range_err"], 4) for n in NAMES},
              "all", round(res[k]["all_mean_range_err"], 4))
    print("scan", res["scan"]["points_per_s"], "points/s, fallback", res["scan"]["fallback_fraction"])


if __name__ == "__main__":
    main()
