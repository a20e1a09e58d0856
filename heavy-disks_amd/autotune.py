"""Shape autotuner for the loop kernels.

The knobs (see dfk_set_shapes in include/diskfft.h) act on separate kernels:
  (ti, tj)  output tile of the fused update kernel
  pr        complex rows (two real rows each) per workgroup of the row r2c/c2r
  cb        columns per workgroup of the column convolution
so each is tuned on its own kernel's time, measured with HIP events on the
library's stream (dfk_profile_step).  In the latency regime the tuner then
compares the whole iteration (row r2c with the control, convolution, row
c2r, update) against the fused row kernels (dfk_set_fused_rows: convolution
with the control + one kernel per block of ti full rows) over ti and threads
per workgroup, and keeps the faster loop.  Two regimes:
  "latency"     one disk alone (few workgroups; launch/latency bound)
  "throughput"  enough copies of the disk to hold >= 8M zones (bandwidth bound),
                which is what a disk sees inside a large batched ensemble (the
                64-disk ensemble has 7.7M; with only ~2M the tuner preferred
                short update strips for their extra waves, which re-read more
                halo rows once the GPU is full anyway).
Results are cached in a JSON file keyed by (grid, regime, device).

    python heavy-disks_amd/autotune.py 256x256 512x512 --regime throughput
"""
import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# the shipped cache; DFK_TUNED_SHAPES points apply()/tune() at another file
CACHE = os.environ.get("DFK_TUNED_SHAPES", os.path.join(HERE, "tuned_shapes.json"))
LDS_MAX = 160 * 1024    # gfx950: LDS per workgroup (the library checks the device's value)


def candidates(Nu, Nphi):
    M = 2 * Nu
    tiles = [(ti, tj) for ti in (2, 4, 8, 16, 32) for tj in (16, 32, 64, 128, 256)
             if tj <= Nphi and ti <= Nu
             and 8 * (3 * (ti + 4) * (tj + 4) + 3 * (ti + 1) * (tj + 4) + 16 * (ti + 4) + 2 * tj) <= LDS_MAX]
    prs = [p for p in (1, 2, 4, 8, 16) if (2 * p + 1) * Nphi * 16 <= LDS_MAX and 2 * p <= Nu]
    cbs = [c for c in (1, 2, 4, 8, 16) if (2 * c + 1) * M * 16 <= LDS_MAX and c <= Nphi // 2 + 1]
    return tiles, prs, cbs


def _key(Nu, Nphi, regime, device_name):
    return f"{Nu}x{Nphi}:{regime}:{device_name}"


def device_name():
    """The library targets gfx950 only; the cache key keeps the field so tuned
    shapes from another architecture are never applied by mistake."""
    return "gfx950"


def load_cache(path=CACHE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def lookup(Nu, Nphi, regime="throughput", path=CACHE, device="gfx950"):
    """Tuned (ti, tj, pr, cb) for a grid, or None."""
    c = load_cache(path).get(_key(Nu, Nphi, regime, device))
    return None if c is None else (c["ti"], c["tj"], c["pr"], c["cb"])


def apply(ens, path=CACHE, regime="throughput", device="gfx950"):
    """Set every disk of an Ensemble to its cached shapes (where present), and
    the handle's threads per workgroup, update kernel, row transform and
    fused row kernels to the most common tuned choice (votes weighted by
    zones).  A disk tuned for fused rows gets its fused row-block height when
    the handle runs them."""
    cache = load_cache(path)
    n, votes, kinds, rows, fused = 0, {}, {}, {}, {}
    found = [cache.get(_key(*ens.shape(d), regime, device)) for d in range(ens.n)]
    for d, c in enumerate(found):
        if c:
            z = ens.shape(d)[0] * ens.shape(d)[1]
            k = c.get("kind", 0)
            kinds[k] = kinds.get(k, 0) + z
            r = c.get("rows_inplace", 0)
            rows[r] = rows.get(r, 0) + z
            f = c.get("fused", {}).get("threads", 0)
            fused[f] = fused.get(f, 0) + z
    if kinds:
        ens.set_update_kernel(max(kinds, key=kinds.get))
    if rows:
        ens.set_row_transform(max(rows, key=rows.get))
    fthr = max(fused, key=fused.get) if fused else 0
    for d, c in enumerate(found):
        if c:
            ti = c["fused"]["ti"] if fthr and c.get("fused", {}).get("threads") else c["ti"]
            ens.set_shapes(ti, c["tj"], c["pr"], c["cb"], d=d)
            ens.set_fft_options(*c.get("fft", list(c.get("tw_lds", (1, 1))) + [0]), d=d)
            if "threads" in c:
                t = tuple(c["threads"])
                votes[t] = votes.get(t, 0) + ens.shape(d)[0] * ens.shape(d)[1]
            n += 1
    if votes:
        ens.set_threads(*max(votes, key=votes.get))
    if fthr:
        ens.set_fused_rows(fthr)
    return n


def _handle(dfk, Nu, Nphi, copies):
    m = dfk.model_defaults(dfk.MODEL_LKA98)
    sig0 = dfk.model_sigma0(m, Nu, Nphi)
    S, U, W, _ = dfk.model_state(m, Nu, Nphi, None)
    p = dfk.disk_params(Nu, Nphi, eps=0.01)
    e = dfk.Ensemble([p] * copies)
    for d in range(copies):
        e.set_state(S, U, W, d=d)
    mD = dfk.model_mass(m, S)
    del sig0
    return e, mD


def tune(dfk, Nu, Nphi, regime="throughput", nrep=30, verbose=True):
    copies = 1 if regime == "latency" else max(1, (8 << 20) // (Nu * Nphi))
    e, mD = _handle(dfk, Nu, Nphi, copies)
    tiles, prs, cbs = candidates(Nu, Nphi)
    base = e.get_shapes(0)

    def prof(ti, tj, pr, cb):
        e.set_shapes(ti, tj, pr, cb)
        e.run_init(1e30, cfl=0.3, mD=mD)
        return e.profile_step(nrep)

    res = {"tile": {}, "stream": {}, "pr": {}, "cb": {}}
    e.set_update_kernel(0)
    for ti, tj in tiles:
        res["tile"][f"{ti}x{tj}"] = float(prof(ti, tj, base[2], base[3])[4])
    e.set_update_kernel(1)
    for ts in (4, 8, 16, 32, 64, 128):
        if ts <= Nu:
            res["stream"][str(ts)] = float(prof(ts, 0, base[2], base[3])[4])
    e.set_update_kernel(0)
    # row batches with the twiddle table in LDS; column batches with it in LDS
    # (1) or read through the caches (0).  A launch's LDS is the largest any
    # of its disks needs, so the knob only pays where the big transforms use
    # it; for the rows it lost in mixed ensembles and is not tuned.
    # ... and the row transforms cooperative (0), in place (1) or by the
    # two- or three-pass register kernels (2, 3: their row pairs per
    # workgroup are fixed)
    for rip in (0, 1, 2, 3):
        e.set_row_transform(rip)
        if e.get_row_transform() != rip:
            continue                   # this Nphi has no such row transform
        for pr in (prs if rip < 2 else [base[2]]):
            ms = prof(base[0], base[1], pr, base[3])
            res["pr"][f"{pr}:{rip}"] = float(ms[1] + ms[3])
    e.set_row_transform(0)
    # column batches: twiddles in LDS or through the caches; workgroup-
    # cooperative columns (ping-pong buffers), one wave per column in place,
    # the workgroup in place on one column after another, or the two- or
    # three-pass register transforms (their batch and twiddle reads are fixed)
    thr0 = e.get_threads()
    M, Nc = 2 * Nu, Nphi // 2 + 1
    cbs_wave = [c for c in (1, 2, 4, 8, 16) if (c + 1) * M * 16 <= LDS_MAX and c <= Nc]
    cbs_group = [c for c in (1, 2, 4) if c <= Nc]
    for wave in (0, 1, 2, 3, 4):
        for twl in (1, 0):
            e.set_fft_options(1, twl, wave)
            if wave and e.get_fft_options(0)[2] == 0:
                break                  # this length has no in-place / register transform
            if wave >= 3 and twl == 0:
                break                  # the register transforms have no twiddle or batch knob
            for cb in {0: cbs, 1: cbs_wave, 2: cbs_group, 3: [1], 4: [1]}[wave]:
                if wave == 1:
                    e.set_threads(thr0[0], min(256, 64 * cb), thr0[2])
                elif wave == 2:
                    e.set_threads(thr0[0], 128, thr0[2])
                res["cb"][f"{cb}:{twl}:{wave}"] = float(prof(base[0], base[1], base[2], cb)[2])
            e.set_threads(*thr0)
    bt = min(res["tile"], key=res["tile"].get)
    bst = min(res["stream"], key=res["stream"].get) if res["stream"] else None
    (bpr, rip), twr = (int(x) for x in min(res["pr"], key=res["pr"].get).split(":")), 1
    e.set_row_transform(rip)
    bcb, twc, cwave = (int(x) for x in min(res["cb"], key=res["cb"].get).split(":"))
    e.set_fft_options(twr, twc, cwave)
    ti, tj = (int(x) for x in bt.split("x"))
    kind = 0
    if bst is not None and res["stream"][bst] < res["tile"][bt]:
        kind, ti, tj = 1, int(bst), 58   # the streaming kernel's output columns per wave
    e.set_update_kernel(kind)
    # threads per workgroup of each kernel, at the chosen shapes
    res["threads"] = {}
    for t in (64, 128, 256):
        e.set_threads(t, t, t)
        res["threads"][str(t)] = [float(x) for x in prof(ti, tj, bpr, bcb)]
    # the row r2c and c2r share one thread count: pick it on their sum
    pick = lambda *idx: int(min(res["threads"], key=lambda k: sum(res["threads"][k][i] for i in idx)))  # noqa: E731
    thr = (pick(1, 3), pick(2), pick(4))
    e.set_threads(*thr)
    best = prof(ti, tj, bpr, bcb)
    # latency regime: the four-launch iteration against fused row kernels
    fused = {}
    if regime == "latency":
        step4 = float(sum(best[1:]))       # the control rides in the row r2c launch
        res["fused"] = {}
        if e.get_row_transform():
            e.set_update_kernel(0)
            for fti in (2, 4, 6, 8):
                for fth in (256, 512, 1024):
                    e.set_shapes(fti, tj if kind == 0 else 0, bpr, bcb)
                    e.set_fused_rows(fth)
                    if e.get_fused_rows():
                        ms = prof(fti, 0, bpr, bcb)
                        res["fused"][f"{fti}:{fth}"] = float(ms[2] + ms[4])
            e.set_fused_rows(0)
            e.set_update_kernel(kind)
        if res["fused"]:
            bf = min(res["fused"], key=res["fused"].get)
            if res["fused"][bf] < step4:
                fti, fth = (int(x) for x in bf.split(":"))
                fused = {"threads": fth, "ti": fti, "step_ms": res["fused"][bf], "step4_ms": step4}
    e.close()
    out = {"kind": kind, "ti": ti, "tj": tj, "pr": bpr, "cb": bcb, "threads": list(thr),
           "fft": [twr, twc, cwave], "rows_inplace": rip, "fused": fused,
           "copies": copies,
           "ms_per_kernel": [float(x) for x in best],
           "default": {"shape": list(base), "tile_ms": res["tile"].get(f"{base[0]}x{base[1]}")},
           "sweep": res}
    if verbose:
        print(f"{Nu}x{Nphi} {regime} x{copies}: kernel {kind} tile {ti}x{tj} pr {bpr} rows in place {rip} cb {bcb} "
              f"twiddles in LDS (rows, cols) {(twr, twc)} wave columns {cwave} threads {thr} "
              f"ms {np.round(best, 4).tolist()} fused rows {fused or 'no'}", flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("grids", nargs="+", help="NuxNphi")
    ap.add_argument("--regime", default="throughput", choices=["latency", "throughput"])
    ap.add_argument("--cache", default=CACHE)
    ap.add_argument("--nrep", type=int, default=30)
    args = ap.parse_args()
    sys.path.insert(0, HERE)
    import diskfft_amd as dfk
    cache = load_cache(args.cache)
    dev = device_name()
    for g in args.grids:
        Nu, Nphi = (int(x) for x in g.lower().split("x"))
        cache[_key(Nu, Nphi, args.regime, dev)] = tune(dfk, Nu, Nphi, args.regime, args.nrep)
        with open(args.cache, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
