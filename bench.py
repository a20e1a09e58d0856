#!/usr/bin/env python3
"""Throughput of the diskfft time-step loop on one MI355X.

Workload: one FP64 disk, 256 x 256 (power-law model: r in [0.1, 1], Sigma ~
r^-1.5 with tanh tapers, M_d/M* = 1, K from min Q = 1.3, 1e-3 seeded noise in
m = 1..8, softening eps = 0.01, CFL 0.4), full step = CFL reduction + Kalnajs
FFT gravity + forces + source + radial and azimuthal MUSCL sweeps + floor.
State, spectra and tables are resident in HBM before the timed region; the
timed region is `--steps` loop iterations run as captured HIP graphs.

Prints one JSON line (value = zone-updates/s) with:
  roofline      the dominant kernel's algorithmic bytes per launch / its average
                launch time (HIP events on the library's stream), vs 8 TB/s;
  cpu_baseline  the CPU oracle (C/OpenMP restatement of the reference loop)
                timed on this host on a bounded number of steps of the same disk;
  ensemble      the 64-disk batched run (grids drawn from 128^2 .. 512^2) and
                the same disks stepped one after another.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "heavy-disks_amd"))

import numpy as np  # noqa: E402

import autotune  # noqa: E402

METRIC = "zone-updates/s incl. FFT gravity, 256² single vs 64-disk batch; % HBM peak"
HBM_PEAK = 8.0e12          # B/s, MI355X
REF_PUBLISHED = 6.12e7     # reference README: 256^2, FFTW + OpenMP 8 threads (CPU)
ONCHIP_B_PER_ZONE = 56     # state in and out (48 B) + real K-hat (8 B): a step kept entirely on chip


def b_alg(Nu, Nphi):
    """Algorithmic HBM bytes of one step of one disk: state (Sigma, vR, vphi)
    read and written once (48 B/zone), the Nu x (Nphi/2+1) complex half
    spectrum written and read once in each transform direction (64 B per
    entry), and the real K-hat read once (16 B per half-spectrum entry)."""
    return 48 * Nu * Nphi + 80 * Nu * (Nphi // 2 + 1)


def kernel_bytes(Nu, Nphi):
    """Algorithmic bytes per launch of each loop kernel for one disk."""
    N, S = Nu * Nphi, Nu * (Nphi // 2 + 1)
    return {
        "control": 0,
        "rows_r2c": 8 * N + 16 * S,
        "cols_conv": 16 * S + 8 * (Nu + 1) * (Nphi // 2 + 1) + 16 * S,
        "rows_c2r": 16 * S + 8 * N,
        "update": 56 * N,      # Sigma, vR, vphi, Phi in; Sigma, vR, vphi out
    }


def warm_graphs(e, steps, warmup):
    """Untimed: launches once every graph that `e.advance(steps)` will launch
    from the current loop position (a graph is captured per length and
    starting state buffer on first use), then `warmup` more steps; returns
    with the loop on the same state buffer as at the call (even step
    counts), so the timed advance(steps) finds all its graphs instantiated
    and uploaded.  Returns the number of steps taken."""
    done = 0
    for _ in range(2 if steps % 2 else 1):   # odd: the graphs from both buffers
        e.advance(steps)
        done += steps
    w = warmup + warmup % 2
    if w:
        e.advance(w)
    return done + w


def single_disk(dfk, Nu, Nphi, eps, cfl, steps, warmup, chunk, model=None):
    m = dfk.model_defaults(dfk.MODEL_POWERLAW if model is None else model)
    p, S, U, W, mD, _ = dfk.model_ic(m, Nu, Nphi, eps)
    e = dfk.Ensemble(p)
    # the latency-regime shapes where tuned for this grid, else the
    # throughput ones (a large single disk fills the GPU by itself)
    tuned = autotune.apply(e, regime="latency") or autotune.apply(e, regime="throughput")
    e.set_state(S, U, W)
    e.set_graph_steps(chunk)
    e.run_init(1e30, dtout=0.25, cfl=cfl, mD=mD, sample_cap=chunk + 8)
    pre = warm_graphs(e, steps, warmup)
    t0 = time.perf_counter()
    e.advance(steps)
    wall = time.perf_counter() - t0
    dev_ms = e.last_advance_ms()
    st = e.status()
    assert st["steps"] == pre + steps and not st["flags"], st
    prof = e.profile_step(200)
    shapes = dict(zip(("ti", "tj", "pr", "cb"), e.get_shapes(0)))
    shapes["threads"] = e.get_threads()
    shapes["kind"] = e.get_update_kernel()
    shapes["fft"] = list(e.get_fft_options(0))
    shapes["rows_inplace"] = e.get_row_transform()
    shapes["tuned"] = bool(tuned)
    e.close()
    return dict(p=p, S=S, U=U, W=W, mD=mD, wall=wall, dev_ms=dev_ms, prof=prof, shapes=shapes)


def host_cpus():
    """(nproc, CPUs this process may run on, threads for the CPU baseline):
    the baseline uses every CPU available to the process, capped by
    OMP_NUM_THREADS where the host sets it (the CPU share of one GPU on a
    shared GPU host)."""
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = nproc
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return nproc, avail, min(avail, share) if share > 0 else avail


def cpu_baseline(p, S, U, W, cfl, seconds, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.set_threads(threads)
    Nu, Nphi = p["Nu"], p["Nphi"]
    h = oracle.Hydro(Nu, Nphi, p["Rin"], p["Rout"], p["K"], p["gamma"], p["gmstar"], fast=True)
    k = oracle.Kalnajs(Nu, Nphi, p["Rin"], p["Rout"], p["eps"], p["G"], fast=True)
    h.set_state(S, U, W)
    t0 = time.perf_counter()
    oracle.run(h, k, 1, 1e30, cfl=cfl, max_steps=20, mD=1.0)
    rate = 20 / (time.perf_counter() - t0)
    n = max(20, int(rate * seconds))
    t0 = time.perf_counter()
    _, _, _, done = oracle.run(h, k, 1, 1e30, cfl=cfl, max_steps=n, mD=1.0)
    dt = time.perf_counter() - t0
    return dict(value=Nu * Nphi * done / dt, unit="zone-updates/s", cores=threads, kind="port",
                sample=f"{done} steps of the same {Nu}x{Nphi} disk from the same initial state "
                       f"({dt:.1f} s), oracle built -O3 -march=x86-64-v3, OpenMP {threads} threads")


def ensemble_legs(dfk, steps_batch, steps_seq, chains=0, cols_mode=None, rows_mode=None):
    import ensemble_cfg
    disks = ensemble_cfg.draw(64, seed=2026)
    ens, ics = ensemble_cfg.build(dfk, disks)
    ntuned = autotune.apply(ens, regime="throughput")
    if chains:
        ens.set_chains(chains)
    if rows_mode is not None:   # row-transform mode of the handle, over the tuned one
        ens.set_row_transform(rows_mode)
    if cols_mode is not None:   # column-convolution mode of every disk, over the tuned one
        for d in range(64):
            r, c, _ = ens.get_fft_options(d)
            ens.set_fft_options(r, c, cols_mode, d=d)
    zones = sum(d["Nu"] * d["Nphi"] for d in disks)
    ens.set_graph_steps(50)
    ens.run_init(1e30, cfl=0.4, mD=[ic[3] for ic in ics], sample_cap=64)
    warm_graphs(ens, steps_batch, 0)
    t0 = time.perf_counter()
    ens.advance(steps_batch)
    wall_b = time.perf_counter() - t0
    batched = zones * steps_batch / wall_b
    nchains = ens.get_chains()
    names = ["control", "rows_r2c", "cols_conv", "rows_c2r", "update"]
    kms = dict(zip(names, [float(x) for x in ens.profile_step(50)]))
    kb = {k: sum(kernel_bytes(d["Nu"], d["Nphi"])[k] for d in disks) for k in names}
    kern = {k: dict(ms=kms[k], alg_bytes=kb[k],
                    achieved_GBs=kb[k] / (kms[k] * 1e-3) / 1e9 if kms[k] > 0 else 0.0)
            for k in names}
    ens.close()
    # the same disks stepped one after another, each a plain single-disk run
    seq_time, seq_zones = 0.0, 0
    balg = sum(b_alg(d["Nu"], d["Nphi"]) for d in disks) / zones
    fracs = {"hbm_frac_alg": batched * balg / HBM_PEAK,
             # the same for a design that kept the whole step on chip
             "hbm_frac_onchip_bound": batched * ONCHIP_B_PER_ZONE / HBM_PEAK}
    if steps_seq <= 0:
        return {"disks": 64, "zones": zones, "steps": steps_batch, "graph_chains": nchains,
                "batched_zone_updates_per_s": batched, "kernels": kern, **fracs}
    for d, ic in zip(disks, ics):
        e = dfk.Ensemble(ic[4])
        autotune.apply(e, regime="latency")
        e.set_state(ic[0], ic[1], ic[2])
        e.set_graph_steps(50)
        e.run_init(1e30, cfl=0.4, mD=ic[3], sample_cap=64)
        warm_graphs(e, steps_seq, 0)
        t0 = time.perf_counter()
        e.advance(steps_seq)
        seq_time += time.perf_counter() - t0
        seq_zones += d["Nu"] * d["Nphi"] * steps_seq
        e.close()
    seq = seq_zones / seq_time
    return {"disks": 64, "zones": zones, "steps": steps_batch, "graph_chains": nchains,
            "batched_zone_updates_per_s": batched,
            "sequential_zone_updates_per_s": seq,
            "speedup_vs_sequential": batched / seq, "tuned_disks": ntuned,
            "kernels": kern, **fracs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=250)
    ap.add_argument("--nu", type=int, default=256)
    ap.add_argument("--nphi", type=int, default=256)
    ap.add_argument("--eps", type=float, default=0.01)
    ap.add_argument("--cfl", type=float, default=0.4)
    ap.add_argument("--chunk", type=int, default=250)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only the batched 64-disk leg (no single disk, no sequential leg)")
    ap.add_argument("--chains", type=int, default=0,
                    help="graph chains of the ensemble handle (0: the library default)")
    ap.add_argument("--rows-mode", type=int, default=None,
                    help="row-transform mode of the ensemble handle (dfk_set_row_transform)")
    ap.add_argument("--cols-mode", type=int, default=None,
                    help="column-convolution mode of every ensemble disk (dfk_set_fft_options)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU; multi-GPU runs are not part of it "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    import diskfft_amd as dfk
    if dfk.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_legs(dfk, args.ensemble_steps, 0, args.chains,
                                                     args.cols_mode, args.rows_mode)}), flush=True)
        return
    Nu, Nphi = args.nu, args.nphi
    r = single_disk(dfk, Nu, Nphi, args.eps, args.cfl, args.steps, args.warmup, args.chunk)
    zps = Nu * Nphi * args.steps / r["wall"]
    names = ["control", "rows_r2c", "cols_conv", "rows_c2r", "update"]
    kms = dict(zip(names, [float(x) for x in r["prof"]]))
    kb = kernel_bytes(Nu, Nphi)
    dom = max(kms, key=kms.get)
    ach = kb[dom] / (kms[dom] * 1e-3) / 1e9 if kms[dom] > 0 else 0.0
    # HBM bytes per launch of that kernel from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE runs (calibrated on a copy kernel), when they
    # cover this workload
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_single_{Nu}x{Nphi}.json")
    if os.path.exists(tpath):
        tk = json.load(open(tpath))["kernels"]
        if dom == "update":   # the loop's build: no force arrays, kernel of the tuned kind
            pre = "k_update_stream<false" if r["shapes"].get("kind") == 1 else "k_update<false"
            names = [k for k in tk if k.startswith(pre)]
        elif dom == "cols_conv":  # the column kernel of the tuned mode
            pre = {0: "k_cols_conv", 1: "k_cols_conv_wave", 2: "k_cols_conv_group<",
                   3: "k_cols_conv_reg"}[r["shapes"].get("fft", [0, 0, 0])[2]]
            names = [k for k in tk if k == pre or (pre.endswith("<") and k.startswith(pre))]
        elif r["shapes"].get("rows_inplace"):   # the in-place (1) or register (2) row kernels
            suf = "_group<" if r["shapes"]["rows_inplace"] == 1 else "_reg<"
            names = [k for k in tk if k.startswith("k_" + dom + suf)]
        else:
            names = [k for k in tk if k == "k_" + dom]
        traffic = tk[names[0]]["traffic_bytes"] if names else None
    out = {
        "metric": METRIC,
        "value": zps,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": zps / REF_PUBLISHED,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"single disk {Nu}x{Nphi}, power-law model, eps {args.eps}, "
                               f"CFL {args.cfl}, Kalnajs FFT self-gravity",
                   "Nu": Nu, "Nphi": Nphi, "global_batch": 1, "parallelism": "none"},
        "hbm_frac_step": zps * b_alg(Nu, Nphi) / (Nu * Nphi) / HBM_PEAK,
        "hbm_frac_onchip_bound": zps * ONCHIP_B_PER_ZONE / HBM_PEAK,
        "device_ms_per_step": r["dev_ms"] / args.steps,
        "kernel_ms": kms,
        "shapes": r["shapes"],
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": ach / (HBM_PEAK / 1e9), "traffic": traffic,
                     "alg_bytes_per_launch": kb[dom]},
    }
    # like for like with the published CPU figure: the reference's own disk
    # (LKA98 ring, eps 0.005, CFL 0.3 as `diskdisk 256 256 0.005 16 1 1e-3`)
    ref = single_disk(dfk, 256, 256, 0.005, 0.3, args.steps, args.warmup, args.chunk,
                      model=dfk.MODEL_LKA98)
    lka = 256 * 256 * args.steps / ref["wall"]
    out["reference_disk"] = {"workload": "LKA98 ring 256x256, eps 0.005, CFL 0.3 (the reference's "
                                         "diskdisk default model)", "value": lka,
                             "vs_published_cpu": lka / REF_PUBLISHED}
    if not args.no_cpu:
        nproc, avail, threads = host_cpus()
        out["cpu_baseline"] = cpu_baseline(r["p"], r["S"], r["U"], r["W"], args.cfl,
                                           args.cpu_seconds, threads)
        out["cpu_baseline"].update(nproc=nproc, cpus_available=avail,
                                   omp_num_threads=os.environ.get("OMP_NUM_THREADS"))
        one = cpu_baseline(r["p"], r["S"], r["U"], r["W"], args.cfl, args.cpu_seconds / 2, 1)
        out["cpu_baseline"]["single_thread"] = {"value": one["value"], "sample": one["sample"]}
    if not args.no_ensemble:
        try:
            out["ensemble"] = ensemble_legs(dfk, args.ensemble_steps, 100, args.chains)
        except ImportError:
            out["ensemble"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
