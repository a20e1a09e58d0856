"""Long runs of the loop on the GPU and on the CPU oracle from the same state;
gamma, Omega_p and saturated |C2|/|C0| of the m = 2 mode from each.

    python tests/longrun/run_growth.py CASE [CASE ...] --out FILE.json

Cases (name: model, grid, eps, tlim or steps, CFL):
  lka128_e<eps>   LKA98 ring 128x128, tlim 9, CFL 0.3 (diskfft/README.md:27-30)
  lka256          LKA98 ring 256x256, eps 0.005, tlim 16 (diskfft/README.md:36-43)
  pl128           power-law model 128x128, eps 0.01, 2000 steps, CFL 0.4
  pl256           power-law model 256x256, eps 0.01, 20 outer orbits (t = 89), CFL 0.4;
                  samples every 0.01, gamma / Omega_p of the first (inner-edge) growth
                  from 10x to 100x the initial |C2|/|C0|, sat over the second half
The oracle runs with OpenMP on this host's cores (<= 16).
--fused-rows THREADS runs the GPU side with the fused row kernels (row tile of
2 rows, THREADS per workgroup) in place of the four-launch step.
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
import oracle  # noqa: E402


def case(name):
    if name.startswith("lka128_e"):
        return dict(model="lka98", N=(128, 128), eps=float(name[8:]), tlim=9.0, steps=0, cfl=0.3)
    return {"lka256": dict(model="lka98", N=(256, 256), eps=0.005, tlim=16.0, steps=0, cfl=0.3),
            "pl128": dict(model="powerlaw", N=(128, 128), eps=0.01, tlim=1e9, steps=2000, cfl=0.4),
            "pl256": dict(model="powerlaw", N=(256, 256), eps=0.01, tlim=89.0, steps=0, cfl=0.4,
                          dtout=0.01, ex=dict(lo_factor=10.0, hi_factor=100.0, sat_frac=0.5)),
            }[name]


def run_case(name, cache_dir=None, skip_gpu=False, fused=0):
    c = case(name)
    Nu, Nphi = c["N"]
    kind = dfk.MODEL_LKA98 if c["model"] == "lka98" else dfk.MODEL_POWERLAW
    m = dfk.model_defaults(kind)
    k = oracle.Kalnajs(Nu, Nphi, m.Rin, m.Rout, c["eps"], m.G, fast=True)
    # the oracle's gravity sets up the equilibrium; both sides start from the same arrays
    p, S, U, W, mD, _ = dfk.model_ic(m, Nu, Nphi, c["eps"],
                                     potential_forces=lambda x: k.forces(k.potential(x)))
    if skip_gpu:
        e = None
    else:
        e = dfk.Ensemble(p)
        e.set_state(S, U, W)
        if fused:
            e.set_shapes(ti=2)
            e.set_fused_rows(fused)
    t0 = time.perf_counter()
    dtout = c.get("dtout", 0.25)
    if not skip_gpu:
        e.set_graph_steps(500)
        ((sg, stg),) = e.run(c["tlim"], dtout=dtout, cfl=c["cfl"], mD=mD, max_steps=c["steps"],
                             chunk=20000)
        print(f"  {name} gpu: {stg['steps']} steps to t = {stg['t']:.2f}", flush=True)
        if e.get_fused_rows() != fused:
            raise RuntimeError(f"{name}: fused row kernels requested {fused}, "
                               f"ran with {e.get_fused_rows()}")
    tg = time.perf_counter() - t0
    cache = os.path.join(cache_dir, f"oracle_{name}.npz") if cache_dir else None
    if cache and os.path.exists(cache):
        # the oracle's run of this case from an earlier invocation (same code,
        # same initial state; any host)
        z = np.load(cache)
        so, tf, c2f, steps, Sig_o = z["samples"], float(z["t"]), float(z["c2f"]), int(z["steps"]), z["Sig"]
        to, nthr = float(z["wall"]), int(z["threads"])
    else:
        h = oracle.Hydro(Nu, Nphi, p["Rin"], p["Rout"], p["K"], p["gamma"], p["gmstar"], fast=True)
        h.set_state(S, U, W)
        t0 = time.perf_counter()
        so, tf, c2f, steps = oracle.run_chunked(
            h, k, 1, c["tlim"], dtout=dtout, cfl=c["cfl"], max_steps=c["steps"], mD=mD,
            chunk=20000,
            progress=lambda t, n: print(f"  {name} oracle: t = {t:.2f} after {n} steps "
                                        f"({time.perf_counter() - t0:.0f} s)", flush=True))
        to, nthr, Sig_o = time.perf_counter() - t0, oracle.lib(True).orc_get_threads(), h.Sig.copy()
        if cache:
            np.savez_compressed(cache, samples=so, t=tf, c2f=c2f, steps=steps, Sig=Sig_o, wall=to,
                                threads=nthr)
    if skip_gpu:
        return dict(case=name, oracle_only=True, steps_oracle=steps, t_oracle=tf)
    ex = c.get("ex", {})
    gg, go = dfk.extract_growth(sg, **ex), dfk.extract_growth(so, **ex)
    rel = {q: abs(gg[q] - go[q]) / abs(go[q]) if go[q] else None
           for q in ("gamma", "omega_p", "sat")}
    Sg, _, _ = e.get_state()
    out = dict(case=name, config=c, K=p["K"], fused_rows=fused, gpu=gg, oracle=go, rel_diff=rel,
               steps_gpu=stg["steps"], steps_oracle=steps, t_gpu=stg["t"], t_oracle=tf,
               c2_final_gpu=stg["c2_final"], c2_final_oracle=c2f,
               mass_drift_gpu=float(sg[-1, 1] - 1.0), mass_drift_oracle=float(so[-1, 1] - 1.0),
               final_sigma_maxrel=float(np.max(np.abs(Sg - Sig_o)) / np.max(np.abs(Sig_o))),
               wall_gpu_s=tg, wall_oracle_s=to, oracle_threads=nthr,
               samples_gpu=sg.tolist(), samples_oracle=so.tolist())
    print(f"{name}: gamma gpu {gg['gamma']:.5f} oracle {go['gamma']:.5f} | Omega_p gpu "
          f"{gg['omega_p']:.5f} oracle {go['omega_p']:.5f} | sat gpu {gg['sat']:.4f} oracle "
          f"{go['sat']:.4f} | steps {stg['steps']}/{steps} | {tg:.1f}s gpu {to:.1f}s cpu", flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cases", nargs="+")
    ap.add_argument("--out", required=True)
    ap.add_argument("--oracle-cache", default=None,
                    help="directory holding (or receiving) the oracle's runs of the cases")
    ap.add_argument("--fused-rows", type=int, default=0,
                    help="GPU side with the fused row kernels of this many threads")
    ap.add_argument("--oracle-only", action="store_true",
                    help="run (and cache) the oracle side only; no GPU needed")
    args = ap.parse_args()
    oracle.set_threads(max(1, min(16, os.cpu_count() or 1)))
    res = []
    for name in args.cases:
        res.append(run_case(name, args.oracle_cache, args.oracle_only, args.fused_rows))
        with open(args.out, "w") as f:
            json.dump(res, f)


if __name__ == "__main__":
    main()
