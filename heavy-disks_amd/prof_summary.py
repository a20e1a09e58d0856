"""Summarise rocprofv3 CSV output for the loop kernels.

    python heavy-disks_amd/prof_summary.py stats DIR            # --kernel-trace --stats run
    python heavy-disks_amd/prof_summary.py pmc DIR [DIR ...] [--largest]   # --pmc runs
    python heavy-disks_amd/prof_summary.py traffic OUT.json WORKLOAD FETCH_DIR WRITE_DIR \
        COPY_FETCH_DIR COPY_WRITE_DIR COPY_MIB [--largest]
    python heavy-disks_amd/prof_summary.py largest DIR          # --kernel-trace run

`largest` (and `traffic ... --largest`) keeps, per kernel, only the
dispatches with the largest grid: the loop's launches over all disks, not
the single-disk calls made while building initial conditions.

`stats` prints calls and average/min/max duration per kernel from
DIR/*kernel_stats.csv; `pmc` prints the per-dispatch average of every counter
per kernel from DIR/*counter_collection.csv (FETCH_SIZE and WRITE_SIZE are in
KiB per dispatch).  `traffic` turns one FETCH_SIZE and one WRITE_SIZE run into
HBM bytes per launch, each counter scaled by what it reads for hbm_copy
(COPY_MIB MiB read and written per launch), and writes OUT.json (read by
bench.py for the roofline's traffic field).
"""
import json
import collections
import csv
import glob
import os
import sys


def short(name):
    base = name.split("(")[0]
    return base.replace("void ", "")


def stats(d):
    f = glob.glob(os.path.join(d, "*kernel_stats.csv"))[0]
    rows = list(csv.DictReader(open(f)))
    print(f"| kernel | calls | avg µs | min µs | max µs | share % |")
    print("|---|---|---|---|---|---|")
    for r in rows:
        print(f"| {short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
              f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | "
              f"{float(r['Percentage']):.1f} |")


def pmc(dirs, largest=False):
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for d in dirs:
        f = glob.glob(os.path.join(d, "*counter_collection.csv"))[0]
        rows = list(csv.DictReader(open(f)))
        big = collections.defaultdict(int)
        for r in rows:
            big[short(r["Kernel_Name"])] = max(big[short(r["Kernel_Name"])], int(r["Grid_Size"]))
        for r in rows:
            k = short(r["Kernel_Name"])
            if not largest or int(r["Grid_Size"]) == big[k]:
                agg[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
    names = sorted({c for v in agg.values() for c in v})
    print("| kernel | dispatches | " + " | ".join(names) + " |")
    print("|---|---|" + "---|" * len(names))
    for k, v in agg.items():
        n = max(len(x) for x in v.values())
        cells = [f"{sum(v[c]) / len(v[c]):.1f}" if c in v else "" for c in names]
        print(f"| {k} | {n} | " + " | ".join(cells) + " |")


def per_kernel(d, counter, largest=False):
    agg = collections.defaultdict(list)
    f = glob.glob(os.path.join(d, "*counter_collection.csv"))[0]
    for r in csv.DictReader(open(f)):
        if r["Counter_Name"] == counter:
            agg[short(r["Kernel_Name"])].append((int(r["Grid_Size"]), float(r["Counter_Value"])))
    out = {}
    for k, v in agg.items():
        g = max(x[0] for x in v) if largest else None
        vals = [x[1] for x in v if g is None or x[0] == g]
        out[k] = sum(vals) / len(vals)
    return out


def largest_stats(d):
    f = glob.glob(os.path.join(d, "*kernel_trace.csv"))[0]
    agg = collections.defaultdict(list)
    for r in csv.DictReader(open(f)):
        g = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        agg[short(r["Kernel_Name"])].append((g, (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3))
    rows = []
    for k, v in agg.items():
        g = max(x[0] for x in v)
        t = [x[1] for x in v if x[0] == g]
        rows.append((sum(t) / len(t), k, len(t), min(t), max(t), g))
    print("| kernel | calls (largest grid) | grid (threads) | avg µs | min µs | max µs |")
    print("|---|---|---|---|---|---|")
    for avg, k, n, lo, hi, g in sorted(rows, reverse=True):
        print(f"| {k} | {n} | {g} | {avg:.2f} | {lo:.2f} | {hi:.2f} |")


def traffic(out, workload, fdir, wdir, cfdir, cwdir, copy_mib, largest=False):
    kib = float(copy_mib) * 1024.0
    cf = next(v for k, v in per_kernel(cfdir, "FETCH_SIZE").items() if k.startswith("copy_kernel"))
    cw = next(v for k, v in per_kernel(cwdir, "WRITE_SIZE").items() if k.startswith("copy_kernel"))
    ff, wf = kib / cf, kib / cw
    fk, wk = per_kernel(fdir, "FETCH_SIZE", largest), per_kernel(wdir, "WRITE_SIZE", largest)
    res = {"workload": workload, "fetch_factor": ff, "write_factor": wf, "kernels": {}}
    for k in sorted(set(fk) & set(wk)):
        if k.startswith("__amd"):
            continue
        res["kernels"][k] = {"fetch_kib": fk[k], "write_kib": wk[k],
                             "traffic_bytes": 1024.0 * (ff * fk[k] + wf * wk[k])}
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
    print(f"copy kernel: FETCH_SIZE {cf:.1f} KiB, WRITE_SIZE {cw:.1f} KiB for {kib:.0f} KiB "
          f"each way -> factors {ff:.3f}, {wf:.3f}")
    for k, v in res["kernels"].items():
        print(f"| {k} | {v['fetch_kib']:.1f} | {v['write_kib']:.1f} | "
              f"{v['traffic_bytes'] / 1e6:.3f} |")


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        stats(sys.argv[2])
    elif sys.argv[1] == "traffic":
        traffic(*sys.argv[2:9], largest="--largest" in sys.argv[9:])
    elif sys.argv[1] == "largest":
        largest_stats(sys.argv[2])
    else:
        pmc([a for a in sys.argv[2:] if a != "--largest"], largest="--largest" in sys.argv[2:])
