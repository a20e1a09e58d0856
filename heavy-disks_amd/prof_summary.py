"""Summarise rocprofv3 CSV output for the loop kernels.

    python heavy-disks_amd/prof_summary.py stats DIR            # --kernel-trace --stats run
    python heavy-disks_amd/prof_summary.py pmc DIR [DIR ...]    # --pmc runs (FETCH_SIZE, WRITE_SIZE, ...)

`stats` prints calls and average/min/max duration per kernel from
DIR/*kernel_stats.csv; `pmc` prints the per-dispatch average of every counter
per kernel from DIR/*counter_collection.csv (FETCH_SIZE and WRITE_SIZE are in
KiB per dispatch).
"""
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


def pmc(dirs):
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for d in dirs:
        f = glob.glob(os.path.join(d, "*counter_collection.csv"))[0]
        for r in csv.DictReader(open(f)):
            agg[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    names = sorted({c for v in agg.values() for c in v})
    print("| kernel | dispatches | " + " | ".join(names) + " |")
    print("|---|---|" + "---|" * len(names))
    for k, v in agg.items():
        n = max(len(x) for x in v.values())
        cells = [f"{sum(v[c]) / len(v[c]):.1f}" if c in v else "" for c in names]
        print(f"| {k} | {n} | " + " | ".join(cells) + " |")


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        stats(sys.argv[2])
    else:
        pmc(sys.argv[2:])
