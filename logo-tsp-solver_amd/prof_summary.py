"""Reduces rocprofv3 CSV output to per-kernel figures.

    python logo-tsp-solver_amd/prof_summary.py counters <run_counter_collection.csv> [out.csv]
    python logo-tsp-solver_amd/prof_summary.py trace <run_kernel_trace.csv> [out.csv]

counters: one row per (kernel, counter): dispatches, mean/min/max of the
per-dispatch value.  trace: one row per kernel: dispatches, mean/min/max
duration in ns (from Start/End timestamps).  The inputs are streamed (they
run to tens of MB for ~10^4 dispatches); the output is a few KB.
"""
import csv
import sys
from collections import defaultdict


def _short(name):
    return name.split("(")[0].replace("void ", "").strip()


def _col(fields, *names):
    low = {f.lower().replace("-", "_"): f for f in fields}
    for n in names:
        if n.lower() in low:
            return low[n.lower()]
    raise KeyError(f"none of {names} in {fields}")


def counters(path, out):
    acc = defaultdict(lambda: defaultdict(float))
    per = defaultdict(lambda: [0, 0.0, float("inf"), float("-inf")])
    with open(path, newline="") as f:
        rd = csv.DictReader(f)
        kcol = _col(rd.fieldnames, "kernel_name")
        ccol = _col(rd.fieldnames, "counter_name")
        vcol = _col(rd.fieldnames, "counter_value")
        dcol = _col(rd.fieldnames, "dispatch_id", "correlation_id")
        for row in rd:
            k = (_short(row[kcol]), row[ccol])
            acc[k][row[dcol]] += float(row[vcol])
    for k, d in acc.items():
        vals = list(d.values())
        per[k] = [len(vals), sum(vals) / len(vals), min(vals), max(vals)]
    with open(out, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["kernel", "counter", "dispatches", "mean", "min", "max"])
        for (kn, cn), (c, m, lo, hi) in sorted(per.items()):
            w.writerow([kn, cn, c, f"{m:.6g}", f"{lo:.6g}", f"{hi:.6g}"])


def trace(path, out):
    per = defaultdict(list)
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            per[_short(row["Kernel_Name"])].append(int(row["End_Timestamp"]) - int(row["Start_Timestamp"]))
    with open(out, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["kernel", "dispatches", "mean_ns", "min_ns", "max_ns"])
        for k, v in sorted(per.items()):
            w.writerow([k, len(v), f"{sum(v) / len(v):.1f}", min(v), max(v)])


if __name__ == "__main__":
    mode, src = sys.argv[1], sys.argv[2]
    dst = sys.argv[3] if len(sys.argv) > 3 else src.replace(".csv", "_summary.csv")
    {"counters": counters, "trace": trace}[mode](src, dst)
    print(dst)
