"""Autotuner for the sweep kernel's tile shape, per n bucket.

    python logo-tsp-solver_amd/tune.py [--out logo-tsp-solver_amd/tuning_gfx950.txt]

For each representative city count it times every compiled shape
{k diagonals per lane, t steps, waves per workgroup} with HIP events
(logo2opt_time_sweep) on two tour states -- a random permutation and a greedy
nearest-neighbour tour -- and keeps the fastest.  Bucket edges are the
geometric midpoints between representatives.  The result is a text table
("max_n k t waves" per line) that the library loads on first use from the
file next to liblogo2opt.so (or from $LOGO2OPT_TUNING).
"""
import argparse
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import logo2opt as L  # noqa: E402

DEFAULT_OUT = os.path.join(HERE, "tuning_gfx950.txt")
REPS = (100, 250, 500, 1000, 2000, 4000, 8000, 16000, 33810, 100000)


def time_shape(ctx, tours, shape, sweeps):
    L.tuning_set([(L.MAX_N, shape)])
    return sum(ctx.time_sweep(t, sweeps) for t in tours)


def tune(reps=REPS, seed=1, verbose=True):
    shapes = L.tuning_shapes()
    best = []
    for n in reps:
        side = 10**4 if n <= 5000 else 10**6
        xy = L.gen_uniform(n, side, seed + n)
        tours = [L.random_perm(n, seed), L.nn_tour(xy, 0)]
        ctx = L.Context(L.coords_of(xy))
        sweeps = max(3, min(50, int(2e9 / (n * n))))
        times = {}
        for sh in shapes:
            time_shape(ctx, tours, sh, 1)  # first launch of this shape
            times[sh] = time_shape(ctx, tours, sh, sweeps) / len(tours)
        ctx.close()
        b = min(times, key=times.get)
        best.append((n, b, times[b]))
        if verbose:
            ranked = sorted(times.items(), key=lambda kv: kv[1])[:4]
            print(f"n={n:>6}: " + "  ".join(f"{s}:{ms * 1e3:.1f}us" for s, ms in ranked), flush=True)
    rows = []
    for q, (n, sh, _) in enumerate(best):
        edge = L.MAX_N if q == len(best) - 1 else int(math.sqrt(n * best[q + 1][0]))
        if rows and rows[-1][1] == sh:
            rows[-1] = (edge, sh)
        else:
            rows.append((edge, sh))
    return rows, best


def write_table(rows, path):
    with open(path, "w") as f:
        for mx, (k, t, w) in rows:
            f.write(f"{mx} {k} {t} {w}\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args()
    rows, best = tune()
    write_table(rows, args.out)
    print("table:", rows)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
