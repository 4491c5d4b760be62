"""Autotuner for the sweep kernel's tile shape, per n bucket.

    python logo-tsp-solver_amd/tune.py [--out-dir logo-tsp-solver_amd]

Writes two tables ("max_n k t waves" per line) that the library loads on first
use from its own directory:
  tuning_gfx950.txt        one instance in flight (logo2opt_create contexts,
                           engines with fewer than 8 live instances): each
                           shape is timed on the steps of a local search from
                           the nearest-neighbour start tour
  tuning_batch_gfx950.txt  batches (8+ live instances): each shape is timed on
                           an engine holding enough random-start instances of
                           that size to fill the GPU, run to their local optima
One table serves both sweep variants, so a shape is scored on the full sweep
plus, where row pruning is in use (batches, and single instances from the
size LOGO2OPT_PRUNE_AUTO prunes), the pruned sweep: the sum of the two times.
A shape is {k columns per lane, t rows, waves per workgroup}.  Bucket edges
are the geometric midpoints between the representative sizes.
"""
import argparse
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import logo2opt as L  # noqa: E402

REPS = (100, 250, 500, 1000, 2000, 4000, 8000, 16000, 33810, 100000)
BATCH_REPS = (100, 250, 500, 1000, 2000, 4000, 8000, 16000)
PRUNE_AUTO_MIN_N = 16000  # LOGO2OPT_PRUNE_AUTO prunes from this size (l2o_engine.hip)
BATCH_STEPS = 3000  # batched steps timed per shape from 8,000 cities


def pairs(n):
    return (n - 2) * (n - 3) // 2


def tune_single(reps=REPS, seed=1, verbose=True):
    """Times every shape on local-search steps from the greedy nearest-neighbour
    start, in a one-instance engine (what logo2opt_optimize runs): the whole
    search up to 40,000 cities, the first 400 steps above; uniform instances
    and, from 16,000 cities, clustered ones.  The search is bit-identical for
    every shape, so all shapes run the same steps; the first 64 (graph capture
    included) are not timed.  Timing single sweeps on fixed tours instead
    ranks shapes differently, because the key published during a search
    tightens the first-tier threshold as the sweep proceeds."""
    best = []
    for n in reps:
        side = 10**4 if n <= 5000 else 10**6
        kinds = ["u", "c"] if n >= 16000 else ["u"]
        cases = []
        for kind in kinds:
            xy = L.gen_uniform(n, side, seed + n) if kind == "u" else L.gen_clustered(n, side, seed + n)
            cases.append((L.coords_of(xy), L.nn_tour(xy, 0)))
        steps = 0xFFFFFFFF if n <= 40000 else 400
        modes = (L.PRUNE_OFF, L.PRUNE_ALL) if n >= PRUNE_AUTO_MIN_N else (L.PRUNE_OFF,)
        times = {}
        for sh in L.tuning_shapes():
            L.tuning_set([(L.MAX_N, sh)], L.TABLE_SINGLE)
            tot = 0.0
            for coords, r0 in cases:
                for mode in modes:
                    with L.Engine(hbm_budget=L.Engine.bytes_for(n) + (16 << 20), max_slots=1) as eng:
                        eng.set_pruning(mode)
                        eng.add(coords, r0)
                        if eng.run(max_steps=64, sync_every=64) == 0:
                            continue  # converged during the warm-up: nothing to time
                        t0 = time.perf_counter()
                        eng.run(max_steps=steps, sync_every=64)
                        tot += time.perf_counter() - t0
            times[sh] = tot
        b = min(times, key=times.get)
        best.append((n, b))
        if verbose:
            ranked = sorted(times.items(), key=lambda kv: kv[1])[:4]
            print(f"single n={n:>6}: " + "  ".join(f"{s}:{t * 1e3:.1f}ms" for s, t in ranked), flush=True)
    return best


def tune_batch(reps=BATCH_REPS, seed=2, verbose=True):
    """Times every shape on whole batched local searches: an engine holding
    enough random-start instances of the size to fill the GPU (about 4e8
    candidates per batched sweep) runs them to their local optima (the first
    8 steps untimed; from 8,000 cities the next BATCH_STEPS steps only), once
    with the full sweep and once pruned, and the sum of the two times decides.
    Timing single sweeps of the random start tours instead (the earlier
    objective) favoured shapes that were slower over whole searches, where
    the tours improve and the filters reject more."""
    best = []
    for n in reps:
        m = max(8, min(1024, int(4e8 / pairs(n))))
        insts = []
        for q in range(m):
            xy = L.gen_uniform(n, 10**4, seed * 100000 + q)
            insts.append((L.coords_of(xy), L.random_perm(n, seed + q)))
        steps = 0xFFFFFFFF if n < 8000 else BATCH_STEPS
        times = {}
        for sh in L.tuning_shapes():
            L.tuning_set([(L.MAX_N, sh)], L.TABLE_BATCH)
            tot = 0.0
            for mode in (L.PRUNE_OFF, L.PRUNE_ALL):
                with L.Engine(hbm_budget=m * L.Engine.bytes_for(n) + (16 << 20), max_slots=m) as eng:
                    eng.set_pruning(mode)
                    for coords, r0 in insts:
                        eng.add(coords, r0)
                    eng.run(max_steps=8, sync_every=8)
                    t0 = time.perf_counter()
                    eng.run(max_steps=steps, sync_every=32)
                    tot += time.perf_counter() - t0
            times[sh] = tot
        b = min(times, key=times.get)
        best.append((n, b))
        if verbose:
            ranked = sorted(times.items(), key=lambda kv: kv[1])[:4]
            print(f"batch  n={n:>6} (x{m}): " + "  ".join(f"{s}:{t * 1e3:.1f}ms" for s, t in ranked), flush=True)
    return best


def table_of(best):
    rows = []
    for q, (n, sh) in enumerate(best):
        edge = L.MAX_N if q == len(best) - 1 else int(math.sqrt(n * best[q + 1][0]))
        if rows and rows[-1][1] == sh:
            rows[-1] = (edge, sh)
        else:
            rows.append((edge, sh))
    return rows


def write_table(rows, path):
    with open(path, "w") as f:
        for mx, (k, t, w) in rows:
            f.write(f"{mx} {k} {t} {w}\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out-dir", default=HERE)
    args = ap.parse_args()
    single_best = tune_single()
    single = table_of(single_best)
    # above the largest batch representative one instance fills the GPU on its
    # own, so the batch table continues with the single-instance choices
    batch = table_of(tune_batch() + [(n, sh) for n, sh in single_best if n > BATCH_REPS[-1]])
    write_table(single, os.path.join(args.out_dir, "tuning_gfx950.txt"))
    write_table(batch, os.path.join(args.out_dir, "tuning_batch_gfx950.txt"))
    print("single:", single)
    print("batch:", batch)


if __name__ == "__main__":
    main()
