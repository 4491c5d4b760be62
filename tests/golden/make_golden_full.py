"""Writes tests/golden/configs_full_golden.json: the oracle's results for two
benchmark workloads at their full sizes, too long to recompute in a test run.

  config2  33,810 clustered cities (60 Gaussian clusters, sigma 2% of the
           side, 10% uniform noise, side 10^6, seed 2), greedy NN tour from
           city 0, best-improvement 2-opt to a local optimum: final length,
           move and sweep counts, SHA-256 of the move log (int32 triples
           i, J, gain, little endian) and of the final tour (uint32 ids).
  config3  100,000 uniform cities (side 10^6, seed 3), greedy NN tour from
           city 0, ILS with 500 seeded double-bridge rounds (seed 5), strict
           accept: length after the first local search, best length after
           every round, total moves, SHA-256 of the best tour.

    python tests/golden/make_golden_full.py [threads]

The loops run on the oracle's exact SIMD sweep (orc_step_fast, the same move
as the plain scan: tests/test_oracle.py pins the two equal).  As a full-size
check of that loop the script first repeats the committed pla33810 local
search of tests/golden/tsplib_lo.json (computed with the plain scan) and
stops if any digest differs.  The configs take about 1 and 30 minutes on 8
threads.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, HERE)

import oracle_lib as O  # noqa: E402


def sha(a, dt):
    return hashlib.sha256(np.ascontiguousarray(a, dt).tobytes()).hexdigest()


def check_pla33810(threads):
    from make_golden_tsplib_lo import load

    with open(os.path.join(HERE, "tsplib_lo.json")) as f:
        want = json.load(f)["pla33810"]
    info, xy = load("pla33810")
    t = time.time()
    r, length, moves, sweeps, log = O.local_search(xy, O.nn_tour(xy, 0), threads=threads, metric=info["metric"],
                                                   log_cap=1 << 20, fast=True)
    got = dict(length=length, moves=moves, sweeps=sweeps, log_sha256=sha(log, "<i4"), tour_sha256=sha(r, "<u4"))
    for k, v in got.items():
        assert want[k] == v, (k, want[k], v)
    print(f"pla33810 fast loop equals the committed plain-scan result ({time.time() - t:.0f}s)", flush=True)


def config2(threads):
    n = 33810
    xy = O.gen_clustered(n, 10**6, 2)
    r0 = O.nn_tour(xy, 0)
    t = time.time()
    r, length, moves, sweeps, log = O.local_search(xy, r0, threads=threads, log_cap=1 << 20, fast=True)
    out = dict(n=n, seed=2, start="nn0", start_length=O.route_length(xy, r0), length=length, moves=moves,
               sweeps=sweeps, log_sha256=sha(log, "<i4"), tour_sha256=sha(r, "<u4"),
               first_moves=log[:8].tolist(), last_moves=log[-8:].tolist())
    print("config2", length, moves, f"{time.time() - t:.0f}s", flush=True)
    return out


def config3(threads, rounds=500):
    n = 100000
    xy = O.gen_uniform(n, 10**6, 3)
    r0 = O.nn_tour(xy, 0)
    t = time.time()
    r, best, rl, tm = O.ils(xy, r0, 5, rounds, threads=threads, fast=True)
    out = dict(n=n, seed=3, start="nn0", ils_seed=5, rounds=rounds, kick="double_bridge",
               start_length=O.route_length(xy, r0), length_after_first_lo=int(rl[0]), best_length=best,
               round_lengths=[int(v) for v in rl], total_moves=tm, tour_sha256=sha(r, "<u4"))
    print("config3", best, tm, f"{time.time() - t:.0f}s", flush=True)
    return out


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    check_pla33810(threads)
    out = {"config2": config2(threads)}
    path = os.path.join(HERE, "configs_full_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    out["config3"] = config3(threads)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
