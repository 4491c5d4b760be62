"""Writes tests/golden/batch_golden.json: the oracle's local optima for a
deterministic subset of the 2,048-instance batched workload -- every 8th
instance whose size lies in (1000, 5000] -- as instance index, size, final
length, move count and SHA-256 of the final tour (uint32 ids).

The workload is the one bench.py's `batched` leg and
tests/test_gpu_batch_golden.py build: sizes n_k = round(exp(U[ln 100, ln 5000]))
from numpy's default_rng(4), coordinates U{0..9999}^2 with seed 10000 + k,
random start permutation with seed 20000 + k, each run to a 2-opt local
optimum.

    python tests/golden/make_golden_batch.py [threads]

Runs the oracle's local search on its exact SIMD sweep (the same moves as the
plain scan); about two minutes on 8 threads.
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

import oracle_lib as O  # noqa: E402

COUNT, SEED = 2048, 4


def sizes():
    rs = np.random.default_rng(SEED)
    return np.rint(np.exp(rs.uniform(np.log(100), np.log(5000), COUNT))).astype(int)


def subset(sz):
    return [k for k in range(0, COUNT, 8) if 1000 < sz[k] <= 5000]


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    sz = sizes()
    out = {"count": COUNT, "seed": SEED, "instances": []}
    t = time.time()
    for k in subset(sz):
        n = int(sz[k])
        xy = O.gen_uniform(n, 10**4, 10_000 + k)
        r, length, moves, sweeps, _ = O.local_search(xy, O.random_perm(n, 20_000 + k), threads=threads, log_cap=1,
                                                     fast=True)
        out["instances"].append(dict(k=k, n=n, length=length, moves=moves,
                                     tour_sha256=hashlib.sha256(np.ascontiguousarray(r, "<u4").tobytes()).hexdigest()))
    print(f"{len(out['instances'])} instances in {time.time() - t:.0f}s", flush=True)
    with open(os.path.join(HERE, "batch_golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
