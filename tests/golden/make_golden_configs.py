"""Writes tests/golden/configs_golden.json: the oracle's results on seeded,
reduced versions of the five benchmark workloads, for the GPU to match bit for
bit and for the oracle itself as a regression pin.

    python tests/golden/make_golden_configs.py

  config1        n = 1000, coordinates U{0..9999}^2 (seed 1), random start
                 (seed 2), local search; also under CEIL_2D and ATT
  config2_small  3,000 clustered cities (seed 2) in [0,1e6)^2, NN start, local search
  config3_small  2,000 uniform cities (seed 3) in [0,1e6)^2, NN start, ILS with 30
                 seeded double-bridge rounds (seed 5)
  config4_small  32 instances, n log-uniform in [100,1000] (numpy seed 4),
                 coordinates U{0..9999}^2, random starts, local searches
  config5_small  8 ILS requests (5 rounds, kick seeds 900+k), n log-uniform in
                 [200,2000], uniform or clustered by a seeded coin, NN starts

Move logs and tours are kept as SHA-256 digests (int32 triples i, J, gain /
uint32 ids, little endian).
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402

THREADS = 8


def digest(a, dt):
    return hashlib.sha256(np.ascontiguousarray(a, dt).tobytes()).hexdigest()


def lo_record(xy, r0, metric=0):
    r, length, moves, sweeps, log = O.local_search(xy, r0, threads=THREADS, metric=metric)
    return dict(length=length, moves=moves, sweeps=sweeps, log_sha256=digest(log, "<i4"),
                tour_sha256=digest(r, "<u4"))


def config4_cases():
    rs = np.random.default_rng(4)
    sizes = np.rint(np.exp(rs.uniform(np.log(100), np.log(1000), 32))).astype(int)
    return [(int(n), O.gen_uniform(int(n), 10**4, 40_000 + k), O.random_perm(int(n), 50_000 + k))
            for k, n in enumerate(sizes)]


def config5_cases():
    rs = np.random.default_rng(5)
    sizes = np.rint(np.exp(rs.uniform(np.log(200), np.log(2000), 8))).astype(int)
    out = []
    for k, n in enumerate(sizes):
        n = int(n)
        xy = O.gen_uniform(n, 10**6, 60_000 + k) if rs.random() < 0.5 else O.gen_clustered(n, 10**6, 60_000 + k)
        out.append((n, xy, O.nn_tour(xy, 0), 900 + k))
    return out


def main():
    res = {}
    xy = O.gen_uniform(1000, 10**4, 1)
    r0 = O.random_perm(1000, 2)
    res["config1"] = {m: lo_record(xy, r0, metric=k) for k, m in enumerate(["EUC_2D", "CEIL_2D", "ATT"])}
    xy = O.gen_clustered(3000, 10**6, 2)
    res["config2_small"] = lo_record(xy, O.nn_tour(xy, 0))
    xy = O.gen_uniform(2000, 10**6, 3)
    rb, Lb, rl, tm = O.ils(xy, O.nn_tour(xy, 0), 5, 30, threads=THREADS)
    res["config3_small"] = dict(best_length=Lb, round_lengths=[int(v) for v in rl], moves=tm,
                                tour_sha256=digest(rb, "<u4"))
    res["config4_small"] = [dict(n=n, length=lo_record(xy, r0)["length"]) for n, xy, r0 in config4_cases()]
    res["config5_small"] = [dict(n=n, seed=seed, best_length=O.ils(xy, r0, seed, 5, threads=THREADS)[1])
                            for n, xy, r0, seed in config5_cases()]
    with open(os.path.join(HERE, "configs_golden.json"), "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res)[:400])


if __name__ == "__main__":
    main()
