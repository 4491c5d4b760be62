"""The 2,048-instance batched workload (n log-uniform in [100, 5000],
coordinates U{0..9999}^2, random starts) run to local optima in one engine,
with its larger instances checked against the oracle: every 8th instance
with n in (1000, 5000] must reach the oracle's local optimum (length, move
count and tour digest committed in tests/golden/batch_golden.json by
tests/golden/make_golden_batch.py), both with full and with pruned sweeps."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import logo2opt as L

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def gold():
    with open(os.path.join(ROOT, "tests", "golden", "batch_golden.json")) as f:
        return json.load(f)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prune", [False, True])
def test_batch_2048_large_instances_equal_oracle(prune):
    GOLD = gold()
    count = GOLD["count"]
    rs = np.random.default_rng(GOLD["seed"])
    sizes = np.rint(np.exp(rs.uniform(np.log(100), np.log(5000), count))).astype(int)
    insts = [(int(n), L.coords_of(L.gen_uniform(int(n), 10**4, 10_000 + k)), L.random_perm(int(n), 20_000 + k))
             for k, n in enumerate(sizes)]
    budget = sum(L.Engine.bytes_for(n) for n, _, _ in insts) + (64 << 20)
    with L.Engine(hbm_budget=budget, max_slots=count) as eng:
        eng.set_pruning(prune)
        slots = [eng.add(c, r) for _, c, r in insts]
        assert eng.run(sync_every=32) == 0
        checked = 0
        for g in GOLD["instances"]:
            k = g["k"]
            assert insts[k][0] == g["n"]
            res = eng.result(slots[k], n=g["n"])
            assert (res["length"], res["moves"]) == (g["length"], g["moves"]), k
            digest = hashlib.sha256(np.ascontiguousarray(res["route"], "<u4").tobytes()).hexdigest()
            assert digest == g["tour_sha256"], k
            checked += 1
    say(f"batch prune={prune}: {checked} instances equal the oracle")
