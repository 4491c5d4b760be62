"""The oracle still reproduces the committed golden results of the seeded,
reduced benchmark workloads (tests/golden/configs_golden.json, written by
tests/golden/make_golden_configs.py); the GPU side of the same fixture is in
test_gpu_golden_configs.py."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden_configs as G  # noqa: E402
import oracle_lib as O  # noqa: E402

GOLD = json.load(open(os.path.join(HERE, "golden", "configs_golden.json")))


def test_oracle_reproduces_config1_all_metrics():
    xy = O.gen_uniform(1000, 10**4, 1)
    r0 = O.random_perm(1000, 2)
    for k, m in enumerate(["EUC_2D", "CEIL_2D", "ATT"]):
        assert G.lo_record(xy, r0, metric=k) == GOLD["config1"][m]


def test_oracle_reproduces_config2_small_and_config4_small():
    xy = O.gen_clustered(3000, 10**6, 2)
    assert G.lo_record(xy, O.nn_tour(xy, 0)) == GOLD["config2_small"]
    got = [dict(n=n, length=O.local_search(xy, r0, threads=G.THREADS)[1]) for n, xy, r0 in G.config4_cases()]
    assert got == GOLD["config4_small"]


def test_golden_internal_consistency():
    c3 = GOLD["config3_small"]
    rl = np.array(c3["round_lengths"], np.int64)
    assert len(rl) == 31 and np.all(np.diff(rl) <= 0) and rl[-1] == c3["best_length"]
    for m in ("EUC_2D", "CEIL_2D", "ATT"):
        assert GOLD["config1"][m]["sweeps"] == GOLD["config1"][m]["moves"] + 1
    # CEIL_2D lengths are never below EUC_2D ones on the same tours; ATT is about a third
    assert GOLD["config1"]["CEIL_2D"]["length"] > GOLD["config1"]["EUC_2D"]["length"]
    assert 2.5 < GOLD["config1"]["EUC_2D"]["length"] / GOLD["config1"]["ATT"]["length"] < 3.5
