"""The device against the committed golden results of the seeded, reduced
benchmark workloads (tests/golden/configs_golden.json, the oracle's), bit for
bit: move-log and tour digests, lengths, ILS round lengths, batched and
served results -- with the sweep's row pruning off and on (pruning must not
change any of them)."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import logo2opt as L

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden_configs as G  # noqa: E402
import oracle_lib as O  # noqa: E402

pytestmark = pytest.mark.gpu
PRUNE = pytest.mark.parametrize("prune", [False, True], ids=["full", "pruned"])
GOLD = json.load(open(os.path.join(HERE, "golden", "configs_golden.json")))


def digest(a, dt):
    return hashlib.sha256(np.ascontiguousarray(a, dt).tobytes()).hexdigest()


def check_lo(out, g):
    assert (out["length"], out["moves"]) == (g["length"], g["moves"])
    assert digest(out["log"], "<i4") == g["log_sha256"]
    assert digest(out["route"], "<u4") == g["tour_sha256"]


@PRUNE
def test_config1_all_metrics(prune):
    xy = O.gen_uniform(1000, 10**4, 1)
    r0 = O.random_perm(1000, 2)
    for k, m in enumerate(["EUC_2D", "CEIL_2D", "ATT"]):
        with L.Context(L.coords_of(xy), metric=k) as ctx:
            ctx.set_pruning(prune)
            check_lo(ctx.optimize(r0, log_cap=1 << 16), GOLD["config1"][m])


@PRUNE
def test_config2_small_and_config3_small(prune):
    xy = O.gen_clustered(3000, 10**6, 2)
    with L.Context(L.coords_of(xy)) as ctx:
        ctx.set_pruning(prune)
        check_lo(ctx.optimize(O.nn_tour(xy, 0), log_cap=1 << 16), GOLD["config2_small"])
    xy = O.gen_uniform(2000, 10**6, 3)
    g = GOLD["config3_small"]
    with L.Context(L.coords_of(xy)) as ctx:
        ctx.set_pruning(prune)
        r, best, rl = ctx.ils(O.nn_tour(xy, 0), 5, 30)
    assert best == g["best_length"] and rl.tolist() == g["round_lengths"]
    assert digest(r, "<u4") == g["tour_sha256"]


@PRUNE
def test_config4_small_batched(prune):
    cases = G.config4_cases()
    with L.Engine(hbm_budget=64 << 20, max_slots=64) as e:
        e.set_pruning(prune)
        slots = [e.add(L.coords_of(xy), r0) for n, xy, r0 in cases]
        assert e.run(sync_every=16) == 0
        got = [dict(n=n, length=e.result(s)["length"]) for s, (n, _, _) in zip(slots, cases)]
    assert got == GOLD["config4_small"]


@PRUNE
def test_config5_small_served(prune):
    cases = G.config5_cases()
    with L.Server(hbm_budget=256 << 20, max_slots=16) as srv:
        srv.set_pruning(prune)
        ids = [srv.submit(L.coords_of(xy), r0, ils_rounds=5, seed=seed) for n, xy, r0, seed in cases]
        for _ in range(100000):
            running, queued = srv.pump(steps=64, sync_every=16)
            if running == 0 and queued == 0:
                break
        got = [dict(n=n, seed=seed, best_length=srv.result(rid, n)["length"])
               for rid, (n, _, _, seed) in zip(ids, cases)]
    assert got == GOLD["config5_small"]
