"""Two benchmark workloads at their full sizes against the oracle's committed
results (tests/golden/configs_full_golden.json, written once by
tests/golden/make_golden_full.py on the oracle's exact SIMD sweep, whose
loop reproduces the plain oracle's committed pla33810 search bit for bit):

  * 33,810 clustered cities, greedy NN start, best-improvement 2-opt to a
    local optimum: the device's length, move count, move log (SHA-256) and
    final tour (SHA-256) equal the oracle's (twoOptLocalSolver.cpp:22-64);
  * 100,000 uniform cities (more than 2^32 candidates per sweep, positions
    past 2^16), NN start, ILS with 500 seeded double-bridge rounds: best
    length after every round and the best tour equal the oracle's
    (ILSGlobalSolverSequential.cpp:60-100); the final tour is also checked
    for the size-independent properties (a permutation, its exact length,
    no improving move left by the oracle's sweep).

Both also run with row pruning on, which must not change a single move.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np
import pytest

import logo2opt as L
import oracle_lib as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
THREADS = min(16, len(os.sched_getaffinity(0)))


def gold():
    with open(os.path.join(ROOT, "tests", "golden", "configs_full_golden.json")) as f:
        return json.load(f)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sha(a, dt):
    return hashlib.sha256(np.ascontiguousarray(a, dt).tobytes()).hexdigest()


@pytest.mark.parametrize("prune", [False, True])
def test_config2_full_local_search_equals_oracle(prune):
    g = gold()["config2"]
    n = g["n"]
    xy = L.gen_clustered(n, 10**6, g["seed"])
    r0 = L.nn_tour(xy, 0)
    assert L.host_route_length(L.coords_of(xy), r0) == g["start_length"]
    with L.Context(L.coords_of(xy)) as ctx:
        ctx.set_pruning(prune)
        t = time.time()
        out = ctx.optimize(r0, log_cap=1 << 20, sync_every=64)
        say(f"config2 n={n} prune={prune}: {out['moves']} moves, {time.time() - t:.2f}s")
    assert out["length"] == g["length"]
    assert out["moves"] == g["moves"]
    assert out["log"][:8].tolist() == g["first_moves"] and out["log"][-8:].tolist() == g["last_moves"]
    assert sha(out["log"], "<i4") == g["log_sha256"]
    assert sha(out["route"], "<u4") == g["tour_sha256"]


@pytest.mark.parametrize("prune", [False, True])
def test_config3_full_ils_equals_oracle(prune):
    g = gold()["config3"]
    n = g["n"]
    xy = L.gen_uniform(n, 10**6, g["seed"])
    r0 = L.nn_tour(xy, 0)
    coords = L.coords_of(xy)
    assert L.host_route_length(coords, r0) == g["start_length"]
    with L.Context(coords) as ctx:
        ctx.set_pruning(prune)
        t = time.time()
        r, best, rl = ctx.ils(r0, seed=g["ils_seed"], rounds=g["rounds"])
        say(f"config3 n={n} prune={prune}: best {best}, {time.time() - t:.2f}s")
    assert int(rl[0]) == g["length_after_first_lo"]
    assert rl.tolist() == g["round_lengths"]
    assert best == g["best_length"]
    assert sha(r, "<u4") == g["tour_sha256"]
    if not prune:
        # size-independent properties of the device's tour
        assert np.array_equal(np.sort(r), np.arange(n, dtype=np.uint32))
        assert O.route_length(np.ascontiguousarray(xy, np.int32), r) == best
        ok, mv = O.step_fast(np.ascontiguousarray(xy, np.int32), r, THREADS)
        assert not ok, mv
