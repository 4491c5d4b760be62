"""Seeded random instances through the C ABI against the oracle, bit for bit:
sizes log-uniform in [4, 1500], every metric, uniform or clustered points,
coordinate spans from 10 to 2^22 - 1 (tier-0 shifts 0 to 14), random or
nearest-neighbour starts, row pruning on or off; the first sweep's move, the
whole local search's move log, length and tour.  LOGO2OPT_RANDOM_PARITY_COUNT sets the number of
instances (default 24), LOGO2OPT_RANDOM_PARITY_SEED the seed that draws them (default 2026)."""
import os
import sys

import numpy as np
import pytest

import logo2opt as L
import oracle_lib as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
COUNT = int(os.environ.get("LOGO2OPT_RANDOM_PARITY_COUNT", "24"))
SEED = int(os.environ.get("LOGO2OPT_RANDOM_PARITY_SEED", "2026"))
BASE = 0 if SEED == 2026 else 100003 * SEED  # offset of the point and start-tour seeds


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cases(count, seed=SEED):
    rs = np.random.default_rng(seed)
    for q in range(count):
        n = int(np.rint(np.exp(rs.uniform(np.log(4), np.log(1500)))))
        side = int(rs.choice([10, 100, 1000, 10**4, 1 << 16, 10**6, (1 << 22) - 1]))
        metric = int(rs.integers(0, 3))
        clustered = bool(rs.random() < 0.5) and side >= 10**4
        nn = bool(rs.random() < 0.5)
        prune = bool(rs.random() < 0.5)
        yield q, n, side, metric, clustered, nn, prune


@pytest.mark.parametrize("q,n,side,metric,clustered,nn,prune", list(cases(COUNT)),
                         ids=lambda v: str(v))
def test_random_instance_matches_oracle(q, n, side, metric, clustered, nn, prune):
    say(f"case {q}: n={n} side={side} metric={metric} clustered={clustered} nn={nn} prune={prune}")
    xy = O.gen_clustered(n, side, BASE + 900 + q) if clustered else O.gen_uniform(n, side, BASE + 900 + q)
    r0 = O.nn_tour(xy, 0) if nn else O.random_perm(n, BASE + 700 + q)
    ok, best = O.step(xy, r0, THREADS, metric=metric)
    rr, Lo, moves, _, log = O.local_search(xy, r0, threads=THREADS, metric=metric)
    with L.Context(L.coords_of(xy), metric=metric) as ctx:
        ctx.set_pruning(prune)
        assert ctx.step(r0) == (best if ok else (0, 0, 0))
        out = ctx.optimize(r0, log_cap=1 << 16)
        np.testing.assert_array_equal(out["log"], log)
        assert out["length"] == Lo and out["moves"] == moves
        assert out["route"].tolist() == rr.tolist()
