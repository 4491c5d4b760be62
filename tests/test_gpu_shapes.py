"""Every compiled sweep tile shape against the oracle.

The tuning tables pick a shape per n, so most shapes would otherwise only run
where the tables send them.  Each shape is forced for all n in turn (the
single-instance table) and checked on sizes that exercise one item, many
items, diagonal waves, partial column blocks and shorter tail items: the
selected move of a sweep, and a whole local search's move log.
"""
import os
import sys

import numpy as np
import pytest

import logo2opt as L
import oracle_lib as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _shapes():
    try:
        return L.tuning_shapes()
    except Exception:  # library not built: the test run reports the load error
        return [(1, 32, 1)]


@pytest.fixture
def single_table():
    saved = L.tuning_get(L.TABLE_SINGLE)
    yield
    L.tuning_set(saved, L.TABLE_SINGLE)


@pytest.mark.parametrize("shape", _shapes(), ids=lambda s: "k%d_t%d_w%d" % tuple(s))
def test_shape_matches_oracle(shape, single_table):
    L.tuning_set([(L.MAX_N, tuple(shape))], L.TABLE_SINGLE)
    say("shape", shape)
    for n, side, metric in [(5, 10, 0), (67, 30, 1), (700, 10**4, 2), (2500, 10**6, 0), (4099, 2**22 - 1, 0)]:
        xy = O.gen_uniform(n, side, n + 3)
        with L.Context(L.coords_of(xy), metric=metric) as ctx:
            for seed in range(2):
                r = O.random_perm(n, 90 + seed) if seed == 0 else O.nn_tour(xy, 0)
                ok, best = O.step(xy, r, THREADS, metric=metric)
                assert ctx.step(r) == (best if ok else (0, 0, 0)), (shape, n, seed)
            if n == 700:
                r0 = O.random_perm(n, 5)
                rr, Lo, moves, _, log = O.local_search(xy, r0, threads=THREADS, metric=metric)
                out = ctx.optimize(r0, log_cap=1 << 16)
                np.testing.assert_array_equal(out["log"], log)
                assert out["length"] == Lo
