"""Row pruning (logo2opt_set_pruning and its engine/server forms) on the GPU.

The sweep kernel's pruned variant skips each block of 4 rows in which no
candidate can pass the integer bound, judged by the rows' L1 distance to the
bounding box of a wave's columns.  It must change no result: every compiled
tile shape, with every metric, must give the oracle's selected move of a
sweep, its local-search move log and length, and the multi-start ILS must
give the oracle's group results, bit for bit.  Near-optimal tours (nearest-
neighbour starts, later ILS rounds) are where most blocks are skipped and
where a wrong bound or a stale V after a skipped run would show."""
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


# (n, coordinate side, metric, clustered, start): small and ragged sizes, a
# diagonal-heavy size, and near-optimal starts on clustered points
CASES = [(5, 10, 0, False, "random"), (67, 30, 1, False, "nn"), (700, 10**4, 2, False, "random"),
         (2000, 10**6, 0, True, "nn"), (4099, 2**22 - 1, 0, False, "nn")]
_oracle = {}


def oracle_case(c):
    """The oracle's first move and (for n >= 700) local search, computed once per case."""
    if c not in _oracle:
        n, side, metric, clustered, start = c
        xy = O.gen_clustered(n, side, n + 5) if clustered else O.gen_uniform(n, side, n + 3)
        r0 = O.random_perm(n, 90) if start == "random" else O.nn_tour(xy, 0)
        ok, best = O.step(xy, r0, THREADS, metric=metric)
        lo = O.local_search(xy, r0, threads=THREADS, metric=metric) if 700 <= n <= 2000 else None
        _oracle[c] = (xy, r0, best if ok else (0, 0, 0), lo)
    return _oracle[c]


@pytest.mark.parametrize("shape", _shapes(), ids=lambda s: "k%d_t%d_w%d" % tuple(s))
def test_pruned_shape_matches_oracle(shape, single_table):
    L.tuning_set([(L.MAX_N, tuple(shape))], L.TABLE_SINGLE)
    for c in CASES:
        n, side, metric = c[:3]
        say("pruned shape", shape, "case", c)
        xy, r0, first, lo = oracle_case(c)
        with L.Context(L.coords_of(xy), metric=metric) as ctx:
            ctx.set_pruning(True)
            assert ctx.step(r0) == first, (shape, c)
            if lo is not None:
                rr, Lo, moves, _, log = lo
                out = ctx.optimize(r0, log_cap=1 << 16)
                np.testing.assert_array_equal(out["log"], log)
                assert out["length"] == Lo and out["route"].tolist() == rr.tolist()
                # a converged tour: every block may be skipped, no move
                assert ctx.step(out["route"]) == (0, 0, 0)


@pytest.mark.parametrize("n,chains,rounds,metric", [(300, 4, 10, 0), (1500, 8, 6, 2)])
def test_pruned_multistart_matches_oracle(n, chains, rounds, metric):
    xy = O.gen_uniform(n, 10**5, n + chains)
    r0 = O.nn_tour(xy, 0)
    say(f"pruned multistart n={n} chains={chains} rounds={rounds} metric={metric}")
    rb, Lb, rl, _ = O.ils_multistart(xy, r0, 31, rounds, chains, threads=THREADS, metric=metric)
    with L.Context(L.coords_of(xy), metric=metric) as ctx:
        ctx.set_pruning(True)
        r, best, rlg = ctx.ils_multistart(r0, 31, rounds, chains)
        assert best == Lb and r.tolist() == rb.tolist()
        np.testing.assert_array_equal(rlg, rl)
        # switching it off on the same context (and its group engine) agrees
        ctx.set_pruning(False)
        r2, best2, _ = ctx.ils_multistart(r0, 31, rounds, chains)
        assert best2 == best and r2.tolist() == r.tolist()


def test_pruning_toggles_between_runs_of_one_engine():
    xy = O.gen_clustered(1200, 10**6, 8)
    r0 = O.nn_tour(xy, 0)
    r, Lo, *_ = O.local_search(xy, r0, threads=THREADS)
    with L.Engine(hbm_budget=1 << 27, max_slots=4) as e:
        for prune in (True, False, True):
            e.set_pruning(prune)
            s = e.add(L.coords_of(xy), r0)
            assert e.run(sync_every=8) == 0
            res = e.result(s, len(r0))
            assert res["length"] == Lo and res["route"].tolist() == r.tolist(), prune
            e.remove(s)


def test_auto_mode_prunes_only_large_instances_and_changes_no_result():
    """LOGO2OPT_PRUNE_AUTO launches the pruned sweep for instances of 16,000
    cities or more and the full one below, in the same engine step; every
    result equals the unpruned engine's."""
    sizes = [(17000, True), (3000, False), (700, False)]
    inst = []
    for k, (n, clustered) in enumerate(sizes):
        xy = O.gen_clustered(n, 10**6, 40 + k) if clustered else O.gen_uniform(n, 10**6, 40 + k)
        inst.append((n, L.coords_of(xy), O.nn_tour(xy, 0)))
    results = {}
    for mode in (L.PRUNE_OFF, L.PRUNE_AUTO, L.PRUNE_ALL):
        with L.Engine(hbm_budget=1 << 28, max_slots=8) as e:
            e.set_pruning(mode)
            slots = [e.add(c, r0, ils_rounds=2, seed=5) for n, c, r0 in inst]
            assert e.run(sync_every=16) == 0
            results[mode] = [(e.result(s, n)["length"], e.result(s, n)["route"].tolist())
                             for s, (n, _, _) in zip(slots, inst)]
    assert results[L.PRUNE_AUTO] == results[L.PRUNE_OFF] == results[L.PRUNE_ALL]
    with L.Engine(hbm_budget=1 << 20, max_slots=1) as e:
        with pytest.raises(L.Logo2OptError):
            e.set_pruning(3)
