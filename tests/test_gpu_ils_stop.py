"""The reference ILS's stop rules on the device (ILSGlobalSolverSequential.cpp:
60-100): a target length solution * (1 + err) and a time limit, checked after
the first local search and after every round.  A stopped run must equal the
fixed-round ILS of as many rounds, which the oracle pins (best length, round
lengths, tour)."""
import sys

import numpy as np
import pytest

import logo2opt as L
import oracle_lib as O

pytestmark = pytest.mark.gpu
THREADS = 8


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ends(solution, err, length):
    thr = np.float32(solution) * (np.float32(1.0) + np.float32(err))
    return not (thr < np.float32(length))


@pytest.fixture(scope="module")
def inst():
    n = 1200
    xy = O.gen_uniform(n, 10**4, 77)
    r0 = O.nn_tour(xy, 0)
    return n, xy, r0


def test_target_stops_at_first_round_that_meets_it(inst):
    n, xy, r0 = inst
    say(f"ils_until target n={n}")
    with L.Context(L.coords_of(xy)) as ctx:
        r_full, best_full, rl = ctx.ils(r0, 31, 40)
        # a target met first at some round k (strict improvements make several candidates)
        ks = [k for k in range(1, 41) if rl[k] < rl[k - 1]]
        assert len(ks) >= 3
        for k in ks[1:4]:
            for err in (0.0, 0.004):
                # the smallest solution whose target round k's best meets
                solution = int(rl[k] / (1 + err))
                while not ends(solution, err, int(rl[k])):
                    solution += 1
                expect = next(r for r in range(k + 1) if ends(solution, err, int(rl[r])))
                r, best, rlu, done = ctx.ils_until(r0, 31, 40, solution, err)
                assert done == expect and best == int(rl[expect]), (k, err, done, expect)
                assert list(rlu) == list(rl[:expect + 1])
                rb, Lb, rlo, _ = O.ils(xy, r0, 31, expect, threads=THREADS)
                assert best == Lb and np.array_equal(r, rb)
                assert ctx.route_length(r) == best


def test_target_met_by_first_local_search_and_unreachable(inst):
    n, xy, r0 = inst
    with L.Context(L.coords_of(xy)) as ctx:
        r1, best1, rl = ctx.ils(r0, 5, 12)
        # met after the first LO: no kick at all
        r, best, rlu, done = ctx.ils_until(r0, 5, 12, int(rl[0]), 0.0)
        assert done == 0 and best == int(rl[0]) and len(rlu) == 1
        # unreachable (the reference's default solution = 1): all rounds
        r, best, rlu, done = ctx.ils_until(r0, 5, 12, 1, 0.005)
        assert done == 12 and best == best1 and np.array_equal(r, r1) and list(rlu) == list(rl)


def test_time_limit_completes_the_round_in_progress(inst):
    n, xy, r0 = inst
    with L.Context(L.coords_of(xy)) as ctx:
        r_all, best_all, rl = ctx.ils(r0, 9, 200)
        # a limit that has passed at the first host check: the round in
        # progress then is the last one
        r, best, rlu, done = ctx.ils_until(r0, 9, 200, 1, 0.0, time_limit_s=1e-9)
        say(f"time limit: {done} rounds of 200")
        assert 0 <= done < 200
        assert best == int(rl[done]) and list(rlu) == list(rl[:done + 1])
        rb, Lb, _, _ = O.ils(xy, r0, 9, done, threads=THREADS)
        assert best == Lb and np.array_equal(r, rb)


def test_engine_group_stops_at_target(inst):
    n, xy, r0 = inst
    coords = L.coords_of(xy)
    with L.Context(coords) as ctx:
        _, _, rlg = ctx.ils_multistart(r0, 3, 30, 4)
    k = next(k for k in range(1, 31) if rlg[k] < rlg[k - 1])
    say(f"group target: round {k}")
    with L.Engine(hbm_budget=4 * L.Engine.bytes_for(n, ils_rounds=30) + (16 << 20), max_slots=4) as eng:
        first = eng.add_group(coords, r0, 4, 30, seed=3, stop_below=int(rlg[k]) + 1)
        assert eng.run(sync_every=16) == 0
        res = eng.group_result(first, n)
    assert res["rounds_done"] == k and res["length"] == int(rlg[k])
    rb, Lb, _, _ = O.ils_multistart(xy, r0, 3, k, 4, threads=THREADS)
    assert res["length"] == Lb and np.array_equal(res["route"], rb)
