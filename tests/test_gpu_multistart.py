"""GPU parity of the multi-start ILS with a shared best (the analogue of
ILSGlobalSolverMT::optimize with comm = 1, ILSGlobalSolverMT.cpp:86-195,
made race free by round-synchronised chains) against the oracle's sequential
restatement (orc_ils_multistart), bit for bit: group best length, tour and
per-round group best."""
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


@pytest.mark.parametrize("n,chains,rounds,kick,metric", [
    (120, 1, 6, "double_bridge", 0),
    (300, 4, 10, "double_bridge", 0),
    (500, 8, 6, "reference", 1),
    (257, 3, 8, "double_bridge", 2),
    (2000, 16, 4, "double_bridge", 0),
])
def test_multistart_matches_oracle(n, chains, rounds, kick, metric):
    xy = O.gen_uniform(n, 10**5, n + chains)
    r0 = O.random_perm(n, 7) if n <= 500 else O.nn_tour(xy, 0)
    say(f"multistart n={n} chains={chains} rounds={rounds} kick={kick} metric={metric}")
    rb, Lb, rl, _ = O.ils_multistart(xy, r0, 31, rounds, chains, kick=kick, threads=THREADS, metric=metric)
    k = L.KICK_DOUBLE_BRIDGE if kick == "double_bridge" else L.KICK_REFERENCE
    with L.Context(L.coords_of(xy), metric=metric) as ctx:
        r, best, rlg = ctx.ils_multistart(r0, 31, rounds, chains, kick=k)
        assert best == Lb and r.tolist() == rb.tolist()
        np.testing.assert_array_equal(rlg, rl)
        if chains == 1:  # one chain is the plain ILS
            r1, b1, rl1 = ctx.ils(r0, 31, rounds, kick=k)
            assert b1 == best and r1.tolist() == r.tolist()
        # a second call on the same context reuses its group engine
        r2, best2, _ = ctx.ils_multistart(r0, 31, rounds, chains, kick=k)
        assert best2 == best and r2.tolist() == r.tolist()


def test_engine_groups_beside_plain_instances():
    xy_a = O.gen_uniform(400, 10**4, 1)
    xy_b = O.gen_clustered(900, 10**6, 2)
    ra, rb = O.random_perm(400, 3), O.nn_tour(xy_b, 0)
    with L.Engine(hbm_budget=1 << 28, max_slots=32) as e:
        p0 = e.add(L.coords_of(xy_b), rb, ils_rounds=3, seed=9)
        g1 = e.add_group(L.coords_of(xy_a), ra, chains=5, ils_rounds=7, seed=11)
        g2 = e.add_group(L.coords_of(xy_b), rb, chains=3, ils_rounds=4, seed=12, kick=L.KICK_REFERENCE)
        p1 = e.add(L.coords_of(xy_a), ra)
        assert None not in (p0, g1, g2, p1)
        assert e.run(max_steps=200000, sync_every=16) == 0
        for g, xy, r0, chains, rounds, seed, kick in [(g1, xy_a, ra, 5, 7, 11, "double_bridge"),
                                                      (g2, xy_b, rb, 3, 4, 12, "reference")]:
            rbest, Lb, rl, _ = O.ils_multistart(xy, r0, seed, rounds, chains, kick=kick, threads=THREADS)
            res = e.group_result(g, len(r0))
            assert res["length"] == Lb and res["route"].tolist() == rbest.tolist() and res["done"] == 1
            for c in range(chains):
                np.testing.assert_array_equal(e.round_lengths(g + c, rounds + 1), rl)
        rp, Lp, _, _ = O.ils(xy_b, rb, 9, 3, threads=THREADS)
        assert e.result(p0, n=900)["length"] == Lp
        assert e.result(p1, n=400)["length"] == O.local_search(xy_a, ra, threads=THREADS)[1]


def test_removing_a_chain_retires_its_group():
    xy = O.gen_uniform(600, 10**4, 5)
    r0 = O.random_perm(600, 6)
    with L.Engine(hbm_budget=1 << 27, max_slots=8) as e:
        g = e.add_group(L.coords_of(xy), r0, chains=4, ils_rounds=50, seed=1)
        p = e.add(L.coords_of(xy), r0)
        e.run(max_steps=20, sync_every=4)
        e.remove(g + 2)
        assert e.stats() is not None
        # the group is gone; the plain instance still finishes
        assert e.run(max_steps=100000, sync_every=16) == 0
        assert e.result(p, n=600)["length"] == O.local_search(xy, r0, threads=THREADS)[1]
        with pytest.raises(L.Logo2OptError):
            e.group_result(g, 600)


def test_ils_slots_out_of_size_order_all_finish():
    # regression: the ILS kernel's grid must cover the highest ILS slot even
    # when a lower slot holds a smaller instance (work lists are built
    # largest instance first)
    small, big = O.gen_uniform(100, 10**4, 8), O.gen_uniform(700, 10**4, 9)
    with L.Engine(hbm_budget=1 << 27, max_slots=4) as e:
        s0 = e.add(L.coords_of(small), O.random_perm(100, 1), ils_rounds=3, seed=2)
        s1 = e.add(L.coords_of(big), O.random_perm(700, 1), ils_rounds=3, seed=3)
        assert e.run(max_steps=100000, sync_every=16) == 0
        for s, xy in [(s0, small), (s1, big)]:
            n = len(xy) // 2
            res = e.result(s, n=n)
            assert res["done"] == 1 and res["rounds_done"] == 3
            assert res["length"] == O.ils(xy, O.random_perm(n, 1), 2 if s == s0 else 3, 3, threads=THREADS)[1]
