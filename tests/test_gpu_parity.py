"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bit-exact bar: the same selected move (i, J, gain) per sweep, the same applied
move sequence, the same final tour and the same uint64 length as the oracle's
restatement of the reference (cpuSolver.cpp / twoOptLocalSolver.cpp /
ILSGlobalSolverSequential.cpp semantics with exact EUC_2D distances).
"""
import os
import sys
import time

import numpy as np
import pytest

import logo2opt as L
import oracle_lib as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
THREADS = min(16, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ctx_for(xy):
    return L.Context(L.coords_of(xy))


# ---------------------------------------------------------------------------
def test_v_sqrt_f32_accuracy_bound():
    # the sweep's slack assumes <= 3 ulp over the squared lengths it sees
    lo = np.float32(1.0).view(np.uint32).item()
    hi = np.float32(2.0 ** 46).view(np.uint32).item()
    m = L.sqrt_ulp_check(lo, hi)
    say("v_sqrt_f32 max ulp error over [1, 2^46):", m)
    assert m <= 3


@pytest.mark.parametrize("n,side,kind", [
    (4, 100, "u"), (5, 100, "u"), (6, 10, "u"), (7, 10**4, "u"), (10, 10**6, "u"), (50, 30, "u"),
    (100, 10**4, "u"), (257, 10**6, "u"), (1000, 10**4, "u"), (2000, 10**6, "c"), (5000, 10**4, "u"),
    (4099, 10**6, "u"),
])
def test_step_matches_oracle(n, side, kind):
    xy = O.gen_uniform(n, side, n) if kind == "u" else O.gen_clustered(n, side, n)
    say(f"step n={n} side={side} kind={kind}")
    with ctx_for(xy) as ctx:
        for seed in range(4 if n <= 2000 else 2):
            r = O.random_perm(n, 1000 + seed)
            ok, (i, j, g) = O.step(xy, r, THREADS)
            got = ctx.step(r)
            if ok:
                assert got == (i, j, g)
            else:
                assert got == (0, 0, 0)


def test_step_on_nn_tours_many_ties():
    # small side: many equal gains, exercises the lowest-(i,J) tie-break
    for n, side in [(300, 20), (1200, 50)]:
        xy = O.gen_uniform(n, side, 5)
        with ctx_for(xy) as ctx:
            for start in range(3):
                r = O.nn_tour(xy, start)
                ok, (i, j, g) = O.step(xy, r, THREADS)
                assert ctx.step(r) == ((i, j, g) if ok else (0, 0, 0))


@pytest.mark.parametrize("sync_every", [1, 64])
def test_local_search_config1_move_sequence(sync_every):
    # the 1,000-city workload: coords U{0..9999}^2, random permutation start
    n = 1000
    xy = O.gen_uniform(n, 10**4, 1)
    r0 = O.random_perm(n, 2)
    t = time.time()
    r, Lo, moves, sweeps, log = O.local_search(xy, r0, threads=THREADS)
    say(f"oracle LO n=1000: {moves} moves in {time.time() - t:.1f}s")
    with ctx_for(xy) as ctx:
        out = ctx.optimize(r0, log_cap=1 << 16, sync_every=sync_every)
        assert out["moves"] == moves
        np.testing.assert_array_equal(out["log"], log)
        assert out["route"].tolist() == r.tolist()
        assert out["length"] == Lo == O.route_length(xy, out["route"])
        st = ctx.stats()
        assert st["sweeps"] == sweeps and st["moves"] == moves


def test_optimize_move_log_any_cap():
    """logo2opt_optimize fills min(moves, log_cap) entries for any log_cap:
    a short cap gives the first moves, a cap beyond the context's initial
    buffers (2,000,000 entries) makes it grow them, the full log equals the
    oracle's, and log_cap >= 2^32 is rejected."""
    n = 1000
    xy = O.gen_uniform(n, 10**4, 1)
    r0 = O.random_perm(n, 2)
    r, Lo, moves, sweeps, log = O.local_search(xy, r0, threads=THREADS)
    with ctx_for(xy) as ctx:
        short = ctx.optimize(r0, log_cap=10, sync_every=64)
        np.testing.assert_array_equal(short["log"], log[:10])
        full = ctx.optimize(r0, log_cap=2_000_000, sync_every=64)
        assert full["moves"] == moves and full["length"] == Lo
        np.testing.assert_array_equal(full["log"], log)
        again = ctx.optimize(r0, log_cap=1 << 16, sync_every=64)  # after the growth
        np.testing.assert_array_equal(again["log"], log)
        rc = L.load().logo2opt_optimize(ctx.h, np.array(r0, np.uint32), None, None,
                                        L.C.cast((L.Best2Out * 1)(), L.vp), 1 << 32, 64)
        assert rc == -1  # LOGO2OPT_EINVAL


@pytest.mark.parametrize("n,side,kind,start", [
    (2000, 10**6, "c", "nn"), (1500, 10**6, "u", "rand"), (3000, 2**22 - 1, "u", "nn"), (800, 40, "u", "rand"),
])
def test_local_search_matches_oracle(n, side, kind, start):
    xy = O.gen_uniform(n, side, 31) if kind == "u" else O.gen_clustered(n, side, 32)
    r0 = O.nn_tour(xy, 0) if start == "nn" else O.random_perm(n, 33)
    r, Lo, moves, sweeps, log = O.local_search(xy, r0, threads=THREADS)
    say(f"LO n={n} side={side} {kind} {start}: {moves} moves")
    with ctx_for(xy) as ctx:
        out = ctx.optimize(r0, log_cap=1 << 16)
        np.testing.assert_array_equal(out["log"], log)
        assert out["route"].tolist() == r.tolist() and out["length"] == Lo
        say("  fallbacks:", ctx.stats()["fallbacks"])


@pytest.mark.parametrize("name", ["berlin52", "kroA100", "kroE100", "a280", "pr226", "ts225", "rat783", "pcb3038"])
def test_tsplib_local_search(name):
    d = np.load(os.path.join(GOLDEN, "tsplib_int.npz"))
    xy = O.xy_of(d[name + "_x"], d[name + "_y"])
    n = len(d[name + "_x"])
    r0 = O.nn_tour(xy, 0) if n > 1000 else O.random_perm(n, 7)
    r, Lo, moves, sweeps, log = O.local_search(xy, r0, threads=THREADS)
    with ctx_for(xy) as ctx:
        out = ctx.optimize(r0, log_cap=1 << 16)
        np.testing.assert_array_equal(out["log"], log)
        assert out["length"] == Lo >= int(d[name + "_opt"][0])


def test_monalisa_length_known_answer():
    m = np.load(os.path.join(GOLDEN, "monalisa.npz"))
    xy = O.xy_of(m["x"].astype(np.int32), m["y"].astype(np.int32))
    with ctx_for(xy) as ctx:
        assert ctx.route_length(m["tour"].astype(np.uint32)) == 5757191


@pytest.mark.parametrize("kick", [L.KICK_DOUBLE_BRIDGE, L.KICK_REFERENCE])
def test_ils_matches_oracle(kick):
    n = 200
    xy = O.gen_uniform(n, 10**4, 21)
    r0 = O.random_perm(n, 22)
    best, Lb, rl, _ = O.ils(xy, r0, seed=99, rounds=30, kick="double_bridge" if kick == 0 else "reference",
                            threads=THREADS)
    with ctx_for(xy) as ctx:
        r, Lg, rlg = ctx.ils(r0, seed=99, rounds=30, kick=kick)
        assert Lg == Lb
        assert r.tolist() == best.tolist()
        np.testing.assert_array_equal(rlg, rl)


def test_edge_cases():
    # n = 3: no candidate; n = 4: one candidate
    for n in (3, 4, 5):
        xy = O.gen_uniform(n, 100, n)
        with ctx_for(xy) as ctx:
            for seed in range(6):
                r0 = O.random_perm(n, seed)
                ok, best = O.step(xy, r0) if n >= 4 else (False, None)
                assert ctx.step(r0) == (best if ok else (0, 0, 0))
                rr, Lo, moves, _, log = O.local_search(xy, r0)
                out = ctx.optimize(r0, log_cap=64)
                assert out["route"].tolist() == rr.tolist() and out["length"] == Lo
    # all points identical: every gain is 0, nothing moves
    xy = O.xy_of(np.full(50, 7), np.full(50, 9))
    with ctx_for(xy) as ctx:
        out = ctx.optimize(np.arange(50, dtype=np.uint32))
        assert out["moves"] == 0 and out["length"] == 0
    # negative and offset coordinates are translated; distances unchanged
    x = np.array([-5000, 3000, -200, 7000, 123, -4321, 999, 0], np.int32)
    y = np.array([100, -9000, 4000, 5, -77, 3210, -1, 0], np.int32)
    xy = O.xy_of(x, y)
    with L.Context(L.float_coords(x, y)) as ctx:
        r0 = O.random_perm(8, 3)
        rr, Lo, moves, _, log = O.local_search(xy, r0)
        out = ctx.optimize(r0, log_cap=64)
        assert out["route"].tolist() == rr.tolist() and out["length"] == Lo


def test_errors_reported():
    xy = O.gen_uniform(20, 100, 1)
    with ctx_for(xy) as ctx:
        bad = np.arange(20, dtype=np.uint32)
        bad[3] = 4
        with pytest.raises(L.Logo2OptError) as e:
            ctx.optimize(bad)
        assert e.value.code == -3
        # context still usable afterwards
        out = ctx.optimize(O.random_perm(20, 1))
        assert out["length"] == O.local_search(xy, O.random_perm(20, 1))[1]


def test_engine_batch_matches_oracle_and_is_alone_mixed_invariant():
    rs = np.random.default_rng(4)
    sizes = [int(v) for v in np.exp(rs.uniform(np.log(60), np.log(700), 48))]
    insts = []
    with L.Engine(hbm_budget=256 << 20, max_slots=64) as eng:
        for k, n in enumerate(sizes):
            xy = O.gen_uniform(n, 10**4, 500 + k)
            r0 = O.random_perm(n, 600 + k)
            s = eng.add(L.coords_of(xy), r0, log_cap=4096)
            insts.append((s, n, xy, r0))
        assert eng.run(sync_every=16) == 0
        mixed = {}
        for s, n, xy, r0 in insts:
            res = eng.result(s, n, log_cap=4096)
            r, Lo, moves, sweeps, log = O.local_search(xy, r0, threads=THREADS)
            assert res["done"] == 1 and res["length"] == Lo and res["moves"] == moves
            assert res["route"].tolist() == r.tolist()
            np.testing.assert_array_equal(res["log"], log)
            mixed[s] = res
    # the same instances alone, one engine each: identical results
    for s, n, xy, r0 in insts[:6]:
        with L.Engine(hbm_budget=16 << 20, max_slots=1) as eng:
            a = eng.add(L.coords_of(xy), r0)
            eng.run(sync_every=1)
            res = eng.result(a, n)
            assert res["length"] == mixed[s]["length"] and res["route"].tolist() == mixed[s]["route"].tolist()


def test_engine_admission_budget():
    n = 1000
    per = L.Engine.bytes_for(n)
    with L.Engine(hbm_budget=3 * per + 1024, max_slots=8) as eng:
        xy = O.gen_uniform(n, 10**4, 1)
        c = L.coords_of(xy)
        slots = [eng.add(c, O.random_perm(n, k)) for k in range(3)]
        assert all(s is not None for s in slots)
        assert eng.add(c, O.random_perm(n, 9)) is None  # budget exhausted -> EAGAIN
        eng.remove(slots[0])
        s = eng.add(c, O.random_perm(n, 9))
        assert s is not None
        assert eng.run() == 0
        r, Lo, *_ = O.local_search(xy, O.random_perm(n, 9), threads=THREADS)
        assert eng.result(s, n)["length"] == Lo


def test_config2_full_size_local_optimum():
    # the 33,810 clustered cities in [0,10^6)^2, greedy NN start, to a
    # local optimum: the size-independent properties (valid permutation,
    # exact length, the oracle finds no improving move, the first moves
    # agree).  test_gpu_full_size.py checks the whole move log and tour
    # against the oracle's committed result.
    n = 33810
    xy = L.gen_clustered(n, 10**6, 2)
    r0 = L.nn_tour(xy, 0)
    with ctx_for(xy) as ctx:
        t = time.time()
        out = ctx.optimize(r0, log_cap=1 << 16, sync_every=64)
        dt = time.time() - t
        st = ctx.stats()
        say(f"config2 LO: {out['moves']} moves, {st['sweeps']} sweeps, {dt:.2f}s, fallbacks {st['fallbacks']}")
        r = out["route"]
        assert O.lib.orc_check_tour(n, r) == 0
        assert out["length"] == O.route_length(xy, r) == ctx.route_length(r)
        ok, _ = O.step(xy, r, THREADS)
        assert not ok
        L0 = O.route_length(xy, r0)
        assert L0 + int(out["log"][:, 2].sum()) == out["length"]
        cur = r0.copy()
        for q in range(3):
            ok, mv = O.step(xy, cur, THREADS)
            assert ok and tuple(out["log"][q]) == (mv[0], mv[1], mv[2])
            cur = O.apply(cur, mv[0], mv[1])
