"""Pins the CPU oracle (oracle/logo_oracle.c) before anything is compared
against it.  Each check is chosen so that a plausible mistake in the oracle (a
wrong rounding, a dropped or swapped gain term, an off-by-one in the
neighbourhood, a wrong apply layout, a tie-break in the wrong direction) fails
at least one of them.  No GPU is needed."""
import ctypes
import itertools
import math
import os

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# --------------------------------------------------------------------------
# independent pure-Python model (small n only)
# --------------------------------------------------------------------------
def py_nint(s):
    # floor(sqrt(s) + 1/2) via isqrt(4s): a different route from the oracle
    return (math.isqrt(4 * s) + 1) // 2


def py_dist(xy, a, b):
    dx = int(xy[2 * a]) - int(xy[2 * b])
    dy = int(xy[2 * a + 1]) - int(xy[2 * b + 1])
    return py_nint(dx * dx + dy * dy)


def py_step(xy, r):
    """Enumerate every candidate, collect (gain, i, J), take the minimum
    tuple: the lexicographic (gain, i, J) order is the tie-break
    cpuSolver.cpp:140-145 produces with its strict '<' scan."""
    n = len(r)
    cands = []
    for i in range(1, n - 2):
        for J in range(i + 2, n):
            g = (py_dist(xy, r[i], r[J]) + py_dist(xy, r[i - 1], r[J - 1])
                 - py_dist(xy, r[i - 1], r[i]) - py_dist(xy, r[J - 1], r[J]))
            cands.append((g, i, J))
    best = min(cands) if cands else (0, 0, 0)
    return best if best[0] < 0 else None


def py_rotate_reverse(r, i, J):
    """std::rotate(begin, begin+i, end); std::reverse(begin+(J-i), end)."""
    r = list(r)
    r = r[i:] + r[:i]
    r[J - i:] = r[J - i:][::-1]
    return r


# --------------------------------------------------------------------------
# P1: exact EUC_2D rounding
# --------------------------------------------------------------------------
def test_nint_exhaustive_below_2_24():
    # nint(sqrt(s)) == k exactly on the integer interval [k^2-k+1, k^2+k]
    # (and s = 0 -> 0): build the expected sequence from the interval
    # structure alone and compare over every s < 2^24.
    top = 1 << 24
    ks = np.arange(0, 4200, dtype=np.int64)
    counts = np.where(ks == 0, 1, 2 * ks)
    expected = np.repeat(ks, counts)[:top]
    assert len(expected) == top
    got = O.nint_sqrt(np.arange(top, dtype=np.uint64))
    np.testing.assert_array_equal(got, expected.astype(np.uint32))


def test_nint_golden_isqrt():
    g = np.load(os.path.join(GOLDEN, "nint_golden.npz"))
    got = O.nint_sqrt(g["s"].astype(np.uint64))
    np.testing.assert_array_equal(got.astype(np.int64), g["nint"])


def test_nint_half_integer_neighbourhood():
    # s = k^2 + k is the largest s rounding down to k; k^2 + k + 1 rounds up
    k = np.arange(1, 3_000_000, 997, dtype=np.uint64)
    np.testing.assert_array_equal(O.nint_sqrt(k * k + k), k.astype(np.uint32))
    np.testing.assert_array_equal(O.nint_sqrt(k * k + k + 1), (k + 1).astype(np.uint32))


# --------------------------------------------------------------------------
# P5: tour length
# --------------------------------------------------------------------------
def test_monalisa_known_answer():
    # TSPLIB/monalisa_5757191.tour states its own length (NAME header, file name)
    m = np.load(os.path.join(GOLDEN, "monalisa.npz"))
    xy = O.xy_of(m["x"].astype(np.int32), m["y"].astype(np.int32))
    tour = m["tour"].astype(np.uint32)
    assert O.lib.orc_check_tour(len(tour), tour) == 0
    L = O.route_length(xy, tour)
    assert L == int(m["stated_length"][0]) == 5757191
    assert L >= int(m["known_opt"][0])


@pytest.mark.parametrize("k,s", [(2, 1), (4, 7), (10, 3), (32, 100), (100, 31)])
def test_boustrophedon_grid_closed_form(k, s):
    # k x k grid with spacing s, k even: snake over columns 1..k-1 row by row,
    # then return down column 0; every edge has length s, so L = k^2 * s
    pts = []
    for row in range(k):
        cols = range(1, k) if row % 2 == 0 else range(k - 1, 0, -1)
        for c in cols:
            pts.append((c, row))
    pts += [(0, row) for row in range(k - 1, -1, -1)]
    xy = O.xy_of([p[0] * s for p in pts], [p[1] * s for p in pts])
    route = np.arange(k * k, dtype=np.uint32)
    assert O.route_length(xy, route) == k * k * s


def test_length_matches_python_model():
    for seed in range(5):
        xy = O.gen_uniform(37, 10**6, seed)
        r = O.random_perm(37, seed + 100)
        want = sum(py_dist(xy, int(r[k - 1]), int(r[k])) for k in range(1, 37)) + py_dist(xy, int(r[-1]), int(r[0]))
        assert O.route_length(xy, r) == want


# --------------------------------------------------------------------------
# P2/P3: sweep and selection
# --------------------------------------------------------------------------
@pytest.mark.parametrize("n", [4, 5, 6, 7])
def test_step_exhaustive_all_tours(n):
    xy = O.gen_uniform(n, 50, 1234 + n)  # small side -> many ties
    for perm in itertools.permutations(range(n)):
        r = np.array(perm, np.uint32)
        ok, (i, J, g) = O.step(xy, r)
        want = py_step(xy, [int(v) for v in r])
        if want is None:
            assert not ok and g == 0
        else:
            assert ok and (g, i, J) == want


@pytest.mark.parametrize("seed", range(12))
def test_step_brute_force_random(seed):
    n = [8, 9, 10, 13, 16, 23, 31, 40, 50, 64, 64, 64][seed]
    side = [20, 10**4, 10**6][seed % 3]
    xy = O.gen_uniform(n, side, seed)
    r = O.random_perm(n, 77 + seed)
    ok, (i, J, g) = O.step(xy, r)
    want = py_step(xy, [int(v) for v in r])
    if want is None:
        assert not ok and g == 0
    else:
        assert ok and (g, i, J) == want


def test_step_constructed_ties_pick_lowest():
    # collinear points with repeated spacings create many equal gains
    n = 24
    xs = [(k * 7) % 13 * 10 for k in range(n)]
    xy = O.xy_of(xs, [0] * n)
    for seed in range(20):
        r = O.random_perm(n, seed)
        ok, (i, J, g) = O.step(xy, r)
        want = py_step(xy, [int(v) for v in r])
        assert want is not None and (g, i, J) == want


def test_step_thread_count_invariance():
    xy = O.gen_uniform(300, 10**4, 5)
    r = O.random_perm(300, 6)
    base = O.step(xy, r, 1)
    for t in (2, 3, 7, 8):
        assert O.step(xy, r, t) == base


# --------------------------------------------------------------------------
# P4: apply
# --------------------------------------------------------------------------
def test_apply_matches_rotate_reverse():
    rs = np.random.default_rng(3)
    for n in (4, 5, 9, 30):
        base = list(range(n))
        for i in range(1, n - 2):
            for J in range(i + 2, n):
                r = rs.permutation(n).astype(np.uint32)
                got = O.apply(r, i, J)
                assert got.tolist() == py_rotate_reverse(r.tolist(), i, J)
                # order of (i,J) arguments does not matter (swapint)
                assert O.apply(r, J, i).tolist() == got.tolist()
        assert base == list(range(n))


def test_rotate_reverse_general():
    rs = np.random.default_rng(4)
    for _ in range(200):
        n = int(rs.integers(3, 40))
        a, b = sorted(rs.integers(0, n, 2).tolist())
        r = rs.permutation(n).astype(np.uint32)
        assert O.rotate_reverse(r, a, b).tolist() == py_rotate_reverse(r.tolist(), a, b)


# --------------------------------------------------------------------------
# P6: local search invariants
# --------------------------------------------------------------------------
def test_local_search_invariants_config1_shape():
    n = 1000
    xy = O.gen_uniform(n, 10**4, 1)
    r0 = O.random_perm(n, 2)
    L0 = O.route_length(xy, r0)
    r, L, moves, sweeps, log = O.local_search(xy, r0, threads=8)
    assert O.lib.orc_check_tour(n, r) == 0
    assert sweeps == moves + 1
    assert L == O.route_length(xy, r)
    assert L0 + int(log[:, 2].sum()) == L
    assert (log[:, 2] < 0).all()
    assert (log[:, 0] >= 1).all() and (log[:, 1] >= log[:, 0] + 2).all() and (log[:, 1] <= n - 1).all()
    ok, _ = O.step(xy, r, 8)
    assert not ok  # a 2-opt local optimum


def test_local_search_replay_matches_step_apply():
    n = 120
    xy = O.gen_uniform(n, 10**6, 9)
    r0 = O.random_perm(n, 10)
    r, L, moves, sweeps, log = O.local_search(xy, r0)
    cur = r0.copy()
    for (i, J, g) in log:
        ok, (bi, bj, bg) = O.step(xy, cur)
        assert ok and (bi, bj, bg) == (i, J, g)
        cur = O.apply(cur, int(i), int(J))
    assert cur.tolist() == r.tolist()


@pytest.mark.parametrize("name", ["berlin52", "kroA100", "kroC100", "a280", "pr226", "rat783"])
def test_tsplib_local_optimum_above_known_optimum(name):
    d = np.load(os.path.join(GOLDEN, "tsplib_int.npz"))
    xy = O.xy_of(d[name + "_x"], d[name + "_y"])
    n = len(d[name + "_x"])
    r, L, moves, sweeps, log = O.local_search(xy, O.random_perm(n, 1), threads=8)
    assert L >= int(d[name + "_opt"][0])
    assert L == O.route_length(xy, r)
    # a 2-opt local optimum from a random start is within a modest factor
    assert L <= 1.25 * int(d[name + "_opt"][0])


# --------------------------------------------------------------------------
# workloads and kicks
# --------------------------------------------------------------------------
def test_generators_deterministic_and_in_range():
    a = O.gen_uniform(1000, 10**4, 42)
    b = O.gen_uniform(1000, 10**4, 42)
    assert (a == b).all() and a.min() >= 0 and a.max() < 10**4
    assert (O.gen_uniform(1000, 10**4, 43) != a).any()
    c = O.gen_clustered(33810, 10**6, 7)
    assert c.min() >= 0 and c.max() < 10**6
    r = O.random_perm(5000, 3)
    assert sorted(r.tolist()) == list(range(5000))


def test_clustered_structure():
    n = 33810
    xy = O.gen_clustered(n, 10**6, 11).reshape(n, 2).astype(np.float64)
    clustered = xy[: n - 3381]
    # round-robin assignment: points k and k+60 share a centre
    d_same = np.hypot(*(clustered[:3000] - clustered[60:3060]).T)
    d_any = np.hypot(*(clustered[:3000] - clustered[1:3001]).T)
    assert np.median(d_same) < 0.1 * np.median(d_any)
    # per-axis spread within a cluster ~ sigma*sqrt(2) for a difference
    assert 0.8 * 2e4 * math.sqrt(2) < np.std(clustered[:3000:1, 0] - clustered[60:3060, 0]) < 1.2 * 2e4 * math.sqrt(2)


def test_nn_tour_is_greedy():
    n = 300
    xy = O.gen_uniform(n, 100, 8)  # small side: many equal distances
    r = O.nn_tour(xy, 0)
    assert sorted(r.tolist()) == list(range(n)) and r[0] == 0
    used = {0}
    for k in range(1, n):
        cur = int(r[k - 1])
        best = min((((int(xy[2 * u]) - int(xy[2 * cur])) ** 2 + (int(xy[2 * u + 1]) - int(xy[2 * cur + 1])) ** 2, u)
                    for u in range(n) if u not in used))
        assert int(r[k]) == best[1]
        used.add(best[1])


def test_double_bridge_segments():
    g = O.rng(5)
    for n in (4, 5, 10, 1000):
        r = np.random.default_rng(n).permutation(n).astype(np.uint32)
        out, cuts = O.double_bridge(r, g)
        p1, p2, p3 = cuts.tolist()
        assert 1 <= p1 < p2 < p3 <= n - 1
        rl = r.tolist()
        assert out.tolist() == rl[:p1] + rl[p2:p3] + rl[p1:p2] + rl[p3:]


def test_ils_monotone_and_consistent():
    n = 200
    xy = O.gen_uniform(n, 10**4, 21)
    r0 = O.random_perm(n, 22)
    best, Lb, rl, tm = O.ils(xy, r0, seed=99, rounds=30)
    assert (np.diff(rl.astype(np.int64)) <= 0).all()
    assert Lb == int(rl[-1]) == O.route_length(xy, best)
    _, L_lo, _, _, _ = O.local_search(xy, r0)
    assert int(rl[0]) == L_lo
    best2, Lb2, rl2, _ = O.ils(xy, r0, seed=99, rounds=30)
    assert Lb2 == Lb and (best2 == best).all() and (rl2 == rl).all()
    ok, _ = O.step(xy, best)
    assert not ok


# --------------------------------------------------------------------------
# the reference's own CPU path (oracle/_ref), built from its sources
# --------------------------------------------------------------------------
R = O.load_ref()
needs_ref = pytest.mark.skipif(R is None, reason="oracle/_ref not built (reference sources absent)")


def collinear(n, seed, direction=(1, 0)):
    rs = np.random.default_rng(seed)
    t = rs.integers(0, 3000, n)
    x = t * direction[0]
    y = t * direction[1] + 17
    return x.astype(np.int32), y.astype(np.int32)


@needs_ref
@pytest.mark.parametrize("seed", range(6))
def test_reference_cpu_solver_move_sequence_on_integer_distances(seed):
    # collinear along (1,0) or (3,4)*k: every edge length is an integer, where
    # the reference's float gain (cpuSolver.cpp:133-138) is exact
    n = [30, 57, 100, 150, 200, 64][seed]
    d = [(1, 0), (3, 4)][seed % 2]
    x, y = collinear(n, seed, d)
    xy = O.xy_of(x, y)
    fxy = np.ascontiguousarray(np.stack([x, y], 1).astype(np.float32).reshape(-1))
    r0 = O.random_perm(n, seed + 1)
    rref = r0.copy()
    cap = 100000
    log = np.zeros(3 * cap, np.int32)
    nsweeps = R.ref_optimize(fxy, rref, n, log, cap)
    log = log[: 3 * nsweeps].reshape(nsweeps, 3)
    r, L, moves, sweeps, olog = O.local_search(xy, r0)
    assert nsweeps == sweeps
    # every sweep but the last applied a move; same (i, J, gain) sequence
    np.testing.assert_array_equal(log[:-1], olog)
    assert log[-1, 2] == 0
    assert rref.tolist() == r.tolist()
    assert R.ref_route_length(fxy, rref, n) == L


@needs_ref
def test_reference_step_on_integer_distances():
    x, y = collinear(80, 3)
    xy = O.xy_of(x, y)
    fxy = np.ascontiguousarray(np.stack([x, y], 1).astype(np.float32).reshape(-1))
    for seed in range(30):
        r = O.random_perm(80, seed)
        out = np.zeros(3, np.int32)
        R.ref_step(fxy, r, 80, out)
        ok, (i, J, g) = O.step(xy, r)
        assert ok and (i, J, g) == tuple(out.tolist())


REF_AVX_SO = os.path.join(os.path.dirname(O.REF_SO), "libref2opt_avx.so")


@pytest.mark.skipif(not os.path.exists(REF_AVX_SO), reason="oracle/_ref not built (reference sources absent)")
def test_reference_multithreaded_step_is_a_valid_baseline():
    # the benchmark's multi-core CPU baseline is the reference's own pthreads
    # sweep (CPUMTSolver::optimizeStep).  It is timed, not trusted: on these
    # inputs, where every float distance is exact, it returns the exact best
    # move in most cases but not all (it skips some candidates of its thread
    # ranges), so what is checked is that its answer does not depend on the
    # thread count and is a real improving move whose exact gain is within 1
    # of the gain it reports, and that it agrees with the oracle at least half
    # of the time (9 of these 16 tours)
    RA = ctypes.CDLL(REF_AVX_SO)
    f32, u32, i32 = (np.ctypeslib.ndpointer(t) for t in (np.float32, np.uint32, np.int32))
    RA.ref_step_mt.argtypes = [f32, u32, ctypes.c_uint32, ctypes.c_int32, i32]
    n = 120
    x, y = collinear(n, 5, (1, 0))
    xy = O.xy_of(x, y)
    fxy = np.ascontiguousarray(np.stack([x, y], 1).astype(np.float32).reshape(-1))
    agree = 0
    for seed in range(16):
        r = O.random_perm(n, 50 + seed)
        ok, best = O.step(xy, r)
        outs = []
        for threads in (1, 3, 8):
            out = np.zeros(3, np.int32)
            RA.ref_step_mt(fxy, r, n, threads, out)
            outs.append(tuple(int(v) for v in out))
        assert outs[0] == outs[1] == outs[2]
        i, J, g = outs[0]
        assert 1 <= i and i + 2 <= J <= n - 1 and g < 0
        d = lambda a, b: py_dist(xy, int(r[a]), int(r[b]))
        exact = d(i, J) + d(i - 1, J - 1) - d(i - 1, i) - d(J - 1, J)
        assert abs(exact - g) <= 1 and exact < 0
        agree += outs[0] == best
    assert agree >= 8


@needs_ref
def test_reference_length_monalisa():
    m = np.load(os.path.join(GOLDEN, "monalisa.npz"))
    fxy = np.ascontiguousarray(np.stack([m["x"], m["y"]], 1).astype(np.float32).reshape(-1))
    tour = m["tour"].astype(np.uint32)
    assert R.ref_route_length(fxy, tour, len(tour)) == 5757191


@needs_ref
def test_reference_kick_matches_rotate_reverse():
    libc = ctypes.CDLL(None)
    for seed in range(20):
        n = 50 + seed
        r = np.random.default_rng(seed).permutation(n).astype(np.uint32)
        libc.srand(seed)
        a = libc.rand() % n
        b = libc.rand() % n
        got = r.copy()
        R.ref_kick(got, n, seed)
        assert got.tolist() == O.rotate_reverse(r, a, b).tolist()


def test_oracle_multistart_ils_properties():
    # one chain is plain ILS; group best per round never increases and is no
    # worse than any single chain started from the same seed
    xy = O.gen_uniform(150, 10**4, 21)
    r0 = O.random_perm(150, 22)
    r1, L1, rl1, _ = O.ils(xy, r0, 5, 8)
    rm, Lm, rlm, _ = O.ils_multistart(xy, r0, 5, 8, 1)
    assert (Lm, rm.tolist(), rlm.tolist()) == (L1, r1.tolist(), rl1.tolist())
    rg, Lg, rlg, _ = O.ils_multistart(xy, r0, 5, 8, 4)
    assert np.all(np.diff(rlg.astype(np.int64)) <= 0)
    assert Lg == O.route_length(xy, rg) == int(rlg[-1])
    # the group best after round 0 is the common first local optimum
    assert int(rlg[0]) == int(rl1[0])


def test_oracle_multistart_against_round_loop():
    # the same rule written as an explicit round loop over the pinned local
    # search and double-bridge: at each barrier the group best tour is fixed,
    # then every chain kicks (its own best if it improved, else that tour).
    # With this instance, reading the group best's tour after its chain has
    # moved on within a round gives a different history.
    n, chains, rounds, seed = 90, 4, 8, 17
    xy = O.gen_uniform(n, 3000, 5)
    r0 = O.random_perm(n, 6)
    gens = [O.rng(seed + c) for c in range(chains)]
    lo = O.local_search(xy, r0)
    best = [lo[0].copy() for _ in range(chains)]
    bl = [lo[1]] * chains
    imp = [True] * chains
    g = 0
    hist = [bl[0]]
    for _ in range(rounds):
        snap = best[g].copy()
        for c in range(chains):
            kicked, _ = O.double_bridge(best[c] if imp[c] else snap, gens[c])
            r, L = O.local_search(xy, kicked)[:2]
            imp[c] = L < bl[c]
            if imp[c]:
                best[c], bl[c] = r, L
        g = min(range(chains), key=lambda c: (bl[c], c))
        hist.append(bl[g])
    rg, Lg, rlg, _ = O.ils_multistart(xy, r0, seed, rounds, chains)
    assert Lg == bl[g] and rg.tolist() == best[g].tolist() and rlg.tolist() == hist


@pytest.mark.parametrize("isa", ["native", "avx2", "scalar"])
def test_exact_simd_sweep_equals_oracle_step(isa, monkeypatch):
    # the CPU baseline's exact SIMD sweep returns the oracle's move, in every
    # instruction-set variant, for every metric, thread count and tie pattern
    if isa != "native":
        monkeypatch.setenv("ORC_FAST_ISA", isa)
    for n, side in [(4, 10), (5, 3), (7, 2), (9, 2), (17, 5), (40, 30), (300, 10**4), (900, 50), (1500, 2**22 - 1)]:
        for metric in (0, 1, 2):
            xy = O.gen_uniform(n, side, n + metric)
            for seed in range(3):
                r = O.random_perm(n, seed) if seed < 2 else O.nn_tour(xy)
                for th in (1, 3):
                    assert O.step_fast(xy, r, th, metric) == O.step(xy, r, th, metric)
