"""Pins the sweep kernel's first filter tier (logo-tsp-solver_amd/csrc/l2o_kernels.hip)
on the CPU, in exact integer arithmetic independent of the kernel.

The kernel skips a candidate (a, b) unless
    L1(a,b) + L1(a+1,b+1) <= ceil(sqrt2 * best) + W[a+1] + W[b+1],
W[k] = ceil(sqrt(2) * D[k]), best = the wave's best exact gain so far (<= 0).
That is sound only if every candidate that could still win (exact gain
G <= best - 1) passes.  The claim rests on nint(d) >= d - 1/2 and
|dx| + |dy| <= sqrt(2) d; these tests check the claim itself, over every
candidate of small instances, at coordinate spans up to the 2^22 limit and on
tours with many near-ties, and check the integer ceil(sqrt(2) d) the kernel and
host compute.  A mistake such as a missing 1/2 per rounding, a floor for a
ceil, or the row and column terms of the wrong edges fails here."""
import math

import numpy as np
import pytest

import oracle_lib as O


def ceil_sqrt2_ref(d):
    """Smallest m with m^2 >= 2 d^2, by definition."""
    m = math.isqrt(2 * d * d)
    return m if m * m >= 2 * d * d else m + 1


def ceil_sqrt2_kernel(d):
    """The kernel's and host's formula (l2o_launch.h): 0 for d = 0, else
    isqrt(2 d^2) + 1 -- valid because 2 d^2 is never a perfect square."""
    return 0 if d <= 0 else math.isqrt(2 * d * d) + 1


def tier1_base(best):
    """ceil(sqrt(2) * best) for best <= 0, as the kernel computes it."""
    return 0 if best == 0 else 1 - ceil_sqrt2_kernel(-best)


def test_ceil_sqrt2_matches_definition():
    rng = np.random.default_rng(3)
    ds = list(range(0, 5000)) + [int(v) for v in rng.integers(1, 1 << 24, 20000)] + [(1 << 24) - 1]
    for d in ds:
        assert ceil_sqrt2_kernel(d) == ceil_sqrt2_ref(d), d
        assert ceil_sqrt2_kernel(d) >= math.sqrt(2) * d


def test_tier1_base_is_ceiling():
    for best in range(0, -20000, -1):
        tb = tier1_base(best)
        # tb is the smallest integer >= sqrt(2) * best
        assert tb <= 0 and tb * tb <= 2 * best * best
        assert (tb - 1) < math.sqrt(2) * best <= tb


def _gains_and_l1(xy, r):
    """Exact gain G and L1 sum of every candidate (a, b), a+2 <= b <= n-2,
    plus the W terms of its row and column edge (python integers)."""
    n = len(r)
    X = [int(xy[2 * c]) for c in r]
    Y = [int(xy[2 * c + 1]) for c in r]

    def nint(p, q):
        dx, dy = X[p] - X[q], Y[p] - Y[q]
        return (math.isqrt(4 * (dx * dx + dy * dy)) + 1) // 2

    def l1(p, q):
        return abs(X[p] - X[q]) + abs(Y[p] - Y[q])

    D = [nint(k - 1 if k else n - 1, k) for k in range(n)]
    W = [ceil_sqrt2_kernel(d) for d in D]
    out = []
    for a in range(0, n - 3):
        for b in range(a + 2, n - 1):
            G = nint(a, b) + nint(a + 1, b + 1) - D[a + 1] - D[b + 1]
            out.append((G, l1(a, b) + l1(a + 1, b + 1), W[a + 1] + W[b + 1]))
    return out


def _check(xy, r):
    cands = _gains_and_l1(xy, r)
    gains = sorted({g for g, _, _ in cands})
    # thresholds at no move found yet, at the optimum, and in between
    bests = {0, min(min(gains), 0)}
    bests.update(g for g in gains[:: max(1, len(gains) // 25)] if g <= 0)
    for best in bests:
        tb = tier1_base(best)
        for G, L, Wsum in cands:
            if G <= best - 1:
                assert L <= tb + Wsum, (best, G, L, Wsum)
    return len(cands)


@pytest.mark.parametrize("n,side,kind,seed", [
    (60, 10**4, "u", 1), (80, 10**6, "c", 2), (70, (1 << 22) - 1, "u", 3), (64, 40, "u", 4), (90, 3, "u", 5),
])
def test_tier1_never_drops_a_winner_random_tours(n, side, kind, seed):
    xy = O.gen_uniform(n, side, seed) if kind == "u" else O.gen_clustered(n, side, seed)
    for s in range(2):
        assert _check(xy, O.random_perm(n, 100 + s)) > 0


def test_tier1_never_drops_a_winner_near_optimum():
    # near a local optimum the best gains are small and many candidates sit
    # close to the threshold
    for n, side in [(120, 10**4), (100, (1 << 22) - 1), (110, 25)]:
        xy = O.gen_uniform(n, side, n)
        r = O.nn_tour(xy, 0)
        _check(xy, r)
        r2 = O.local_search(xy, r, max_moves=max(1, n // 10))[0]
        _check(xy, r2)


def test_tier1_tight_instance():
    # A=(40,40) B=(0,2) C=(41,41) E=(39,41), tour A B C E: the one candidate
    # has both new edges at 45 degrees (L1 = sqrt2 d exactly), n(B,E) =
    # nint(39 sqrt2 = 55.15) rounds down, and G = 1 + 55 - 55 - 2 = -1.  Since
    # nint rounds half up, d < nint(d) + 1/2 strictly and the bound always
    # keeps a margin of at least 1; here the margin is exactly 1, so any
    # threshold lower by 2 or more (a dropped rounding term, a W floored) fails.
    xy = O.xy_of(np.array([40, 0, 41, 39], np.int32), np.array([40, 2, 41, 41], np.int32))
    r = np.arange(4, dtype=np.uint32)
    (G, L, Wsum), = _gains_and_l1(xy, r)
    assert G == -1 and L == 80 and Wsum == 81
    assert tier1_base(0) + Wsum - L == 1
    _check(xy, r)


def test_tier1_extreme_geometry():
    # collinear points, duplicates and the largest allowed span
    span = (1 << 22) - 1
    x = np.array([0, span, 0, span, 1, span - 1, 0, 0, span, 7, 7, 7, span // 2, span // 3, 2, span], np.int32)
    y = np.array([0, span, span, 0, 1, span - 1, 0, 5, 3, 7, 7, 8, span // 2, 2 * span // 3, span, 1], np.int32)
    xy = O.xy_of(x, y)
    for s in range(6):
        _check(xy, O.random_perm(len(x), s))


# --- tier 0: the coarse test on coordinates >> q that gates tier 1 ----------

T0_BIAS = 1 << 16


def tier0_shift(max_coord):
    """The host's shift (l2o_launch.h tier0_shift): smallest q with
    max_coord >> q <= 255."""
    q = 0
    while (max_coord >> q) > 255:
        q += 1
    return q


def _tier0_keeps(Sq, wq, tb, wa, q):
    """The kernel's test, with wq = the column's (or the lane's) ceil(W[b+1]/Q)
    and wa = W[a+1]: BIAS + Sq - wq - ceil(wa/Q) <= BIAS + floor((tb + 4 (Q -
    1)) / Q), floor by arithmetic shift (numpy arrays or ints)."""
    Q = 1 << q
    return T0_BIAS + Sq - wq - (-(-wa // Q)) <= T0_BIAS + ((tb + 4 * (Q - 1)) >> q)


def _tier0_cands(X, Y, q):
    """Per candidate (a, b), a+2 <= b <= n-2, of the tour order (X, Y): the
    tier-1 sum L1(a,b) + L1(a+1,b+1), W[a+1], W[b+1] and the L1 sum Sq of the
    shifted coordinates, as int64 arrays."""
    n = len(X)
    X = np.asarray(X, np.int64)
    Y = np.asarray(Y, np.int64)
    dx = X - np.roll(X, 1)
    dy = Y - np.roll(Y, 1)
    D = [(math.isqrt(4 * int(u * u + v * v)) + 1) // 2 for u, v in zip(dx, dy)]
    W = np.array([ceil_sqrt2_kernel(d) for d in D], np.int64)
    a, b = np.array([(a, b) for a in range(n - 3) for b in range(a + 2, n - 1)]).T
    L = np.abs(X[a] - X[b]) + np.abs(Y[a] - Y[b]) + np.abs(X[a + 1] - X[b + 1]) + np.abs(Y[a + 1] - Y[b + 1])
    Xs, Ys = X >> q, Y >> q
    Sq = (np.abs(Xs[a] - Xs[b]) + np.abs(Ys[a] - Ys[b]) + np.abs(Xs[a + 1] - Xs[b + 1]) +
          np.abs(Ys[a + 1] - Ys[b + 1]))
    return L, W[a + 1], W[b + 1], Sq, b


def _check_tier0(X, Y, lane_cols=6):
    """Every candidate that passes tier 1 (L - W[b+1] <= tb + W[a+1]) at any
    threshold passes tier 0, with wq = ceil(W[b+1]/2^q) and with the lane's
    largest such value (columns grouped lane_cols to a lane), and with W[a+1]
    replaced by the largest of its block of four rows (a // 4), as the kernel
    stages it."""
    q = tier0_shift(int(max(np.max(X), np.max(Y))))
    assert max(np.max(X), np.max(Y)) >> q <= 255
    L, wa, wb, Sq, b = _tier0_cands(X, Y, q)
    n = len(X)
    a = np.array([a for a in range(n - 3) for _ in range(a + 2, n - 1)])
    blk = np.zeros(a.max() // 4 + 1, np.int64)
    np.maximum.at(blk, a // 4, wa)
    wa_blk = blk[a // 4]
    assert np.all(wa_blk >= wa)
    Q = 1 << q
    wq = -(-wb // Q)
    lane = (b - 2) // lane_cols
    wq_lane = np.zeros(lane.max() + 1, np.int64)
    np.maximum.at(wq_lane, lane, wq)
    slack = L - wb - wa  # tier 1 passes iff slack <= tb
    for tb in sorted(set(int(v) for v in np.unique(slack)[::max(1, len(slack) // 40)]) | {0, int(slack.min())}):
        if tb > 0:
            continue
        t1 = slack <= tb
        assert np.all(_tier0_keeps(Sq, wq, tb, wa, q)[t1]), (q, tb)
        assert np.all(_tier0_keeps(Sq, wq_lane[lane], tb, wa, q)[t1]), (q, tb)
        assert np.all(_tier0_keeps(Sq, wq_lane[lane], tb, wa_blk, q)[t1]), (q, tb)
        if q == 0:  # no shift: tier 0 is tier 1
            assert np.array_equal(_tier0_keeps(Sq, wq, tb, wa, q), t1)
    return q


def test_tier0_shift():
    assert [tier0_shift(m) for m in (0, 1, 255, 256, 511, 512, (1 << 22) - 1)] == [0, 0, 0, 1, 1, 2, 14]


@pytest.mark.parametrize("n,side,kind,seed", [
    (70, 200, "u", 1), (80, 10**4, "c", 2), (70, (1 << 22) - 1, "u", 3), (90, 3000, "u", 4), (60, 2**16, "c", 5),
])
def test_tier0_never_drops_a_tier1_hit(n, side, kind, seed):
    xy = O.gen_uniform(n, side, seed) if kind == "u" else O.gen_clustered(n, side, seed)
    x, y = xy[0::2].astype(np.int64), xy[1::2].astype(np.int64)
    x -= x.min()
    y -= y.min()
    for s in range(2):
        r = O.random_perm(n, 200 + s)
        _check_tier0(x[r], y[r])
    r = O.local_search(xy, O.nn_tour(xy, 0), max_moves=n // 8)[0]
    _check_tier0(x[r], y[r])


def _tier0_one(P, q):
    """(L, W[a+1], W[b+1], Sq) of candidate (0, 3) of the 6-point tour P."""
    X = np.array([p[0] for p in P], np.int64)
    Y = np.array([p[1] for p in P], np.int64)
    L, wa, wb, Sq, b = _tier0_cands(X, Y, q)
    pairs = [(a, bb) for a in range(3) for bb in range(a + 2, 5)]
    k = pairs.index((0, 3))
    return int(L[k]), int(wa[k]), int(wb[k]), int(Sq[k])


def test_tier0_tight_instances():
    # Candidate (0, 3) of a 6-city tour, Q = 16.  Each coordinate of P0 is
    # Q - 1 mod Q and the matching one of P3 is 0 mod Q, and likewise P1 and
    # P4, so each of the four differences loses exactly Q - 1 to the shift:
    # L = Q Sq - 4 (Q - 1).  tb is set so that tier 1 passes with no margin.
    q, Q = 4, 16
    fill = [(0, 240), (240, 0)]

    def case(P):
        L, wa, wb, Sq = _tier0_one(P, q)
        assert L == Q * Sq - 4 * (Q - 1)
        tb = L - wa - wb
        assert tb <= 0
        return L, wa, wb, Sq, tb
    # (1) d(P0,P1) = d(P3,P4) = 181, W = 256 = 16 Q: tier 0 passes with no
    # margin either, so a rounding term dropped from its threshold or W[a+1]
    # left out of it fails
    L, wa, wb, Sq, tb = case([(63, 95), (191, 223), fill[0], (64, 96), (192, 224), fill[1]])
    assert (L, wa, wb, Sq) == (4, 256, 256, 4)
    wq = -(-wb // Q)
    assert _tier0_keeps(Sq, wq, tb, wa, q)
    assert Sq - wq - wa // Q == (tb + 4 * (Q - 1)) >> q  # margin 0
    assert not Sq - wq - wa // Q <= (tb + 3 * (Q - 1)) >> q  # 3 (Q-1) for 4 (Q-1)
    assert not Sq - wq <= (tb + 4 * (Q - 1)) >> q  # W[a+1] dropped
    # (2) W[a+1] = 160 = 10 Q, W[b+1] = 177 = 11 Q + 1: ceil(W[b+1]/Q) keeps
    # the candidate, floor does not
    L, wa, wb, Sq, tb = case([(63, 95), (143, 175), fill[0], (64, 96), (144, 192), fill[1]])
    assert (wa, wb) == (160, 177)
    assert _tier0_keeps(Sq, -(-wb // Q), tb, wa, q)
    assert not _tier0_keeps(Sq, wb // Q, tb, wa, q)
    # (3) W[a+1] = 132 = 8 Q + 4, W[b+1] = 160: ceil(W[a+1]/Q) keeps it, floor
    # does not
    L, wa, wb, Sq, tb = case([(63, 95), (143, 143), fill[0], (64, 96), (144, 176), fill[1]])
    assert (wa, wb) == (132, 160)
    assert _tier0_keeps(Sq, wb // Q, tb, wa, q)
    assert not Sq - wb // Q - wa // Q <= (tb + 4 * (Q - 1)) >> q
    # tier 0 keeps a candidate at every looser threshold too
    for t in range(tb, 1):
        assert _tier0_keeps(Sq, wb // Q, t, wa, q)
