"""Pins the sweep's row-pruning bound (the PRUNE variant of k_sweep in
logo-tsp-solver_amd/csrc/l2o_kernels.hip) on the CPU, in exact integer
arithmetic independent of the kernel.

A wave owns columns b in [b0, b0 + C) (C = 64K - 1).  The kernel takes the
bounding box B of the points P[b0 .. b0 + C] (the last one is the b+1 of its
last column) and Wmax = max W[b+1] over b in [b0, b0 + C], b <= n-2, and skips
row a when
    lb(P[a]) + lb(P[a+1]) - Wmax > tb + W[a+1],   lb = L1 distance to B.
That is sound only if no such row holds a candidate that passes the first
tier, L1(a,b) + L1(a+1,b+1) - W[b+1] <= tb + W[a+1].  These tests check the
claim over every row and column of small instances (random, nearest-neighbour
and locally optimal tours, several thresholds), that the bound does skip rows
where it should, and that plausible mistakes in it (no Wmax term, the W of the
wrong row edge, a box missing the last column's b+1 point) are caught."""
import math

import numpy as np
import pytest

import oracle_lib as O


def ceil_sqrt2(d):
    return 0 if d <= 0 else math.isqrt(2 * d * d) + 1


def tier1_base(best):
    return 0 if best == 0 else 1 - ceil_sqrt2(-best)


def tour_arrays(xy, r):
    P = np.asarray(xy, np.int64).reshape(-1, 2)[np.asarray(r, np.int64)]
    X, Y = P[:, 0], P[:, 1]
    n = len(r)
    D = np.array([(math.isqrt(4 * int((X[k] - X[k - 1]) ** 2 + (Y[k] - Y[k - 1]) ** 2)) + 1) // 2
                  for k in range(n)], np.int64)
    W = np.array([ceil_sqrt2(int(d)) for d in D], np.int64)
    return X, Y, W


def hits_and_skips(X, Y, W, b0, C, tb, variant=None):
    """Per row a in [0, n-3): does a candidate of columns [b0, b0+C) pass
    tier 1, and does the pruning test skip the row."""
    n = len(X)
    a = np.arange(n - 3)[:, None]
    b = np.arange(b0, min(b0 + C, n - 1))[None, :]
    S = (np.abs(X[a] - X[b]) + np.abs(Y[a] - Y[b]) + np.abs(X[a + 1] - X[b + 1]) + np.abs(Y[a + 1] - Y[b + 1])
         - W[b + 1])
    hit = ((b >= a + 2) & (S <= tb + W[a + 1])).any(axis=1)
    last = b0 + C - (1 if variant == "box_without_last" else 0)
    cols = np.clip(np.arange(b0, last + 1), 0, n - 1)
    bx0, bx1, by0, by1 = X[cols].min(), X[cols].max(), Y[cols].min(), Y[cols].max()
    bw = np.arange(b0, b0 + C + 1)
    wmax = W[bw[bw <= n - 2] + 1].max()
    if variant == "no_wmax":
        wmax = 0

    def lb(x, y):
        return np.maximum(np.maximum(bx0 - x, x - bx1), 0) + np.maximum(np.maximum(by0 - y, y - by1), 0)

    rows = np.arange(n - 3)
    wr = W[rows] if variant == "row_w_of_a" else W[rows + 1]
    skip = lb(X[rows], Y[rows]) + lb(X[rows + 1], Y[rows + 1]) - wmax > tb + wr
    return hit, skip


def cases():
    out = []
    for n, side, kind, seed in [(300, 10**4, "u", 1), (400, 10**6, "c", 2), (250, (1 << 22) - 1, "c", 3),
                                (200, 50, "u", 4)]:
        xy = O.gen_uniform(n, side, seed) if kind == "u" else O.gen_clustered(n, side, seed)
        nn = O.nn_tour(xy, 0)
        lo = O.local_search(xy, nn, threads=4)[0]
        out += [(xy, O.random_perm(n, seed + 10)), (xy, nn), (xy, lo)]
    return out


CASES = cases()
THRESH = [tier1_base(b) for b in (0, -1, -7, -100, -5000)]


def scan(variant=None):
    """(rows skipped, rows with a hit that the test skipped) over all cases."""
    skipped = wrong = 0
    for xy, r in CASES:
        X, Y, W = tour_arrays(xy, r)
        n = len(X)
        for C in (5, 63, 127, 383):
            for b0 in range(2, n - 1, C):
                for tb in THRESH:
                    hit, skip = hits_and_skips(X, Y, W, b0, C, tb, variant)
                    skipped += int(skip.sum())
                    wrong += int((hit & skip).sum())
    return skipped, wrong


def test_pruning_never_skips_a_row_with_a_tier1_hit():
    skipped, wrong = scan()
    assert wrong == 0
    assert skipped > 0


def test_pruning_skips_most_rows_far_from_the_columns_on_a_good_tour():
    xy, lo = CASES[5]  # clustered, locally optimal tour
    X, Y, W = tour_arrays(xy, lo)
    n = len(X)
    hit, skip = hits_and_skips(X, Y, W, n // 2, 63, 0)
    # rows that lie far before the columns in tour order are spatially far
    assert skip[: n // 4].mean() > 0.5


@pytest.mark.parametrize("variant", ["no_wmax", "row_w_of_a", "box_without_last"])
def test_plausible_mistakes_are_caught(variant):
    skipped, wrong = scan(variant)
    assert wrong > 0, variant
