"""Tour construction (CPU): the library's multiple-fragment (greedy edge) tour
against the oracle's brute-force restatement (all edges sorted) and, at the
serving and config sizes, against its two-phase restatement (pairs within a
radius, then all pairs among the cities left with degree < 2), itself pinned
to the brute force for every radius; the greedy nearest-neighbour tour
against the oracle's O(n^2) scan.  The library takes both from nearest-city
lists built on several host threads, so these are the threaded host tests.

Reference: routeInit's "mf" / "nn" options (construction.c:118-150, default
"mf" in defs.h:70; multipleFragment_KD and initRoute_KD_NN,
kdtree_utils.c:56-301).  The reference's kd-tree drops duplicate points and
orders edges by rounded float lengths, so its tours are not reproduced
exactly; the rule (shortest valid edge first) is.
"""
import numpy as np
import pytest

import logo2opt as L
import oracle_lib as O


@pytest.mark.parametrize("n,side,kind", [
    (3, 10, "u"), (4, 2, "u"), (9, 3, "u"), (64, 8, "u"), (300, 10**4, "u"), (1000, 10**6, "c"),
    (1200, 40, "u"), (2000, 2**22 - 1, "u"),
])
def test_mf_matches_brute_force(n, side, kind):
    xy = O.gen_uniform(n, side, 100 + n) if kind == "u" else O.gen_clustered(n, side, 100 + n)
    assert L.mf_tour(xy).tolist() == O.mf_tour(xy).tolist()


def test_mf_duplicates_and_collinear():
    # every point equal: all edges tie at 0 and the id order decides: (0,1),
    # (0,2), then (1,3), (2,4), (3,5), ... build the path ..5-3-1-0-2-4..
    xy = np.zeros(2 * 50, np.int32)
    r = L.mf_tour(xy)
    assert r.tolist() == O.mf_tour(xy).tolist() == [0] + list(range(1, 50, 2)) + list(range(48, 0, -2))
    # points on a line, shuffled ids: consecutive points joined, ends closed
    n = 200
    perm = O.random_perm(n, 3)
    xy = np.zeros(2 * n, np.int32)
    xy[0::2] = perm.astype(np.int32) * 7
    r = L.mf_tour(xy)
    assert r.tolist() == O.mf_tour(xy).tolist()
    assert O.route_length(xy, r) == 2 * 7 * (n - 1)


def test_mf_is_a_permutation_and_shorter_than_nn_on_average():
    ratios = []
    for seed in range(4):
        xy = O.gen_uniform(20000, 10**6, seed)
        r = L.mf_tour(xy)
        assert np.array_equal(np.sort(r), np.arange(20000))
        ratios.append(O.route_length(xy, r) / O.route_length(xy, L.nn_tour(xy)))
    # greedy edge is known to beat nearest neighbour by several percent on
    # uniform instances (Johnson & McGeoch: ~14-16% vs ~25% above the HK bound)
    assert max(ratios) < 0.97


def test_nn_matches_oracle():
    for n, side, start in [(5, 3, 0), (700, 30, 5), (3000, 10**6, 17)]:
        xy = O.gen_uniform(n, side, n)
        assert L.nn_tour(xy, start).tolist() == O.nn_tour(xy, start).tolist()


def test_mf_rejects_bad_arguments():
    with pytest.raises(L.Logo2OptError):
        L.mf_tour(np.zeros(4, np.int32))  # n = 2


def test_mf_radius_oracle_equals_brute_force_for_every_radius():
    for n, side, kind in [(9, 3, "u"), (64, 8, "u"), (300, 10**4, "u"), (1000, 10**6, "c"), (1200, 40, "u")]:
        xy = O.gen_uniform(n, side, 100 + n) if kind == "u" else O.gen_clustered(n, side, 100 + n)
        want = O.mf_tour(xy).tolist()
        for radius in [0, 1, 2, 5, 50, 10**3, 10**5, 10**7]:
            r = np.zeros(n, np.uint32)
            assert O.lib.orc_mf_tour_radius(n, xy, radius, 1 << 40, r) == 0
            assert r.tolist() == want, (n, radius)


@pytest.mark.parametrize("n,kind,seed", [(33810, "c", 2), (50000, "c", 77), (50000, "u", 5)])
def test_constructions_at_full_size(n, kind, seed):
    xy = O.gen_clustered(n, 10**6, seed) if kind == "c" else O.gen_uniform(n, 10**6, seed)
    assert L.mf_tour(xy).tolist() == O.mf_tour_radius(xy).tolist()
    assert L.nn_tour(xy, 0).tolist() == O.nn_tour(xy, 0).tolist()
