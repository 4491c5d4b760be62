"""logo2opt_step_noalloc (no device allocation, fixed scratch array) against
the existing entry point it varies, logo2opt_step, through the C ABI: the
same move on two calls in a row, the second with a smaller instance (the
scratch array keeps the first call's contents), under every metric; inputs
beyond the array's maximum are refused with LOGO2OPT_ESIZE, invalid ones
with the same codes as logo2opt_create / logo2opt_step."""
import ctypes as C

import numpy as np
import pytest

import logo2opt as L
import oracle_lib as O

pytestmark = pytest.mark.gpu


def lib():
    f = L.load()
    f.logo2opt_step_noalloc.argtypes = [C.c_int, C.c_uint32, L.vp, C.c_int, L.vp, C.POINTER(L.Best2Out)]
    f.logo2opt_step_noalloc.restype = C.c_int
    return f


def noalloc_step(xy, route, metric=L.METRIC_EUC_2D):
    coords = L.coords_of(xy)
    r = np.ascontiguousarray(route, np.uint32)
    o = L.Best2Out()
    rc = lib().logo2opt_step_noalloc(0, len(r), C.cast(coords, L.vp), metric, r.ctypes.data, C.byref(o))
    return rc, (o.i, o.j, o.minchange)


def context_step(xy, route, metric=L.METRIC_EUC_2D):
    with L.Context(L.coords_of(xy), metric=metric) as ctx:
        return ctx.step(route)


@pytest.mark.parametrize("metric", [L.METRIC_EUC_2D, L.METRIC_CEIL_2D, L.METRIC_ATT])
def test_noalloc_step_equals_step_twice_shrinking(metric):
    cases = [(6000, 10**6, 1), (700, 10**4, 2)]  # second call smaller
    for n, side, seed in cases:
        xy = L.gen_clustered(n, side, seed) if seed % 2 else L.gen_uniform(n, side, seed)
        for route in (L.random_perm(n, 10 + seed), L.nn_tour(xy, 0)):
            rc, got = noalloc_step(xy, route, metric)
            assert rc == 0
            want = context_step(xy, route, metric)
            assert got == want, (n, metric, got, want)
            ok, mv = O.step(np.ascontiguousarray(xy, np.int32), route, 8, metric=metric)
            assert got == (mv if ok else (0, 0, 0))


def test_noalloc_step_largest_and_limits():
    n = L.NOALLOC_MAX_N
    xy = L.gen_uniform(n, 10**6, 7)
    r = L.nn_tour(xy, 0)
    rc, got = noalloc_step(xy, r)
    assert rc == 0 and got == context_step(xy, r)
    # then a tiny instance, and a locally optimal tour (no move)
    xy2 = L.gen_uniform(40, 100, 8)
    r2 = L.random_perm(40, 9)
    rc, got = noalloc_step(xy2, r2)
    assert rc == 0 and got == context_step(xy2, r2)
    with L.Context(L.coords_of(xy2)) as ctx:
        opt = ctx.optimize(r2)["route"]
    assert noalloc_step(xy2, opt) == (0, (0, 0, 0))
    # beyond the maximum, and invalid inputs
    big = L.gen_uniform(n + 1, 10**6, 10)
    assert noalloc_step(big, np.arange(n + 1, dtype=np.uint32))[0] == -11  # LOGO2OPT_ESIZE
    bad = r2.copy()
    bad[0] = bad[1]
    assert noalloc_step(xy2, bad)[0] == -3  # LOGO2OPT_ETOUR
    c = L.coords_of(xy2)
    cf = np.frombuffer(c, dtype=np.float32).copy()
    cf[0] += 0.5
    o = L.Best2Out()
    rc = lib().logo2opt_step_noalloc(0, 40, cf.ctypes.data, 0, r2.ctypes.data, C.byref(o))
    assert rc == -2  # LOGO2OPT_ECOORD
