"""ctypes binding of the oracle (oracle/liblogo_oracle.so) and of the reference
driver (oracle/_ref/libref2opt.so).  Test infrastructure only: the product
library never loads either."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liblogo_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref2opt.so")

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


class Best2Out(C.Structure):
    _fields_ = [("i", C.c_uint32), ("j", C.c_uint32), ("minchange", C.c_int32)]


class Rng(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4)]


def _load_oracle():
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError("oracle not built: run `make -C oracle` or __graft_entry__.build()")
    L = C.CDLL(ORACLE_SO)
    L.orc_rng_seed.argtypes = [C.POINTER(Rng), C.c_uint64]
    L.orc_rng_next.argtypes = [C.POINTER(Rng)]
    L.orc_rng_next.restype = C.c_uint64
    L.orc_gen_uniform.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, i32p]
    L.orc_gen_clustered.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                    C.c_uint64, i32p]
    L.orc_random_perm.argtypes = [C.c_uint32, C.c_uint64, u32p]
    L.orc_nn_tour.argtypes = [C.c_uint32, i32p, C.c_uint32, u32p]
    L.orc_mf_tour.argtypes = [C.c_uint32, i32p, u32p]
    L.orc_nint_sqrt.argtypes = [C.c_uint64]
    L.orc_nint_sqrt.restype = C.c_uint32
    L.orc_nint_sqrt_batch.argtypes = [u64p, u32p, C.c_uint64]
    L.orc_route_length.argtypes = [C.c_uint32, i32p, u32p]
    L.orc_route_length.restype = C.c_uint64
    L.orc_check_tour.argtypes = [C.c_uint32, u32p]
    L.orc_step.argtypes = [C.c_uint32, i32p, u32p, C.c_int, C.POINTER(Best2Out)]
    L.orc_apply.argtypes = [C.c_uint32, u32p, C.c_uint32, C.c_uint32]
    L.orc_local_search.argtypes = [C.c_uint32, i32p, u32p, C.c_int, C.c_uint64,
                                   C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint64)]
    L.orc_local_search.restype = C.c_uint64
    L.orc_double_bridge.argtypes = [C.c_uint32, u32p, C.POINTER(Rng), u32p]
    L.orc_rotate_reverse.argtypes = [C.c_uint32, u32p, C.c_uint32, C.c_uint32]
    L.orc_ref_kick.argtypes = [C.c_uint32, u32p, C.POINTER(Rng)]
    L.orc_ils.argtypes = [C.c_uint32, i32p, u32p, C.c_uint64, C.c_uint32, C.c_int, C.c_int,
                          C.c_void_p, C.POINTER(C.c_uint64)]
    L.orc_ils.restype = C.c_uint64
    L.orc_metric_len.argtypes = [C.c_uint64, C.c_int]
    L.orc_metric_len.restype = C.c_uint32
    L.orc_metric_len_batch.argtypes = [u64p, u32p, C.c_uint64, C.c_int]
    L.orc_route_length_m.argtypes = [C.c_uint32, i32p, u32p, C.c_int]
    L.orc_route_length_m.restype = C.c_uint64
    L.orc_step_m.argtypes = [C.c_uint32, i32p, u32p, C.c_int, C.c_int, C.POINTER(Best2Out)]
    L.orc_local_search_m.argtypes = [C.c_uint32, i32p, u32p, C.c_int, C.c_int, C.c_uint64,
                                     C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_uint64)]
    L.orc_local_search_m.restype = C.c_uint64
    L.orc_ils_m.argtypes = [C.c_uint32, i32p, u32p, C.c_uint64, C.c_uint32, C.c_int, C.c_int, C.c_int,
                            C.c_void_p, C.POINTER(C.c_uint64)]
    L.orc_ils_m.restype = C.c_uint64
    L.orc_ils_multistart.argtypes = [C.c_uint32, i32p, u32p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_int,
                                     C.c_int, C.c_void_p, C.POINTER(C.c_uint64)]
    L.orc_ils_multistart.restype = C.c_uint64
    L.orc_step_fast.argtypes = [C.c_uint32, i32p, u32p, C.c_int, C.c_int, C.POINTER(Best2Out),
                                C.POINTER(C.c_uint64)]
    L.orc_fast_isa.restype = C.c_char_p
    L.orc_mf_tour_radius.argtypes = [C.c_uint32, i32p, C.c_uint64, C.c_uint64, u32p]
    L.orc_local_search_fast.argtypes = L.orc_local_search_m.argtypes
    L.orc_local_search_fast.restype = C.c_uint64
    L.orc_ils_fast.argtypes = L.orc_ils_m.argtypes
    L.orc_ils_fast.restype = C.c_uint64
    return L


lib = _load_oracle()


def xy_of(x, y):
    return np.ascontiguousarray(np.stack([np.asarray(x, np.int32), np.asarray(y, np.int32)], 1).reshape(-1))


def gen_uniform(n, side, seed):
    xy = np.zeros(2 * n, np.int32)
    lib.orc_gen_uniform(n, side, seed, xy)
    return xy


def gen_clustered(n, side, seed, nclusters=60, sigma_frac=0.02, noise_frac=0.10):
    xy = np.zeros(2 * n, np.int32)
    lib.orc_gen_clustered(n, side, nclusters, sigma_frac, noise_frac, seed, xy)
    return xy


def random_perm(n, seed):
    r = np.zeros(n, np.uint32)
    lib.orc_random_perm(n, seed, r)
    return r


def nn_tour(xy, start=0):
    n = len(xy) // 2
    r = np.zeros(n, np.uint32)
    lib.orc_nn_tour(n, xy, start, r)
    return r


def mf_tour(xy):
    n = len(xy) // 2
    r = np.zeros(n, np.uint32)
    assert lib.orc_mf_tour(n, xy, r) == 0
    return r


def mf_tour_radius(xy, radius=None, max_pairs=1 << 27):
    """The greedy-edge tour for large n: pairs within `radius` first (default:
    4x the mean spacing over the bounding box), then all pairs among the
    cities left with degree < 2; the radius doubles until the second phase
    fits max_pairs."""
    n = len(xy) // 2
    r = np.zeros(n, np.uint32)
    if radius is None:
        x, y = xy[0::2].astype(np.int64), xy[1::2].astype(np.int64)
        area = max(1, int(x.max() - x.min() + 1)) * max(1, int(y.max() - y.min() + 1))
        radius = max(1, int(4 * (area / n) ** 0.5))
    while lib.orc_mf_tour_radius(n, xy, radius, max_pairs, r) != 0:
        radius *= 2
    return r


def nint_sqrt(s):
    s = np.ascontiguousarray(s, np.uint64)
    out = np.zeros(len(s), np.uint32)
    lib.orc_nint_sqrt_batch(s, out, len(s))
    return out


METRIC_EUC_2D, METRIC_CEIL_2D, METRIC_ATT = 0, 1, 2


def metric_len(s, metric):
    s = np.ascontiguousarray(s, np.uint64)
    out = np.zeros(len(s), np.uint32)
    lib.orc_metric_len_batch(s, out, len(s), metric)
    return out


def route_length(xy, route, metric=METRIC_EUC_2D):
    return int(lib.orc_route_length_m(len(route), xy, np.ascontiguousarray(route, np.uint32), metric))


def step(xy, route, threads=1, metric=METRIC_EUC_2D):
    o = Best2Out()
    ok = lib.orc_step_m(len(route), xy, np.ascontiguousarray(route, np.uint32), threads, metric, C.byref(o))
    return bool(ok), (o.i, o.j, o.minchange)


def step_fast(xy, route, threads=1, metric=METRIC_EUC_2D):
    """Exact SIMD sweep (the CPU baseline): same result as step()."""
    o = Best2Out()
    sv = C.c_uint64()
    ok = lib.orc_step_fast(len(route), xy, np.ascontiguousarray(route, np.uint32), threads, metric, C.byref(o),
                           C.byref(sv))
    return bool(ok), (o.i, o.j, o.minchange)


def fast_isa():
    return lib.orc_fast_isa().decode()


def apply(route, i, j):
    r = np.array(route, np.uint32)
    lib.orc_apply(len(r), r, i, j)
    return r


def local_search(xy, route, threads=1, max_moves=0, log_cap=1 << 20, metric=METRIC_EUC_2D, fast=False):
    """fast=True runs the loop on the exact SIMD sweep (same moves)."""
    r = np.array(route, np.uint32)
    log = np.zeros(3 * log_cap, np.int32)
    sweeps = C.c_uint64()
    length = C.c_uint64()
    fn = lib.orc_local_search_fast if fast else lib.orc_local_search_m
    moves = fn(len(r), xy, r, threads, metric, max_moves, log.ctypes.data, log_cap, C.byref(sweeps),
               C.byref(length))
    m = min(moves, log_cap)
    return r, int(length.value), int(moves), int(sweeps.value), log[:3 * m].reshape(m, 3).copy()


def ils(xy, route, seed, rounds, kick="double_bridge", threads=1, metric=METRIC_EUC_2D, fast=False):
    """fast=True runs the local searches on the exact SIMD sweep (same moves)."""
    r = np.array(route, np.uint32)
    rl = np.zeros(rounds + 1, np.uint64)
    tm = C.c_uint64()
    fn = lib.orc_ils_fast if fast else lib.orc_ils_m
    best = fn(len(r), xy, r, seed, rounds, 0 if kick == "double_bridge" else 1, threads, metric, rl.ctypes.data,
              C.byref(tm))
    return r, int(best), rl, int(tm.value)


def ils_multistart(xy, route, seed, rounds, chains, kick="double_bridge", threads=1, metric=METRIC_EUC_2D):
    r = np.array(route, np.uint32)
    rl = np.zeros(rounds + 1, np.uint64)
    tm = C.c_uint64()
    best = lib.orc_ils_multistart(len(r), xy, r, seed, rounds, chains, 0 if kick == "double_bridge" else 1,
                                  threads, metric, rl.ctypes.data, C.byref(tm))
    return r, int(best), rl, int(tm.value)


def rng(seed):
    g = Rng()
    lib.orc_rng_seed(C.byref(g), seed)
    return g


def double_bridge(route, g):
    r = np.array(route, np.uint32)
    cuts = np.zeros(3, np.uint32)
    lib.orc_double_bridge(len(r), r, C.byref(g), cuts)
    return r, cuts


def rotate_reverse(route, a, b):
    r = np.array(route, np.uint32)
    lib.orc_rotate_reverse(len(r), r, a, b)
    return r


# ---- the reference's own CPU path, compiled from its sources -------------
def load_ref():
    if not os.path.exists(REF_SO):
        return None
    R = C.CDLL(REF_SO)
    R.ref_route_length.argtypes = [f32p, u32p, C.c_uint32]
    R.ref_route_length.restype = C.c_uint64
    R.ref_step.argtypes = [f32p, u32p, C.c_uint32, i32p]
    R.ref_optimize.argtypes = [f32p, u32p, C.c_uint32, i32p, C.c_uint64]
    R.ref_optimize.restype = C.c_uint64
    R.ref_kick.argtypes = [u32p, C.c_uint32, C.c_uint]
    return R
