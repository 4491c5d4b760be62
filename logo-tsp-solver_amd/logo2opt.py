"""Python binding (ctypes) of the C ABI in include/logo2opt.h.

Thin host-side mirror for tests and the benchmark: every call goes straight to
liblogo2opt.so (HIP kernels on the GPU).  There is no CPU fallback: if the
library is missing or a call fails, an exception is raised.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblogo2opt.so")

KICK_DOUBLE_BRIDGE = 0
KICK_REFERENCE = 1
MAX_N = 1 << 20
NOALLOC_MAX_N = 1 << 16  # LOGO2OPT_NOALLOC_MAX_N
METRIC_EUC_2D = 0
METRIC_CEIL_2D = 1
METRIC_ATT = 2
METRICS = {"EUC_2D": METRIC_EUC_2D, "CEIL_2D": METRIC_CEIL_2D, "ATT": METRIC_ATT}
PRUNE_OFF, PRUNE_ALL, PRUNE_AUTO = 0, 1, 2
ORDER_FIFO, ORDER_SHORTEST = 0, 1  # LOGO2OPT_ORDER_FIFO / LOGO2OPT_ORDER_SHORTEST


def _prune_mode(on):
    """True / False map to PRUNE_ALL / PRUNE_OFF; ints pass through."""
    return int(on) if not isinstance(on, bool) else (PRUNE_ALL if on else PRUNE_OFF)
TSP_ROUND = 1
EFORMAT = -9
EIO = -10

# every entry point the header declares (checked by the CPU test suite)
EXPORTS = [
    "logo2opt_strerror", "logo2opt_create", "logo2opt_destroy", "logo2opt_step", "logo2opt_optimize",
    "logo2opt_ils", "logo2opt_route_length", "logo2opt_get_stats", "logo2opt_set_timing", "logo2opt_time_sweep",
    "logo2opt_engine_create", "logo2opt_engine_destroy", "logo2opt_engine_bytes_for", "logo2opt_engine_add",
    "logo2opt_engine_run", "logo2opt_engine_result", "logo2opt_engine_round_lengths", "logo2opt_engine_remove",
    "logo2opt_engine_poll_done", "logo2opt_engine_stats", "logo2opt_engine_set_timing", "logo2opt_engine_set_graphs", "logo2opt_gen_uniform", "logo2opt_gen_clustered",
    "logo2opt_random_perm", "logo2opt_nn_tour", "logo2opt_xy_to_coords", "logo2opt_sqrt_ulp_check",
    "logo2opt_tuning_shapes", "logo2opt_tuning_set", "logo2opt_tuning_get", "logo2opt_tuning_load",
    "logo2opt_engine_time_sweep", "logo2opt_create_metric", "logo2opt_tsp_read", "logo2opt_tour_read",
    "logo2opt_tour_write", "logo2opt_host_route_length", "logo2opt_mf_tour", "logo2opt_ils_multistart",
    "logo2opt_step_noalloc",
    "logo2opt_engine_add_group", "logo2opt_engine_group_result", "logo2opt_set_pruning",
    "logo2opt_engine_set_pruning", "logo2opt_ils_until", "logo2opt_stop_below",
]


class CityCoords(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class Best2Out(C.Structure):
    _fields_ = [("i", C.c_uint32), ("j", C.c_uint32), ("minchange", C.c_int32)]


class ProcessTime(C.Structure):
    _fields_ = [("kernel_time", C.c_double), ("HtD_time", C.c_double), ("DtH_time", C.c_double),
                ("out", Best2Out)]


class Stats(C.Structure):
    _fields_ = [("sweeps", C.c_uint64), ("moves", C.c_uint64), ("evals", C.c_uint64),
                ("fallbacks", C.c_uint64), ("sweep_ms", C.c_double), ("wall_ms", C.c_double),
                ("sweep_launches", C.c_uint64)]


class Job(C.Structure):
    _fields_ = [("ils_rounds", C.c_uint32), ("kick", C.c_uint32), ("seed", C.c_uint64),
                ("log_cap", C.c_uint32), ("metric", C.c_uint32), ("stop_below", C.c_uint64)]


class TspInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("n", C.c_uint32), ("metric", C.c_int32), ("offset_x", C.c_double),
                ("offset_y", C.c_double), ("integral", C.c_uint32), ("reserved", C.c_uint32)]


class Result(C.Structure):
    _fields_ = [("length", C.c_uint64), ("moves", C.c_uint64), ("sweeps", C.c_uint64),
                ("fallbacks", C.c_uint64), ("rounds_done", C.c_uint32), ("done", C.c_uint32)]


class Logo2OptError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {strerror(code)} ({code})")
        self.code = code


_lib = None
vp = C.c_void_p
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def load():
    """Loads liblogo2opt.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with logo-tsp-solver_amd/build.py")
    L = C.CDLL(LIB_PATH)
    L.logo2opt_strerror.argtypes = [C.c_int]
    L.logo2opt_strerror.restype = C.c_char_p
    L.logo2opt_create.argtypes = [C.c_int, C.c_uint32, vp, C.POINTER(vp)]
    L.logo2opt_create_metric.argtypes = [C.c_int, C.c_uint32, vp, C.c_int, C.POINTER(vp)]
    L.logo2opt_tsp_read.argtypes = [C.c_char_p, C.c_uint32, C.POINTER(TspInfo), vp, C.c_uint32]
    L.logo2opt_tour_read.argtypes = [C.c_char_p, C.c_uint32, u32p]
    L.logo2opt_tour_write.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32, u32p, C.c_uint64, C.c_int]
    L.logo2opt_host_route_length.argtypes = [C.c_uint32, vp, C.c_int, u32p, C.POINTER(C.c_uint64)]
    L.logo2opt_destroy.argtypes = [vp]
    L.logo2opt_step.argtypes = [vp, u32p, C.POINTER(Best2Out), C.POINTER(ProcessTime)]
    L.logo2opt_optimize.argtypes = [vp, u32p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), vp, C.c_uint64,
                                    C.c_uint32]
    L.logo2opt_ils.argtypes = [vp, u32p, C.c_uint64, C.c_uint32, C.c_int, C.POINTER(C.c_uint64), vp]
    L.logo2opt_route_length.argtypes = [vp, u32p, C.POINTER(C.c_uint64)]
    L.logo2opt_ils_multistart.argtypes = [vp, u32p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int,
                                          C.POINTER(C.c_uint64), vp]
    L.logo2opt_ils_until.argtypes = [vp, u32p, C.c_uint64, C.c_uint32, C.c_int, C.c_uint64, C.c_float,
                                     C.c_double, C.POINTER(C.c_uint64), vp, C.POINTER(C.c_uint32)]
    L.logo2opt_stop_below.argtypes = [C.c_uint64, C.c_float]
    L.logo2opt_stop_below.restype = C.c_uint64
    L.logo2opt_engine_add_group.argtypes = [vp, C.c_uint32, vp, u32p, C.POINTER(Job), C.c_uint32,
                                            C.POINTER(C.c_int32)]
    L.logo2opt_engine_group_result.argtypes = [vp, C.c_int32, vp, C.POINTER(Result), C.POINTER(C.c_int32)]
    L.logo2opt_get_stats.argtypes = [vp, C.POINTER(Stats)]
    L.logo2opt_set_timing.argtypes = [vp, C.c_int]
    L.logo2opt_engine_set_timing.argtypes = [vp, C.c_int]
    L.logo2opt_engine_set_graphs.argtypes = [vp, C.c_int]
    L.logo2opt_engine_set_pruning.argtypes = [vp, C.c_int]
    L.logo2opt_set_pruning.argtypes = [vp, C.c_int]
    L.logo2opt_time_sweep.argtypes = [vp, u32p, C.c_uint32, C.POINTER(C.c_double)]
    L.logo2opt_engine_create.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.POINTER(vp)]
    L.logo2opt_engine_destroy.argtypes = [vp]
    L.logo2opt_engine_bytes_for.argtypes = [C.c_uint32, C.POINTER(Job)]
    L.logo2opt_engine_bytes_for.restype = C.c_uint64
    L.logo2opt_engine_add.argtypes = [vp, C.c_uint32, vp, u32p, C.POINTER(Job), C.POINTER(C.c_int32)]
    L.logo2opt_engine_run.argtypes = [vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]
    L.logo2opt_engine_result.argtypes = [vp, C.c_int32, vp, C.POINTER(Result), vp]
    L.logo2opt_engine_round_lengths.argtypes = [vp, C.c_int32, u64p, C.c_uint32]
    L.logo2opt_engine_remove.argtypes = [vp, C.c_int32]
    L.logo2opt_engine_poll_done.argtypes = [vp, vp, C.c_uint32, C.POINTER(C.c_uint32)]
    L.logo2opt_engine_stats.argtypes = [vp, C.POINTER(Stats)]
    L.logo2opt_gen_uniform.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, i32p]
    L.logo2opt_gen_uniform.restype = None
    L.logo2opt_gen_clustered.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                         C.c_uint64, i32p]
    L.logo2opt_gen_clustered.restype = None
    L.logo2opt_random_perm.argtypes = [C.c_uint32, C.c_uint64, u32p]
    L.logo2opt_random_perm.restype = None
    L.logo2opt_nn_tour.argtypes = [C.c_uint32, i32p, C.c_uint32, u32p]
    L.logo2opt_mf_tour.argtypes = [C.c_uint32, i32p, u32p]
    L.logo2opt_xy_to_coords.argtypes = [C.c_uint32, i32p, vp]
    L.logo2opt_xy_to_coords.restype = None
    L.logo2opt_sqrt_ulp_check.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]
    L.logo2opt_tuning_shapes.argtypes = [i32p, C.c_uint32, C.POINTER(C.c_uint32)]
    L.logo2opt_tuning_set.argtypes = [C.c_int, u32p, i32p, C.c_uint32]
    L.logo2opt_tuning_get.argtypes = [C.c_int, u32p, i32p, C.c_uint32, C.POINTER(C.c_uint32)]
    L.logo2opt_tuning_load.argtypes = [C.c_int, C.c_char_p]
    L.logo2opt_engine_time_sweep.argtypes = [vp, C.c_uint32, C.POINTER(C.c_double)]
    _lib = L
    return L


def strerror(code):
    return load().logo2opt_strerror(code).decode()


def _check(rc, what):
    if rc != 0:
        raise Logo2OptError(rc, what)


# ---- workloads -------------------------------------------------------------
def gen_uniform(n, side, seed):
    xy = np.zeros(2 * n, np.int32)
    load().logo2opt_gen_uniform(n, side, seed, xy)
    return xy


def gen_clustered(n, side, seed, nclusters=60, sigma_frac=0.02, noise_frac=0.10):
    xy = np.zeros(2 * n, np.int32)
    load().logo2opt_gen_clustered(n, side, nclusters, sigma_frac, noise_frac, seed, xy)
    return xy


def random_perm(n, seed):
    r = np.zeros(n, np.uint32)
    load().logo2opt_random_perm(n, seed, r)
    return r


def nn_tour(xy, start=0):
    xy = np.ascontiguousarray(xy, np.int32)
    r = np.zeros(len(xy) // 2, np.uint32)
    _check(load().logo2opt_nn_tour(len(r), xy, start, r), "nn_tour")
    return r


def mf_tour(xy):
    """Multiple-fragment (greedy edge) tour of int32 interleaved xy."""
    xy = np.ascontiguousarray(xy, np.int32)
    r = np.zeros(len(xy) // 2, np.uint32)
    _check(load().logo2opt_mf_tour(len(r), xy, r), "mf_tour")
    return r


def coords_of(xy):
    """int32 interleaved xy -> array of reference city_coords (float x, y)."""
    xy = np.ascontiguousarray(xy, np.int32)
    n = len(xy) // 2
    c = (CityCoords * n)()
    load().logo2opt_xy_to_coords(n, xy, C.cast(c, vp))
    return c


def float_coords(x, y):
    n = len(x)
    c = (CityCoords * n)()
    arr = np.frombuffer(c, dtype=np.float32).reshape(n, 2)
    arr[:, 0] = x
    arr[:, 1] = y
    return c


def sqrt_ulp_check(lo_bits, hi_bits, device=0):
    """Max ulp distance of v_sqrt_f32 from the correctly rounded sqrt over the
    float bit patterns [lo_bits, hi_bits), measured on the GPU."""
    m = C.c_uint32()
    _check(load().logo2opt_sqrt_ulp_check(device, lo_bits, hi_bits, C.byref(m)), "sqrt_ulp_check")
    return m.value


# ---- TSPLIB files ------------------------------------------------------------
def tsp_read(path, round_coords=False):
    """Reads a TSPLIB .tsp file: returns (info dict, city_coords array).  The
    coordinates are translated so that the minimum is 0 (info offset_x/_y)."""
    L = load()
    info = TspInfo()
    flags = TSP_ROUND if round_coords else 0
    rc = L.logo2opt_tsp_read(os.fsencode(path), flags, C.byref(info), None, 0)
    _check(rc, f"tsp_read({path})")
    c = (CityCoords * info.n)()
    _check(L.logo2opt_tsp_read(os.fsencode(path), flags, C.byref(info), C.cast(c, vp), info.n), f"tsp_read({path})")
    d = dict(name=info.name.decode(errors="replace"), n=info.n, metric=info.metric, offset_x=info.offset_x,
             offset_y=info.offset_y, integral=bool(info.integral))
    return d, c


def tour_read(path, n):
    r = np.zeros(n, np.uint32)
    _check(load().logo2opt_tour_read(os.fsencode(path), n, r), f"tour_read({path})")
    return r


def tour_write(path, route, name="tour", length=0, headerless=False):
    route = np.ascontiguousarray(route, np.uint32)
    _check(load().logo2opt_tour_write(os.fsencode(path), name.encode(), len(route), route, length,
                                      1 if headerless else 0), f"tour_write({path})")


def host_route_length(coords, route, metric=METRIC_EUC_2D):
    route = np.ascontiguousarray(route, np.uint32)
    L = C.c_uint64()
    _check(load().logo2opt_host_route_length(len(route), C.cast(coords, vp), metric, route, C.byref(L)),
           "host_route_length")
    return int(L.value)


# ---- sweep tile shapes -------------------------------------------------------
def stop_below(solution, err):
    """The job field for the reference's ILS target rule (logo2opt_stop_below)."""
    return int(load().logo2opt_stop_below(solution, err))


def tuning_shapes():
    out = np.zeros(3 * 64, np.int32)
    cnt = C.c_uint32()
    _check(load().logo2opt_tuning_shapes(out, 64, C.byref(cnt)), "tuning_shapes")
    return [tuple(int(v) for v in out[3 * q:3 * q + 3]) for q in range(cnt.value)]


TABLE_SINGLE = 0  # fewer than 8 live instances in an engine
TABLE_BATCH = 1


def tuning_set(rows, table=TABLE_SINGLE):
    """rows: [(max_n, (k, t, waves)), ...]"""
    mx = np.array([r[0] for r in rows], np.uint32)
    sh = np.array([v for r in rows for v in r[1]], np.int32)
    _check(load().logo2opt_tuning_set(table, mx, sh, len(rows)), "tuning_set")


def tuning_get(table=TABLE_SINGLE):
    mx = np.zeros(64, np.uint32)
    sh = np.zeros(3 * 64, np.int32)
    cnt = C.c_uint32()
    _check(load().logo2opt_tuning_get(table, mx, sh, 64, C.byref(cnt)), "tuning_get")
    return [(int(mx[q]), tuple(int(v) for v in sh[3 * q:3 * q + 3])) for q in range(cnt.value)]


def tuning_load(path, table=TABLE_SINGLE):
    _check(load().logo2opt_tuning_load(table, path.encode()), "tuning_load")


# ---- single-instance context (the CUDASolver replacement) -----------------
class Context:
    def __init__(self, coords, n=None, device=0, metric=None):
        L = load()
        if n is None:
            n = len(coords)
        self.n = n
        self._coords = coords  # keep alive
        h = vp()
        if metric is None:
            _check(L.logo2opt_create(device, n, C.cast(coords, vp), C.byref(h)), "logo2opt_create")
        else:
            _check(L.logo2opt_create_metric(device, n, C.cast(coords, vp), metric, C.byref(h)),
                   "logo2opt_create_metric")
        self.h = h

    def close(self):
        if self.h:
            load().logo2opt_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, route, timing=False):
        route = np.ascontiguousarray(route, np.uint32)
        o = Best2Out()
        t = ProcessTime()
        _check(load().logo2opt_step(self.h, route, C.byref(o), C.byref(t) if timing else None), "logo2opt_step")
        out = (o.i, o.j, o.minchange)
        return (out, (t.HtD_time, t.kernel_time, t.DtH_time)) if timing else out

    def optimize(self, route, log_cap=0, sync_every=64):
        r = np.array(route, np.uint32)
        L = C.c_uint64()
        m = C.c_uint64()
        log = (Best2Out * max(1, log_cap))() if log_cap else None
        _check(load().logo2opt_optimize(self.h, r, C.byref(L), C.byref(m), C.cast(log, vp) if log else None,
                                        log_cap, sync_every), "logo2opt_optimize")
        out = dict(route=r, length=int(L.value), moves=int(m.value))
        if log_cap:
            k = min(out["moves"], log_cap)
            out["log"] = np.array([(log[q].i, log[q].j, log[q].minchange) for q in range(k)], np.int64).reshape(k, 3)
        return out

    def ils(self, route, seed, rounds, kick=KICK_DOUBLE_BRIDGE):
        r = np.array(route, np.uint32)
        best = C.c_uint64()
        rl = np.zeros(rounds + 1, np.uint64)
        _check(load().logo2opt_ils(self.h, r, seed, rounds, kick, C.byref(best), rl.ctypes.data),
               "logo2opt_ils")
        return r, int(best.value), rl

    def ils_until(self, route, seed, max_rounds, solution, err=0.0, time_limit_s=0.0, kick=KICK_DOUBLE_BRIDGE):
        """ILS with the reference's stop rules: returns (best tour, best length,
        round lengths of the rounds run, rounds run)."""
        r = np.array(route, np.uint32)
        best = C.c_uint64()
        done = C.c_uint32()
        rl = np.zeros(max_rounds + 1, np.uint64)
        _check(load().logo2opt_ils_until(self.h, r, seed, max_rounds, kick, solution, err, time_limit_s,
                                         C.byref(best), rl.ctypes.data, C.byref(done)), "logo2opt_ils_until")
        return r, int(best.value), rl[:done.value + 1], int(done.value)

    def ils_multistart(self, route, seed, rounds, chains, kick=KICK_DOUBLE_BRIDGE):
        r = np.array(route, np.uint32)
        best = C.c_uint64()
        rl = np.zeros(rounds + 1, np.uint64)
        _check(load().logo2opt_ils_multistart(self.h, r, seed, rounds, chains, kick, C.byref(best), rl.ctypes.data),
               "logo2opt_ils_multistart")
        return r, int(best.value), rl

    def route_length(self, route):
        L = C.c_uint64()
        _check(load().logo2opt_route_length(self.h, np.ascontiguousarray(route, np.uint32), C.byref(L)),
               "logo2opt_route_length")
        return int(L.value)

    def stats(self):
        s = Stats()
        _check(load().logo2opt_get_stats(self.h, C.byref(s)), "logo2opt_get_stats")
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def set_timing(self, on):
        _check(load().logo2opt_set_timing(self.h, 1 if on else 0), "logo2opt_set_timing")

    def set_pruning(self, on):
        """Skip blocks of rows the bounding-box bound rules out (same results)."""
        _check(load().logo2opt_set_pruning(self.h, _prune_mode(on)), "logo2opt_set_pruning")

    def time_sweep(self, route, sweeps):
        ms = C.c_double()
        _check(load().logo2opt_time_sweep(self.h, np.ascontiguousarray(route, np.uint32), sweeps, C.byref(ms)),
               "logo2opt_time_sweep")
        return ms.value


# ---- engine (batched instances, ILS, serving) ------------------------------
class Engine:
    def __init__(self, device=0, hbm_budget=1 << 30, max_slots=4096):
        h = vp()
        _check(load().logo2opt_engine_create(device, hbm_budget, max_slots, C.byref(h)), "logo2opt_engine_create")
        self.h = h
        self.max_slots = max_slots
        self._keep = {}

    def close(self):
        if self.h:
            load().logo2opt_engine_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def bytes_for(n, ils_rounds=0, log_cap=0):
        j = Job(ils_rounds, 0, 0, log_cap, 0)
        return int(load().logo2opt_engine_bytes_for(n, C.byref(j)))

    def add(self, coords, route, ils_rounds=0, kick=KICK_DOUBLE_BRIDGE, seed=0, log_cap=0, metric=METRIC_EUC_2D,
            stop_below=0):
        """Returns the slot id, or None when the engine is full (EAGAIN)."""
        route = np.ascontiguousarray(route, np.uint32)
        j = Job(ils_rounds, kick, seed, log_cap, metric, stop_below)
        s = C.c_int32()
        rc = load().logo2opt_engine_add(self.h, len(route), C.cast(coords, vp), route, C.byref(j), C.byref(s))
        if rc == -6:
            return None
        _check(rc, "logo2opt_engine_add")
        return s.value

    def add_group(self, coords, route, chains, ils_rounds, kick=KICK_DOUBLE_BRIDGE, seed=0, metric=METRIC_EUC_2D,
                  stop_below=0):
        """Multi-start ILS group; returns its first slot, or None when full."""
        route = np.ascontiguousarray(route, np.uint32)
        j = Job(ils_rounds, kick, seed, 0, metric, stop_below)
        s = C.c_int32()
        rc = load().logo2opt_engine_add_group(self.h, len(route), C.cast(coords, vp), route, C.byref(j), chains,
                                              C.byref(s))
        if rc == -6:
            return None
        _check(rc, "logo2opt_engine_add_group")
        return s.value

    def group_result(self, first_slot, n):
        res = Result()
        r = np.zeros(n, np.uint32)
        b = C.c_int32()
        _check(load().logo2opt_engine_group_result(self.h, first_slot, r.ctypes.data, C.byref(res), C.byref(b)),
               "logo2opt_engine_group_result")
        out = {f: getattr(res, f) for f, _ in Result._fields_}
        out.update(route=r, best_chain=b.value)
        return out

    def run(self, max_steps=0xFFFFFFFF, sync_every=32):
        live = C.c_uint32()
        _check(load().logo2opt_engine_run(self.h, max_steps, sync_every, C.byref(live)), "logo2opt_engine_run")
        return live.value

    def result(self, slot, n=None, log_cap=0):
        res = Result()
        r = np.zeros(n, np.uint32) if n else None
        log = (Best2Out * max(1, log_cap))() if log_cap else None
        _check(load().logo2opt_engine_result(self.h, slot, r.ctypes.data if n else None, C.byref(res),
                                             C.cast(log, vp) if log else None), "logo2opt_engine_result")
        out = {f: getattr(res, f) for f, _ in Result._fields_}
        if n:
            out["route"] = r
        if log_cap:
            k = min(out["moves"], log_cap)
            out["log"] = np.array([(log[q].i, log[q].j, log[q].minchange) for q in range(k)], np.int64).reshape(k, 3)
        return out

    def time_sweep(self, sweeps):
        ms = C.c_double()
        _check(load().logo2opt_engine_time_sweep(self.h, sweeps, C.byref(ms)), "engine_time_sweep")
        return ms.value

    def round_lengths(self, slot, cap):
        out = np.zeros(cap, np.uint64)
        _check(load().logo2opt_engine_round_lengths(self.h, slot, out, cap), "round_lengths")
        return out

    def remove(self, slot):
        _check(load().logo2opt_engine_remove(self.h, slot), "logo2opt_engine_remove")

    def poll_done(self, cap=4096):
        buf = (C.c_int32 * cap)()
        cnt = C.c_uint32()
        _check(load().logo2opt_engine_poll_done(self.h, C.cast(buf, vp), cap, C.byref(cnt)), "poll_done")
        return [buf[k] for k in range(cnt.value)]

    def set_timing(self, on):
        _check(load().logo2opt_engine_set_timing(self.h, 1 if on else 0), "logo2opt_engine_set_timing")

    def set_graphs(self, on):
        _check(load().logo2opt_engine_set_graphs(self.h, 1 if on else 0), "logo2opt_engine_set_graphs")

    def set_pruning(self, on):
        _check(load().logo2opt_engine_set_pruning(self.h, _prune_mode(on)), "logo2opt_engine_set_pruning")

    def stats(self):
        s = Stats()
        _check(load().logo2opt_engine_stats(self.h, C.byref(s)), "logo2opt_engine_stats")
        return {f: getattr(s, f) for f, _ in Stats._fields_}


# ---- server (continuous batching of requests) -------------------------------
SERVER_EXPORTS = ["logo2opt_server_create", "logo2opt_server_destroy", "logo2opt_server_submit",
                  "logo2opt_server_pump", "logo2opt_server_poll", "logo2opt_server_result",
                  "logo2opt_server_release", "logo2opt_server_counts", "logo2opt_server_replay",
                  "logo2opt_server_set_limits", "logo2opt_server_set_pruning", "logo2opt_server_set_order"]
EXPORTS += SERVER_EXPORTS


def _bind_server(L):
    if getattr(L, "_server_bound", False):
        return L
    L.logo2opt_server_create.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(vp)]
    L.logo2opt_server_destroy.argtypes = [vp]
    L.logo2opt_server_submit.argtypes = [vp, C.c_uint32, vp, vp, C.POINTER(Job), C.POINTER(C.c_int64)]
    L.logo2opt_server_pump.argtypes = [vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.logo2opt_server_poll.argtypes = [vp, vp, C.c_uint32, C.POINTER(C.c_uint32)]
    L.logo2opt_server_result.argtypes = [vp, C.c_int64, vp, C.POINTER(Result), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]
    L.logo2opt_server_release.argtypes = [vp, C.c_int64]
    L.logo2opt_server_counts.argtypes = [vp, u64p]
    L.logo2opt_server_set_limits.argtypes = [vp, C.c_uint32, C.c_uint64]
    L.logo2opt_server_set_pruning.argtypes = [vp, C.c_int]
    L.logo2opt_server_set_order.argtypes = [vp, C.c_int, C.c_uint32]
    L.logo2opt_server_replay.argtypes = [vp, C.c_uint32, vp, vp, vp, vp, vp, C.c_double, C.c_uint32, C.c_uint32,
                                         vp, vp, C.POINTER(C.c_double)]
    L._server_bound = True
    return L


class Server:
    """Continuous-batching request loop (logo2opt_server_*)."""

    def __init__(self, device=0, hbm_budget=4 << 30, max_slots=1024, queue_cap=0):
        L = _bind_server(load())
        h = vp()
        _check(L.logo2opt_server_create(device, hbm_budget, max_slots, queue_cap, C.byref(h)),
               "logo2opt_server_create")
        self.h = h

    def close(self):
        if self.h:
            load().logo2opt_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, coords, route, ils_rounds=0, kick=KICK_DOUBLE_BRIDGE, seed=0, metric=METRIC_EUC_2D, n=None,
               stop_below=0):
        """route None: the server builds the greedy NN tour from city 0 (give n).
        stop_below: the ILS target bound (see stop_below())."""
        j = Job(ils_rounds, kick, seed, 0, metric, stop_below)
        rid = C.c_int64()
        if route is None:
            rc = load().logo2opt_server_submit(self.h, n, C.cast(coords, vp), None, C.byref(j), C.byref(rid))
        else:
            route = np.ascontiguousarray(route, np.uint32)
            rc = load().logo2opt_server_submit(self.h, len(route), C.cast(coords, vp), route.ctypes.data, C.byref(j),
                                               C.byref(rid))
        if rc == -6:
            return None
        _check(rc, "logo2opt_server_submit")
        return rid.value

    def pump(self, steps=64, sync_every=16):
        run, q = C.c_uint32(), C.c_uint32()
        _check(load().logo2opt_server_pump(self.h, steps, sync_every, C.byref(run), C.byref(q)), "server_pump")
        return run.value, q.value

    def poll(self, cap=1024):
        buf = (C.c_int64 * cap)()
        cnt = C.c_uint32()
        _check(load().logo2opt_server_poll(self.h, C.cast(buf, vp), cap, C.byref(cnt)), "server_poll")
        return [buf[k] for k in range(cnt.value)]

    def result(self, rid, n):
        """A finished request's result; for a request rejected at admission
        {"error": code, "latency_s": ...} (e.g. -4, LOGO2OPT_ENOMEM: larger
        than the whole HBM budget)."""
        r = np.zeros(n, np.uint32)
        res = Result()
        lat, qs = C.c_double(), C.c_double()
        rc = load().logo2opt_server_result(self.h, rid, r.ctypes.data, C.byref(res), C.byref(lat), C.byref(qs))
        if rc in (-4, -2, -3):  # rejected at admission: ENOMEM, ECOORD, ETOUR
            return {"error": rc, "latency_s": lat.value}
        _check(rc, "server_result")
        out = {f: getattr(res, f) for f, _ in Result._fields_}
        out.update(route=r, latency_s=lat.value, queue_s=qs.value, error=0)
        return out

    def set_limits(self, max_running=0, max_inflight_pairs=0):
        _check(load().logo2opt_server_set_limits(self.h, max_running, max_inflight_pairs), "server_set_limits")

    def set_pruning(self, on):
        _check(load().logo2opt_server_set_pruning(self.h, _prune_mode(on)), "server_set_pruning")

    def set_order(self, shortest_first, max_bypass=0):
        """Admission order: arrival order (False) or fewest cities first
        (True), a request overtaken max_bypass times going next (0: no bound)."""
        _check(load().logo2opt_server_set_order(self.h, ORDER_SHORTEST if shortest_first else ORDER_FIFO,
                                                max_bypass), "server_set_order")

    def counts(self):
        c = np.zeros(6, np.uint64)
        _check(load().logo2opt_server_counts(self.h, c), "server_counts")
        return dict(zip(["submitted", "rejected", "admitted", "done", "queued", "running"], c.tolist()))

    def replay(self, arrivals, requests, horizon_s, steps=32, sync_every=16):
        """requests: list of (coords, route, Job), or (coords, None, Job, n) to
        let the server build the NN start tour.  Returns (latency[], queue[],
        elapsed)."""
        m = len(requests)
        arr = np.ascontiguousarray(arrivals, np.float64)
        ns = np.array([len(r[1]) if r[1] is not None else r[3] for r in requests], np.uint32)
        cptr = (vp * m)(*[C.cast(r[0], vp) for r in requests])
        routes = [np.ascontiguousarray(r[1], np.uint32) if r[1] is not None else None for r in requests]
        rptr = (vp * m)(*[r.ctypes.data if r is not None else None for r in routes])
        jobs = (Job * m)(*[r[2] for r in requests])
        lat = np.zeros(m, np.float64)
        qs = np.zeros(m, np.float64)
        el = C.c_double()
        _check(load().logo2opt_server_replay(self.h, m, arr.ctypes.data, ns.ctypes.data, C.cast(cptr, vp),
                                             C.cast(rptr, vp), C.cast(jobs, vp), horizon_s, steps, sync_every,
                                             lat.ctypes.data, qs.ctypes.data, C.byref(el)), "server_replay")
        return lat, qs, el.value
