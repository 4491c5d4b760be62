#!/usr/bin/env python
"""Benchmark of the MI355X 2-opt local-search path (one GPU).

Headline (`value`): 2-opt move evaluations per second of a whole best-
improvement local search on the GPU, on the 33,810-city clustered instance
(60 Gaussian clusters with sigma = 2% of the side plus 10% uniform noise in
[0,10^6)^2, greedy nearest-neighbour start) run to a 2-opt local optimum.
One "step" is one full local search from the same start tour.  The
coordinates are copied to the device once, before the timed region; each
step hands the library the 135 KB start tour, as TwoOptLocalSolver::optimize's
caller does, and the library rebuilds and uploads the tour-ordered point
array with its exact edge lengths (541 KB) before the first sweep -- this is
inside the timed region (about 1 ms of the ~0.3 s step).

Also reported on the same line:
  roofline      sweep kernel, 14 algorithmic FP32 FLOP per evaluated move,
                average kernel duration from HIP events on the engine stream
                over one timed local search (same trajectory as the steps)
  cpu_baseline  the reference's own CPU sweep (CPUSolver::optimizeStep built
                from its sources, oracle/_ref) timed on this host on a bounded
                sample of sweeps of the same instance and start tour
  batched       2,048 instances with n log-uniform in [100,5000], coords
                U{0..9999}^2, random starts, all to a local optimum in one
                engine (batched sweeps, graph-replayed steps) versus the plain
                one-instance-per-launch path (host round trip per move) on an
                every-16th subset; final lengths cross-checked between the two
  pruned, batched.pruned, serving_pruned
                the same local search, batch and serving replay with the sweep's
                row pruning on: separate figures (fewer candidates evaluated,
                identical results, asserted), never the headline
  serving_order the serving replay at 50 and 100 req/s with at most 4 running
                requests, admitted FIFO and fewest-cities-first (a request
                overtaken 16 times goes next); p50 ratio FIFO / shortest-first
  ils_100k      500 double-bridge ILS rounds on 100,000 uniform cities from the
                NN tour, full and pruned sweeps (--skip-config3 leaves it out)
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "logo-tsp-solver_amd"))

import logo2opt as L  # noqa: E402

METRIC = "2-opt move evals/sec + instances/sec batched vs plain, bit-exact tour length"
# algorithmic work of one move evaluation, the unit the roofline is priced in:
# 2 new edges x (2 sub + mul + fma(2) + sqrt + add); the kernel's own filter
# does less (integer L1 bounds, see l2o_kernels.hip) and is priced the same
FLOP_PER_EVAL = 14
# HBM bytes per sweep launch of the 33,810-city search, from two rocprofv3 --pmc runs of this
# benchmark (FETCH_SIZE 2,680.2 KiB x 2.0 + WRITE_SIZE 4.42 KiB x 1.0, factors from a 256 MiB copy;
# profiles/k_sweep_config2_counters.md).  Reported only for the kernel that was profiled.
PROFILED_TRAFFIC = {"l2o::k_sweep<4,256,2,0>": 2680.17 * 1024 * 2.0 + 4.41877 * 1024 * 1.0}
REF_FLOP_PER_EVAL = 36    # the reference's GPU GFLOPS convention (solver.c:305-306)
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector peak
# The sweep's own bound: VALU instruction issue.  Each of the 1,024 SIMDs
# (256 CUs x 4) issues one wave64 VALU instruction per 4 cycles at 2.4 GHz
# (the FP32 peak above is this rate x 64 lanes x 2 for FMA x 2 for packed
# FP32).  VALU instructions per candidate come from SQ_INSTS_VALU of the
# profiled kernel (per dispatch / candidates per dispatch, wave instructions).
VALU_ISSUE_PEAK = 1024 * 2.4e9 / 4
PROFILED_VALU_PER_EVAL = {"l2o::k_sweep<4,256,2,0>": 1.96806e7 / 571473528}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pairs(n):
    return (n - 2) * (n - 3) // 2


def sweep_kernel_name(n, table=None):
    """Name of the sweep kernel instance the library launches for size n."""
    rows = L.tuning_get(L.TABLE_SINGLE if table is None else table)
    k, t, w = next((sh for mx, sh in rows if n <= mx), rows[-1][1])
    return f"l2o::k_sweep<{k},{t},{w},0>"  # mode 0: the full sweep


def config2(seed=2):
    n = 33810
    xy = L.gen_clustered(n, 10**6, seed)
    r0 = L.nn_tour(xy, 0)
    return n, xy, r0


def exact_simd_port(n, xy, r0, budget_s):
    """The strongest CPU sweep here: the oracle's exact SIMD sweep
    (orc_step_fast: integer L1 first test over 16 or 8 columns, exact lengths
    for the survivors) on up to 16 threads, on the same tour."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    threads = min(16, len(os.sched_getaffinity(0)))
    xy = np.ascontiguousarray(xy, np.int32)
    sweeps, t0 = 0, time.perf_counter()
    while True:
        O.step_fast(xy, r0, threads)
        sweeps += 1
        el = time.perf_counter() - t0
        if el > budget_s or sweeps >= 400:
            break
    return {"value": sweeps * pairs(n) / el, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"{sweeps} exact SIMD sweeps ({O.fast_isa()}, {el:.1f} s, {threads} threads) of the "
                      "33,810-city clustered instance at its NN start tour; same move as the oracle"}


def cpu_baseline(n, xy, r0, budget_s):
    """Bounded samples of the reference's CPU sweep on this host: its pthreads
    CPUMTSolver::optimizeStep on up to 16 threads (the CPU share of one GPU
    here) as the baseline, and its single-thread CPUSolver::optimizeStep; the
    exact SIMD port beside them."""
    ref_avx = os.path.join(ROOT, "oracle", "_ref", "libref2opt_avx.so")
    ref_sc = os.path.join(ROOT, "oracle", "_ref", "libref2opt.so")
    path = ref_avx if os.path.exists(ref_avx) else ref_sc if os.path.exists(ref_sc) else None
    fxy = np.ascontiguousarray(xy.astype(np.float32))
    r0 = np.ascontiguousarray(r0, np.uint32)
    if path:
        R = C.CDLL(path)
        f32, u32, i32 = (np.ctypeslib.ndpointer(t) for t in (np.float32, np.uint32, np.int32))
        R.ref_step.argtypes = [f32, u32, C.c_uint32, i32]
        out = np.zeros(3, np.int32)

        def sample(step, cap, budget):
            sweeps, t0 = 0, time.perf_counter()
            while True:
                step()
                sweeps += 1
                el = time.perf_counter() - t0
                if el > budget or sweeps >= cap:
                    return sweeps, el

        variant = "AVX build (USE_CPU_SIMD, USE_AVX)" if path == ref_avx else "scalar build"
        s1, e1 = sample(lambda: R.ref_step(fxy, r0, n, out), 30, budget_s / 3)
        single = {"value": s1 * pairs(n) / e1, "cores": 1,
                  "sample": f"{s1} CPUSolver::optimizeStep sweeps ({e1:.1f} s), one thread as --mode=cpu"}
        if hasattr(R, "ref_step_mt"):
            R.ref_step_mt.argtypes = [f32, u32, C.c_uint32, C.c_int32, i32]
            threads = min(16, len(os.sched_getaffinity(0)))
            sm, em = sample(lambda: R.ref_step_mt(fxy, r0, n, threads, out), 200, budget_s / 3)
            return {"value": sm * pairs(n) / em, "unit": "evals/s", "cores": threads, "kind": "reference",
                    "sample": f"{sm} CPUMTSolver::optimizeStep sweeps ({em:.1f} s, {threads} pthreads as "
                              f"--cores={threads}) of the 33,810-city clustered instance at its NN start tour, "
                              f"reference sources {variant}",
                    "single_thread": single, "exact_simd_port": exact_simd_port(n, xy, r0, budget_s / 3)}
        single.update({"unit": "evals/s", "kind": "reference",
                       "sample": single["sample"] + f" of the 33,810-city instance, reference sources {variant}"})
        return single
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    threads = len(os.sched_getaffinity(0))
    sweeps, t0 = 0, time.perf_counter()
    while True:
        O.step(xy, r0, threads)
        sweeps += 1
        el = time.perf_counter() - t0
        if el > budget_s or sweeps >= 50:
            break
    return {"value": sweeps * pairs(n) / el, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"{sweeps} exact oracle sweeps ({el:.1f} s) of the 33,810-city clustered instance"}


def batched_vs_plain(count, plain_every, seed=4, prune=True):
    rs = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rs.uniform(np.log(100), np.log(5000), count))).astype(int)
    insts = []
    for k, n in enumerate(sizes):
        xy = L.gen_uniform(int(n), 10**4, 10_000 + k)
        insts.append((int(n), L.coords_of(xy), L.random_perm(int(n), 20_000 + k)))
    budget = sum(L.Engine.bytes_for(n) for n, _, _ in insts) + (64 << 20)
    with L.Engine(hbm_budget=budget, max_slots=count) as eng:
        slots = [eng.add(c, r) for n, c, r in insts]
        assert all(s is not None for s in slots)
        t0 = time.perf_counter()
        live = eng.run(sync_every=32)
        tb = time.perf_counter() - t0
        assert live == 0
        st = eng.stats()
        lengths = [eng.result(s)["length"] for s in slots]
    tbp = None
    if prune:
        # the same batch with row pruning on every instance: a separate
        # figure (fewer candidates evaluated), lengths asserted equal
        with L.Engine(hbm_budget=budget, max_slots=count) as eng:
            eng.set_pruning(L.PRUNE_ALL)
            slots = [eng.add(c, r) for n, c, r in insts]
            t0 = time.perf_counter()
            assert eng.run(sync_every=32) == 0
            tbp = time.perf_counter() - t0
            assert [eng.result(s)["length"] for s in slots] == lengths
    # the plain path, one instance at a time, on every plain_every-th
    # instance: with a host round trip per move (the reference's loop,
    # sync_every = 1) and with 64 graph-replayed steps between host checks
    # (the single-instance path of the headline figure)
    subset = list(range(0, count, plain_every))
    ctxs = [(k, L.Context(insts[k][1])) for k in subset]
    tp = tp64 = 0.0
    mism = 0
    for k, ctx in ctxs:
        t0 = time.perf_counter()
        out = ctx.optimize(insts[k][2], sync_every=1)
        tp += time.perf_counter() - t0
        t0 = time.perf_counter()
        out64 = ctx.optimize(insts[k][2], sync_every=64)
        tp64 += time.perf_counter() - t0
        mism += (out["length"] != lengths[k]) + (out64["length"] != lengths[k])
        ctx.close()
    ib = count / tb
    ip = len(subset) / tp
    ip64 = len(subset) / tp64
    extra = {} if tbp is None else {"pruned": {"batched_s": tbp, "batched_instances_per_s": count / tbp,
                                               "speedup_vs_plain": count / tbp / ip,
                                               "speedup_vs_plain_graph": count / tbp / ip64, "same_lengths": True}}
    return {
        "workload": f"{count} instances, n log-uniform in [100,5000] (seed {seed}), coords U{{0..9999}}^2, "
                    "random starts, each to a 2-opt local optimum",
        "instances": count, "batched_s": tb, "batched_instances_per_s": ib,
        "batched_evals_per_s": st["evals"] / tb, "moves": st["moves"], "sweeps": st["sweeps"],
        "plain_sample": len(subset), "plain_s": tp, "plain_instances_per_s": ip,
        "speedup": ib / ip, "plain_graph_s": tp64, "plain_graph_instances_per_s": ip64,
        "speedup_vs_plain_graph": ib / ip64, "plain_vs_batched_length_mismatches": int(mism), **extra,
    }


def serving(rates, window_s, horizon_s, pool=48, seed=5, rounds=50, max_slots=512, max_running=0,
            max_inflight_pairs=0, prune=False, order=None):
    """Open-loop Poisson replay through the continuous-batching server (the
    same seeded pool and arrivals for every call with the same arguments).
    order: None admits in arrival order; an integer admits the fewest cities
    first with that overtaking bound (max_bypass, 0 = none)."""
    rs = np.random.default_rng(seed)
    insts, tcon = [], []
    for k in range(pool):
        n = int(np.rint(np.exp(rs.uniform(np.log(200), np.log(50000)))))
        xy = L.gen_clustered(n, 10**6, 5000 + k) if rs.random() < 0.5 else L.gen_uniform(n, 10**6, 5000 + k)
        t0 = time.perf_counter()
        L.nn_tour(xy, 0)  # what the server does per request (its time is part of each latency)
        tcon.append(time.perf_counter() - t0)
        insts.append((n, L.coords_of(xy)))
    out = []
    for lam in rates:
        gaps = rs.exponential(1.0 / lam, int(lam * window_s * 1.5) + 16)
        arrivals = np.cumsum(gaps)
        arrivals = arrivals[arrivals <= window_s]
        pick = rs.integers(0, pool, len(arrivals))
        # no start tour: the server builds each request's NN tour at admission
        reqs = [(insts[p][1], None, L.Job(rounds, 0, 7000 + q, 0, 0), insts[p][0]) for q, p in enumerate(pick)]
        with L.Server(hbm_budget=8 << 30, max_slots=max_slots, queue_cap=1 << 20) as srv:
            srv.set_limits(max_running, max_inflight_pairs)
            srv.set_pruning(prune)
            if order is not None:
                srv.set_order(True, max_bypass=order)
            lat, qs, el = srv.replay(arrivals, reqs, horizon_s, steps=32, sync_every=16)
            cnt = srv.counts()
        ok = lat >= 0
        done_n = [insts[p][0] for p in pick[ok]]
        # every arrived request counts: an unfinished one with (end of the run
        # - arrival), a lower bound of its latency
        cens = np.where(ok, lat, el - arrivals)
        p50, p99 = (float(np.percentile(cens, q)) if len(cens) else None for q in (50, 99))
        # exact when every censored value lies above it, else a lower bound
        cmin = float(np.min(cens[~ok])) if (~ok).any() else None
        lb = {q: cmin is not None and v is not None and cmin <= v for q, v in ((50, p50), (99, p99))}
        out.append({
            "rate_req_per_s": lam, "arrived": int(len(arrivals)), "completed": int(ok.sum()),
            "unfinished": int((~ok).sum()),
            "throughput_req_per_s": float(ok.sum() / el), "elapsed_s": el,
            "p50_latency_s": p50, "p99_latency_s": p99,
            "p50_is_lower_bound": lb[50], "p99_is_lower_bound": lb[99],
            "p50_latency_completed_s": float(np.percentile(lat[ok], 50)) if ok.any() else None,
            "p99_latency_completed_s": float(np.percentile(lat[ok], 99)) if ok.any() else None,
            "p50_queue_s": float(np.percentile(qs[ok], 50)) if ok.any() else None,
            "mean_n_completed": float(np.mean(done_n)) if done_n else None,
            "mean_construction_s": float(np.mean([tcon[p] for p in pick])) if len(pick) else None,
            "admitted": int(cnt["admitted"]), "rejected": int(cnt["rejected"]),
            "left_running": int(cnt["running"]), "left_queued": int(cnt["queued"]),
        })
        log(f"serving lambda={lam}: {out[-1]}")
    # the knee: the highest offered rate the server keeps up with (completes
    # at least 90% of what arrived within the horizon), and the saturation
    # throughput (the most completed requests per second at any rate)
    keep = [r["rate_req_per_s"] for r in out if r["completed"] >= 0.9 * r["arrived"]]
    return {"workload": f"ILS requests ({rounds} double-bridge rounds, greedy NN start built by the server at "
                        "admission and counted in the latency), n log-uniform in "
                        f"[200,50000], uniform or clustered coords in [0,1e6)^2 (pool of {pool} instances, "
                        f"distinct kick seeds), Poisson arrivals for {window_s}s, horizon {horizon_s}s",
            "latency_method": "p50/p99 over every arrived request; a request unfinished at the end of the run "
                              "counts with (end - arrival), a lower bound (flagged *_is_lower_bound when such a "
                              "value reaches the percentile); *_completed_s are over completed requests only",
            "max_slots": max_slots, "max_running": max_running, "max_inflight_pairs": max_inflight_pairs,
            "order": "fifo" if order is None else f"shortest-first, max_bypass {order}", "pruned": bool(prune),
            "knee_req_per_s": max(keep) if keep else None,
            "saturation_throughput_req_per_s": max(r["throughput_req_per_s"] for r in out) if out else None,
            "rates": out}


def config3(rounds, start, seed=3, prune=True):
    """ILS on 100,000 uniform cities: LO, then `rounds` double-bridge kicks."""
    n = 100000
    xy = L.gen_uniform(n, 10**6, seed)
    r0 = L.nn_tour(xy, 0) if start == "nn" else L.random_perm(n, seed + 1)
    with L.Context(L.coords_of(xy)) as ctx:
        t0 = time.perf_counter()
        r, best, rl = ctx.ils(r0, seed=seed + 2, rounds=rounds)
        dt = time.perf_counter() - t0
        st = ctx.stats()
        out = {"workload": f"ILS on 100,000 uniform cities in [0,1e6)^2, {start} start, {rounds} seeded "
                           "double-bridge rounds, strict accept", "seconds": dt, "best_length": best,
               "length_after_first_lo": int(rl[0]), "moves": st["moves"], "sweeps": st["sweeps"],
               "evals_per_s": st["evals"] / dt, "fallbacks": st["fallbacks"]}
        if prune:
            # the same ILS with row pruning on: same tours, separate figure
            ctx.set_pruning(True)
            t0 = time.perf_counter()
            rp, bp, rlp = ctx.ils(r0, seed=seed + 2, rounds=rounds)
            dtp = time.perf_counter() - t0
            assert bp == best and rlp.tolist() == rl.tolist() and rp.tolist() == r.tolist()
            out["pruned"] = {"seconds": dtp, "speedup_vs_full": dt / dtp, "same_result": True}
    return out


def copy_calibration(nbytes):
    """One device-to-device hipMemcpy of nbytes (a plain copy kernel reading
    and writing nbytes each): the yardstick for FETCH_SIZE / WRITE_SIZE."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes}), flush=True)


PLA33810_OPT = 66048945  # TSPLIB known optimum (src/utils/cut.c:299-300)


def tsplib_pla33810(ils_chains, ils_rounds):
    """The TSPLIB instance of config 2's size, pla33810 (CEIL_2D, shipped in the
    reference's TSPLIB/ and committed as tests/golden/tsplib/pla33810.tsp.xz):
    local search from the nearest-neighbour and multiple-fragment tours, then a
    short multi-start ILS with a shared best from the MF optimum."""
    import lzma
    import tempfile

    with open(os.path.join(ROOT, "tests", "golden", "tsplib", "pla33810.tsp.xz"), "rb") as f:
        data = lzma.decompress(f.read())
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "pla33810.tsp")
        with open(p, "wb") as f:
            f.write(data)
        info, coords = L.tsp_read(p)
    n, metric = info["n"], info["metric"]
    xy = np.ascontiguousarray(np.frombuffer(coords, dtype=np.float32).astype(np.int32))
    out = {"instance": "pla33810 (TSPLIB, CEIL_2D)", "n": n, "known_optimum": PLA33810_OPT}
    with L.Context(coords, metric=metric) as ctx:
        for name, build in (("nn", lambda: L.nn_tour(xy, 0)), ("mf", lambda: L.mf_tour(xy))):
            t0 = time.perf_counter()
            start = build()
            tc = time.perf_counter() - t0
            s0 = ctx.stats()
            t0 = time.perf_counter()
            lo = ctx.optimize(start, sync_every=64)
            dt = time.perf_counter() - t0
            s1 = ctx.stats()
            out[name] = {"construction_s": tc, "start_length": L.host_route_length(coords, start, metric),
                         "local_search_s": dt, "moves": lo["moves"], "length": lo["length"],
                         "gap_pct": 100.0 * (lo["length"] / PLA33810_OPT - 1.0),
                         "evals_per_s": (s1["evals"] - s0["evals"]) / dt}
            if name == "mf":
                mf_lo = lo["route"]
        t0 = time.perf_counter()
        r, best, rl = ctx.ils_multistart(mf_lo, 1, ils_rounds, ils_chains)
        dt = time.perf_counter() - t0
        out["ils_multistart"] = {"chains": ils_chains, "rounds": ils_rounds, "seconds": dt, "length": best,
                                 "gap_pct": 100.0 * (best / PLA33810_OPT - 1.0),
                                 "valid": L.host_route_length(coords, r, metric) == best}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--batch-count", type=int, default=2048)
    ap.add_argument("--plain-every", type=int, default=16)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--serve-rates", default="10,20,50,100,200,500,1000,2000")
    ap.add_argument("--serve-window", type=float, default=2.0)
    ap.add_argument("--serve-horizon", type=float, default=6.0)
    ap.add_argument("--serve-max-running", type=int, default=0)
    ap.add_argument("--serve-max-pairs", type=float, default=0, help="cap on in-flight candidates per sweep")
    ap.add_argument("--no-serve-prune", action="store_true",
                    help="skip the replay of the same arrivals with row pruning on (serving_pruned)")
    ap.add_argument("--skip-order", action="store_true",
                    help="skip the FIFO vs shortest-first admission comparison (serving_order)")
    ap.add_argument("--order-rates", default="50,100")
    ap.add_argument("--order-max-running", type=int, default=4)
    ap.add_argument("--order-bypass", type=int, default=16)
    ap.add_argument("--skip-tsplib", action="store_true")
    ap.add_argument("--skip-pruned", action="store_true")
    ap.add_argument("--tsplib-chains", type=int, default=16)
    ap.add_argument("--tsplib-rounds", type=int, default=100)
    ap.add_argument("--config3", action="store_true", help="run the 100k-city ILS (the default; kept for old command lines)")
    ap.add_argument("--skip-config3", action="store_true", help="skip the 100k-city ILS (about 10 s)")
    ap.add_argument("--config3-rounds", type=int, default=500)
    ap.add_argument("--config3-start", default="nn", choices=["nn", "random"])
    ap.add_argument("--no-graph", action="store_true", help="direct launches instead of HIP graph replay")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration << 20)
        return
    if args.no_graph:
        os.environ["LOGO2OPT_GRAPHS"] = "0"
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (--gpus {args.gpus} requested); the sweep's "
              "multi-GPU sharding (pair triangle split, one 8-byte min-allreduce per move) is not run.",
              file=sys.stderr)
        sys.exit(2)

    n, xy, r0 = config2()
    ctx = L.Context(L.coords_of(xy))
    log(f"config2 n={n}: warmup {args.warmup}")
    for _ in range(args.warmup):
        ref = ctx.optimize(r0, sync_every=64)
    s0 = ctx.stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = ctx.optimize(r0, sync_every=64)
    dt = time.perf_counter() - t0
    s1 = ctx.stats()
    if args.warmup:
        assert out["length"] == ref["length"] and out["moves"] == ref["moves"]
    evals = s1["evals"] - s0["evals"]
    sweeps_per = (s1["sweeps"] - s0["sweeps"]) // args.steps
    fb_per = (s1["fallbacks"] - s0["fallbacks"]) // args.steps
    log(f"timed {args.steps} local searches: {dt:.3f}s, {out['moves']} moves, {sweeps_per} sweeps each")

    # roofline: the same local search again with every sweep launch event-timed
    ctx.set_timing(True)
    a0 = ctx.stats()
    out_t = ctx.optimize(r0, sync_every=64)
    a1 = ctx.stats()
    ctx.set_timing(False)
    assert out_t["length"] == out["length"]
    launches = a1["sweep_launches"] - a0["sweep_launches"]
    avg_ms = (a1["sweep_ms"] - a0["sweep_ms"]) / launches
    work_sweeps = a1["sweeps"] - a0["sweeps"]
    # algorithmic FLOP per launch averaged over launches (launches after
    # convergence inside a sync batch do no work and are counted too)
    flop_per_launch = FLOP_PER_EVAL * pairs(n) * work_sweeps / launches
    achieved = flop_per_launch / (avg_ms * 1e-3) / 1e12
    kname = sweep_kernel_name(n)
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": PROFILED_TRAFFIC.get(kname),
                "kernel": kname, "avg_kernel_ms": avg_ms, "launches": launches,
                "evals_per_s_kernel": pairs(n) * work_sweeps / launches / (avg_ms * 1e-3),
                "ref_convention_tflops": achieved * REF_FLOP_PER_EVAL / FLOP_PER_EVAL}
    vpe = PROFILED_VALU_PER_EVAL.get(kname)
    if vpe:
        rate = roofline["evals_per_s_kernel"] * vpe
        roofline["valu_issue"] = {
            "achieved": rate, "peak": VALU_ISSUE_PEAK, "unit": "wave64 VALU instructions/s",
            "frac": rate / VALU_ISSUE_PEAK, "valu_per_eval_x64": vpe * 64,
            "note": "the kernel's integer VALU issue rate (SQ_INSTS_VALU per candidate from the profiled "
                    "counters x candidates/s in the kernel) against the issue peak; frac above is the 14-FLOP "
                    "contract figure"}
    ctx.close()

    pruned = None
    if not args.skip_pruned:
        # the same local searches with row pruning on: a separate figure,
        # never the headline (it evaluates fewer candidates per sweep)
        ctx = L.Context(L.coords_of(xy))
        ctx.set_pruning(True)
        for _ in range(args.warmup):
            ctx.optimize(r0, sync_every=64)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out_p = ctx.optimize(r0, sync_every=64)
        dtp = time.perf_counter() - t0
        ctx.close()
        same = (out_p["length"], out_p["moves"], out_p["route"].tolist()) == \
            (out["length"], out["moves"], out["route"].tolist())
        assert same, "pruned local search differs from the full one"
        pruned = {"ms_per_search": dtp / args.steps * 1e3, "speedup_vs_full": dt / dtp, "same_result": same,
                  "note": "row pruning on (bounding-box bound skips blocks of rows); not the headline value"}
        log(f"pruned: {dtp / args.steps * 1e3:.1f} ms per search ({dt / dtp:.2f}x)")

    res = {
        "metric": METRIC, "value": evals / dt, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "i32", "data": "synthetic",
        "dtype_note": "the sweep's hot loop is integer VALU (v_sad_u8 on 8-bit coordinates for every candidate, "
                      "v_sad_u32 L1 bounds for the blocks that pass); FP32 sqrt re-checks the rare survivors and "
                      "FP64 + integer fix-up gives their exact lengths; lengths are exact integers",
        "config": {"workload": "33,810 clustered cities (60 Gaussian clusters, sigma 2% of side, + 10% uniform "
                               "noise) in [0,1e6)^2, greedy NN start, best-improvement 2-opt to a local optimum",
                   "n": n, "seed": 2, "moves": out["moves"], "sweeps": sweeps_per, "length": out["length"],
                   "fallbacks_per_search": fb_per, "parallelism": "single GPU"},
        "roofline": roofline,
    }
    if pruned:
        res["pruned"] = pruned
    if not args.skip_cpu:
        log("cpu baseline")
        res["cpu_baseline"] = cpu_baseline(n, xy, r0, args.cpu_budget)
    if not args.skip_batch:
        log("batched vs plain")
        res["batched"] = batched_vs_plain(args.batch_count, args.plain_every, prune=not args.skip_pruned)
    if not args.skip_serve:
        log("serving")
        rates = [float(v) for v in args.serve_rates.split(",")]
        res["serving"] = serving(rates, args.serve_window, args.serve_horizon, max_running=args.serve_max_running,
                                 max_inflight_pairs=int(args.serve_max_pairs))
        if not args.no_serve_prune:
            log("serving, row pruning on")
            res["serving_pruned"] = serving(rates, args.serve_window, args.serve_horizon,
                                            max_running=args.serve_max_running,
                                            max_inflight_pairs=int(args.serve_max_pairs), prune=True)
    if not args.skip_serve and not args.skip_order:
        # admission order under an admission cap: the same arrivals admitted
        # FIFO and fewest-cities-first (bounded overtaking)
        log("serving, FIFO vs shortest-first under a cap on running requests")
        orates = [float(v) for v in args.order_rates.split(",")]
        kw = dict(max_running=args.order_max_running, prune=not args.no_serve_prune)
        fifo = serving(orates, args.serve_window, args.serve_horizon, **kw)
        sjf = serving(orates, args.serve_window, args.serve_horizon, order=args.order_bypass, **kw)
        res["serving_order"] = {
            "fifo": fifo, "shortest_first": sjf,
            "p50_ratio_fifo_over_shortest": [a["p50_latency_s"] / b["p50_latency_s"]
                                             for a, b in zip(fifo["rates"], sjf["rates"])],
        }
    if not args.skip_tsplib:
        log("TSPLIB pla33810")
        res["tsplib"] = tsplib_pla33810(args.tsplib_chains, args.tsplib_rounds)
    if not args.skip_config3:
        log("config3 ILS 100k")
        res["ils_100k"] = config3(args.config3_rounds, args.config3_start, prune=not args.skip_pruned)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
