"""GPU tests of the batched engine and the continuous-batching server:
results are identical alone and in a mix of concurrently admitted requests
(tolerance 0), ILS results equal the oracle's, admission respects the HBM
budget, and requests admitted while others run finish correctly."""
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


def make_requests(count, lo, hi, seed, rounds):
    rs = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(np.rint(np.exp(rs.uniform(np.log(lo), np.log(hi)))))
        xy = L.gen_clustered(n, 10**6, 100 + k) if k % 2 else L.gen_uniform(n, 10**6, 100 + k)
        out.append(dict(n=n, xy=xy, coords=L.coords_of(xy), route=L.nn_tour(xy, 0), seed=1000 + k, rounds=rounds))
    return out


def alone(req):
    with L.Context(req["coords"]) as ctx:
        r, best, rl = ctx.ils(req["route"], seed=req["seed"], rounds=req["rounds"])
    return r, best


def test_server_staggered_admission_equals_alone():
    reqs = make_requests(10, 200, 3000, 5, rounds=6)
    with L.Server(hbm_budget=512 << 20, max_slots=16) as srv:
        ids = []
        for q, rq in enumerate(reqs):
            ids.append(srv.submit(rq["coords"], rq["route"], ils_rounds=rq["rounds"], seed=rq["seed"]))
            if q % 3 == 2:
                srv.pump(steps=24, sync_every=8)  # later requests join while these run
        done = {}
        for _ in range(100000):
            running, queued = srv.pump(steps=64, sync_every=16)
            for rid in srv.poll():
                done[rid] = True
            if running == 0 and queued == 0:
                break
        assert len(done) == len(reqs)
        for rid, rq in zip(ids, reqs):
            got = srv.result(rid, rq["n"])
            r, best = alone(rq)
            assert got["length"] == best and got["route"].tolist() == r.tolist()
            assert got["rounds_done"] == rq["rounds"]
            assert O.route_length(rq["xy"], got["route"]) == got["length"]
        c = srv.counts()
        assert c["done"] == len(reqs) and c["rejected"] == 0


def test_server_ils_matches_oracle():
    reqs = make_requests(6, 120, 400, 9, rounds=8)
    with L.Server(hbm_budget=64 << 20, max_slots=8) as srv:
        ids = [srv.submit(rq["coords"], rq["route"], ils_rounds=rq["rounds"], seed=rq["seed"]) for rq in reqs]
        while True:
            running, queued = srv.pump(steps=64, sync_every=16)
            if running == 0 and queued == 0:
                break
        for rid, rq in zip(ids, reqs):
            got = srv.result(rid, rq["n"])
            best, Lb, rl, _ = O.ils(rq["xy"], rq["route"], seed=rq["seed"], rounds=rq["rounds"], threads=THREADS)
            assert got["length"] == Lb and got["route"].tolist() == best.tolist()


def test_server_admission_by_hbm_budget_queues_the_rest():
    reqs = make_requests(6, 1500, 1600, 13, rounds=2)
    per = max(L.Engine.bytes_for(rq["n"], ils_rounds=2) for rq in reqs)
    with L.Server(hbm_budget=2 * per + 4096, max_slots=8) as srv:
        ids = [srv.submit(rq["coords"], rq["route"], ils_rounds=rq["rounds"], seed=rq["seed"]) for rq in reqs]
        running, queued = srv.pump(steps=1, sync_every=1)
        assert running <= 2 and queued >= len(reqs) - 2
        while True:
            running, queued = srv.pump(steps=64, sync_every=16)
            assert running <= 2
            if running == 0 and queued == 0:
                break
        lat = [srv.result(rid, rq["n"])["queue_s"] for rid, rq in zip(ids, reqs)]
        assert max(lat) > 0.0
        for rid, rq in zip(ids[:2], reqs[:2]):
            r, best = alone(rq)
            assert srv.result(rid, rq["n"])["length"] == best


def _serve_one_at_a_time(reqs, order):
    """Submit every request, then serve them one at a time (max_running 1), so
    completions come in admission order.  order: None (FIFO) or max_bypass for
    shortest-first.  Returns the request indices in completion order and the
    results by index."""
    with L.Server(hbm_budget=256 << 20, max_slots=8) as srv:
        srv.set_limits(max_running=1)
        if order is not None:
            srv.set_order(True, max_bypass=order)
        ids = [srv.submit(rq["coords"], rq["route"], ils_rounds=rq["rounds"], seed=rq["seed"]) for rq in reqs]
        pos = {rid: q for q, rid in enumerate(ids)}
        seen = []
        while True:
            running, queued = srv.pump(steps=64, sync_every=16)
            assert running <= 1
            seen += [pos[rid] for rid in srv.poll()]
            if running == 0 and queued == 0:
                break
        return seen, {q: srv.result(ids[q], rq["n"]) for q, rq in enumerate(reqs)}


def test_server_shortest_first_order_and_bypass_bound():
    sizes = [2400, 300, 1200, 150, 600]
    reqs = []
    for k, n in enumerate(sizes):
        xy = L.gen_uniform(n, 10**6, 300 + k)
        reqs.append(dict(n=n, xy=xy, coords=L.coords_of(xy), route=L.nn_tour(xy, 0), seed=2000 + k, rounds=2))
    fifo, res_fifo = _serve_one_at_a_time(reqs, None)
    assert fifo == [0, 1, 2, 3, 4]
    sjf, res_sjf = _serve_one_at_a_time(reqs, 0)
    assert [sizes[q] for q in sjf] == sorted(sizes)
    # bound 2: the 2,400-city head is overtaken by 150 and 300, then goes next;
    # 1,200 is then overtaken by 600 (its second time) and goes last
    bounded, res_bounded = _serve_one_at_a_time(reqs, 2)
    assert [sizes[q] for q in bounded] == [150, 300, 2400, 600, 1200]
    for q, rq in enumerate(reqs):
        r, best = alone(rq)
        for res in (res_fifo, res_sjf, res_bounded):
            assert res[q]["length"] == best and res[q]["route"].tolist() == r.tolist()
    with L.Server(hbm_budget=16 << 20, max_slots=2) as srv:
        assert L.load().logo2opt_server_set_order(srv.h, 7, 0) == -1  # LOGO2OPT_EINVAL


def test_server_queue_cap_rejects():
    reqs = make_requests(5, 200, 300, 17, rounds=1)
    with L.Server(hbm_budget=64 << 20, max_slots=4, queue_cap=2) as srv:
        got = [srv.submit(rq["coords"], rq["route"], ils_rounds=1, seed=1) for rq in reqs]
        assert got[2:] == [None, None, None] and all(g is not None for g in got[:2])
        assert srv.counts()["rejected"] == 3


def test_server_rejects_request_larger_than_budget_and_serves_the_rest():
    """A request that needs more than the whole HBM budget is rejected at
    admission with LOGO2OPT_ENOMEM, reported by poll, and does not block the
    requests queued behind it (FIFO admission); the replay loop ends."""
    small = make_requests(3, 200, 400, 23, rounds=2)
    big = make_requests(1, 6000, 6001, 29, rounds=2)[0]
    budget = 3 * max(L.Engine.bytes_for(rq["n"], ils_rounds=2) for rq in small) + 4096
    assert L.Engine.bytes_for(big["n"], ils_rounds=2) > budget
    with L.Server(hbm_budget=budget, max_slots=8) as srv:
        rid_big = srv.submit(big["coords"], big["route"], ils_rounds=2, seed=big["seed"])
        ids = [srv.submit(rq["coords"], rq["route"], ils_rounds=2, seed=rq["seed"]) for rq in small]
        done = []
        for _ in range(10000):
            srv.pump(steps=64, sync_every=16)
            done += srv.poll()
            if len(done) == 4:
                break
        assert sorted(done) == sorted([rid_big] + ids)
        assert srv.result(rid_big, big["n"])["error"] == -4  # LOGO2OPT_ENOMEM
        assert srv.counts()["rejected"] == 1
        for rid, rq in zip(ids, small):
            r, best = alone(rq)
            assert srv.result(rid, rq["n"])["length"] == best
    # the open-loop replay finishes with the rejected request marked -1
    with L.Server(hbm_budget=budget, max_slots=8) as srv:
        reqs = [big] + small
        jobs = [L.Job(2, 0, rq["seed"], 0, 0) for rq in reqs]
        lat, qs, el = srv.replay(np.arange(len(reqs)) * 0.001, [(rq["coords"], rq["route"], j)
                                                               for rq, j in zip(reqs, jobs)],
                                 horizon_s=120.0, steps=32, sync_every=16)
    assert lat[0] == -1.0 and (lat[1:] >= 0).all() and el < 60.0


def test_server_mix_up_to_50000_clustered_cities():
    """Serving at the top of the serving workload's size range: a 50,000-city
    clustered ILS request whose NN start the server builds itself, admitted
    with small requests arriving around it; every result is bit-identical to
    the same request run alone, and the small ones equal the oracle's ILS."""
    big_xy = L.gen_clustered(50000, 10**6, 77)
    big = dict(n=50000, xy=big_xy, coords=L.coords_of(big_xy), route=None, seed=4242, rounds=3)
    small = make_requests(4, 300, 2000, 31, rounds=3)
    with L.Server(hbm_budget=1 << 30, max_slots=16) as srv:
        rid_big = srv.submit(big["coords"], None, ils_rounds=3, seed=big["seed"], n=big["n"])
        ids = [srv.submit(small[0]["coords"], small[0]["route"], ils_rounds=3, seed=small[0]["seed"])]
        srv.pump(steps=8, sync_every=8)
        ids += [srv.submit(rq["coords"], rq["route"], ils_rounds=3, seed=rq["seed"]) for rq in small[1:]]
        done = []
        while len(done) < 5:
            srv.pump(steps=64, sync_every=16)
            done += srv.poll()
        got_big = srv.result(rid_big, big["n"])
        got = [srv.result(rid, rq["n"]) for rid, rq in zip(ids, small)]
    r0 = L.nn_tour(big_xy, 0)
    with L.Context(big["coords"]) as ctx:
        r, best, rl = ctx.ils(r0, seed=big["seed"], rounds=3)
    assert got_big["length"] == best and got_big["route"].tolist() == r.tolist()
    say(f"50,000 clustered cities served: {best}")
    for g, rq in zip(got, small):
        r, best = alone(rq)
        assert g["length"] == best and g["route"].tolist() == r.tolist()
        rb, Lb, _, _ = O.ils(np.ascontiguousarray(rq["xy"], np.int32), rq["route"], rq["seed"], rq["rounds"],
                             threads=THREADS)
        assert Lb == best


def test_server_open_loop_replay():
    reqs = make_requests(16, 200, 2500, 21, rounds=3)
    rs = np.random.default_rng(3)
    arrivals = np.cumsum(rs.exponential(1 / 200.0, len(reqs)))
    jobs = [L.Job(rq["rounds"], 0, rq["seed"], 0, 0) for rq in reqs]
    with L.Server(hbm_budget=1 << 30, max_slots=64) as srv:
        lat, qs, el = srv.replay(arrivals, [(rq["coords"], rq["route"], j) for rq, j in zip(reqs, jobs)],
                                 horizon_s=120.0, steps=32, sync_every=16)
    say("replay latencies (s):", np.round(lat, 3).tolist(), "elapsed", round(el, 2))
    assert (lat >= 0).all() and (qs >= 0).all() and (qs <= lat).all()


def test_engine_large_mixed_batch_matches_plain_and_oracle():
    # config-4 shaped: n log-uniform in [100, 5000], coords U{0..9999}^2, random starts
    rs = np.random.default_rng(8)
    sizes = np.rint(np.exp(rs.uniform(np.log(100), np.log(5000), 160))).astype(int)
    insts = []
    with L.Engine(hbm_budget=256 << 20, max_slots=256) as eng:
        for k, n in enumerate(sizes):
            xy = L.gen_uniform(int(n), 10**4, 300 + k)
            r0 = L.random_perm(int(n), 400 + k)
            insts.append((eng.add(L.coords_of(xy), r0), int(n), xy, r0))
        assert eng.run(sync_every=32) == 0
        results = {s: eng.result(s, n) for s, n, _, _ in insts}
    checked = 0
    for s, n, xy, r0 in insts[::8]:
        if n <= 700:
            r, Lo, *_ = O.local_search(xy, r0, threads=THREADS)
        else:
            with L.Context(L.coords_of(xy)) as ctx:
                out = ctx.optimize(r0, sync_every=1)
            r, Lo = out["route"], out["length"]
        assert results[s]["length"] == Lo and results[s]["route"].tolist() == list(r)
        checked += 1
    assert checked == 20


def test_engine_no_graph_mode_identical():
    n = 900
    xy = L.gen_uniform(n, 10**4, 77)
    r0 = L.random_perm(n, 78)
    outs = []
    for graphs in (True, False):
        with L.Engine(hbm_budget=32 << 20, max_slots=2) as eng:
            eng.set_graphs(graphs)
            s = eng.add(L.coords_of(xy), r0, log_cap=8192)
            eng.run(sync_every=32)
            outs.append(eng.result(s, n, log_cap=8192))
    assert outs[0]["length"] == outs[1]["length"] and outs[0]["route"].tolist() == outs[1]["route"].tolist()
    np.testing.assert_array_equal(outs[0]["log"], outs[1]["log"])
