"""One batched engine holding a seeded random mix of jobs, against the
oracle, bit for bit: plain local searches, ILS with either kick and
multi-start ILS groups, under every metric, sizes log-uniform in [5, 1200],
uniform or clustered points, random or nearest-neighbour starts, with row
pruning on or off for the whole engine.  Instances of different sizes,
metrics and job kinds share every launch, so the per-instance state, the
work lists, the ILS bookkeeping and the group barrier are all exercised
together.  LOGO2OPT_RANDOM_ENGINE_JOBS sets the number of jobs (default 24),
LOGO2OPT_RANDOM_ENGINE_SEED adds to the seeds that draw them (default 0)."""
import os
import sys

import numpy as np
import pytest

import logo2opt as L
import oracle_lib as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
JOBS = int(os.environ.get("LOGO2OPT_RANDOM_ENGINE_JOBS", "24"))
SEED = int(os.environ.get("LOGO2OPT_RANDOM_ENGINE_SEED", "0"))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def jobs(count, seed):
    rs = np.random.default_rng(seed)
    out = []
    for q in range(count):
        n = int(np.rint(np.exp(rs.uniform(np.log(5), np.log(1200)))))
        side = int(rs.choice([100, 10**4, 10**6]))
        metric = int(rs.integers(0, 3))
        base = 100003 * SEED  # 0 by default: the default jobs' point and start-tour seeds
        xy = O.gen_clustered(n, side, base + 300 + q) if (side >= 10**4 and rs.random() < 0.5) else \
            O.gen_uniform(n, side, base + 300 + q)
        r0 = O.nn_tour(xy, 0) if rs.random() < 0.5 else O.random_perm(n, base + 500 + q)
        kind = str(rs.choice(["lo", "ils", "group"]))
        if n < 8:  # ILS needs n >= 8 (three double-bridge cuts); the engine rejects less
            kind = "lo"
        rounds = int(rs.integers(1, 6))
        chains = int(rs.integers(2, 5))
        kick = str(rs.choice(["double_bridge", "reference"]))
        out.append(dict(q=q, n=n, xy=xy, r0=r0, metric=metric, kind=kind, rounds=rounds, chains=chains, kick=kick,
                        seed=40 + q))
    return out


@pytest.mark.parametrize("prune", [False, True], ids=["full", "pruned"])
def test_random_engine_mix_matches_oracle(prune):
    js = jobs(JOBS, (77 if prune else 76) + 1000 * SEED)
    say(f"engine mix: {len(js)} jobs, prune={prune}, kinds", [j["kind"] for j in js])
    # a multi-start group takes one slot per chain
    slots = sum(j["chains"] if j["kind"] == "group" else 1 for j in js)
    with L.Engine(hbm_budget=1 << 29 if slots <= 256 else 1 << 31, max_slots=max(256, slots)) as e:
        e.set_pruning(prune)
        for j in js:
            c = L.coords_of(j["xy"])
            j["coords"] = c  # kept alive while the engine holds the instance
            k = L.KICK_DOUBLE_BRIDGE if j["kick"] == "double_bridge" else L.KICK_REFERENCE
            if j["kind"] == "lo":
                j["slot"] = e.add(c, j["r0"], metric=j["metric"])
            elif j["kind"] == "ils":
                j["slot"] = e.add(c, j["r0"], ils_rounds=j["rounds"], kick=k, seed=j["seed"], metric=j["metric"])
            else:
                j["slot"] = e.add_group(c, j["r0"], chains=j["chains"], ils_rounds=j["rounds"], kick=k,
                                        seed=j["seed"], metric=j["metric"])
            assert j["slot"] is not None
        assert e.run(max_steps=2_000_000, sync_every=16) == 0
        for j in js:
            xy, r0, m, n = j["xy"], j["r0"], j["metric"], j["n"]
            if j["kind"] == "lo":
                rr, Lo, *_ = O.local_search(xy, r0, threads=THREADS, metric=m)
                res = e.result(j["slot"], n)
            elif j["kind"] == "ils":
                rr, Lo, _, _ = O.ils(xy, r0, j["seed"], j["rounds"], kick=j["kick"], threads=THREADS, metric=m)
                res = e.result(j["slot"], n)
            else:
                rr, Lo, _, _ = O.ils_multistart(xy, r0, j["seed"], j["rounds"], j["chains"], kick=j["kick"],
                                                threads=THREADS, metric=m)
                res = e.group_result(j["slot"], n)
            assert res["length"] == Lo and res["route"].tolist() == rr.tolist(), \
                {k: j[k] for k in ("q", "n", "metric", "kind", "rounds", "chains", "kick")}
