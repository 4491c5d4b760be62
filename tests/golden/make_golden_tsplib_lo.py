"""Writes tests/golden/tsplib_lo.json: the oracle's full 2-opt local search from
the nearest-neighbour tour (city 0 first) on TSPLIB files under their own
metric -- att532 (ATT), pla7397 and pla33810 (CEIL_2D) -- as final length,
move and sweep counts, and SHA-256 digests of the move log (int32 triples
i, J, gain, little endian) and of the final tour (uint32 ids).

    python tests/golden/make_golden_tsplib_lo.py [threads]

Reads only the committed fixtures (tests/golden/tsplib/*.tsp.xz) and runs the
oracle (oracle/liblogo_oracle.so); the pla33810 search takes tens of minutes
on 8 threads, which is why its result is committed rather than recomputed by
the tests.
"""
import hashlib
import json
import lzma
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "logo-tsp-solver_amd"))

import logo2opt as L  # noqa: E402  (host-only calls: the .tsp reader)
import oracle_lib as O  # noqa: E402

FILES = ["att532", "pla7397", "pla33810"]


def load(name):
    with open(os.path.join(HERE, "tsplib", name + ".tsp.xz"), "rb") as f:
        data = lzma.decompress(f.read())
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, name + ".tsp")
        with open(p, "wb") as f:
            f.write(data)
        info, c = L.tsp_read(p)
    xy = np.ascontiguousarray(np.frombuffer(c, dtype=np.float32).astype(np.int32))
    return info, xy


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    out = {}
    for name in FILES:
        info, xy = load(name)
        r0 = O.nn_tour(xy, 0)
        t = time.time()
        r, length, moves, sweeps, log = O.local_search(xy, r0, threads=threads, metric=info["metric"],
                                                       log_cap=1 << 20)
        out[name] = dict(
            n=info["n"], metric=info["metric"], start="nn0",
            start_length=O.route_length(xy, r0, info["metric"]),
            length=length, moves=moves, sweeps=sweeps,
            log_sha256=hashlib.sha256(np.ascontiguousarray(log, "<i4").tobytes()).hexdigest(),
            tour_sha256=hashlib.sha256(np.ascontiguousarray(r, "<u4").tobytes()).hexdigest(),
            first_moves=log[:8].tolist())
        print(name, out[name]["length"], moves, f"{time.time() - t:.0f}s", flush=True)
    with open(os.path.join(HERE, "tsplib_lo.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
