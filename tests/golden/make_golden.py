"""Writes the committed golden fixtures under tests/golden/.

Run once in a container that has the reference checkout (read as data only):

    python tests/golden/make_golden.py [/root/reference] [all|npz|tsplib]

Outputs
  tsplib_int.npz   integral-coordinate EUC_2D TSPLIB instances (TSPLIB/*.tsp)
                   with the known optimal lengths the reference keys by file
                   name (src/utils/cut.c:205-310), used as lower bounds.
  monalisa.npz     TSPLIB/mona-lisa100K.tsp coordinates and the 0-based tour of
                   TSPLIB/monalisa_5757191.tour; its length 5,757,191 is stated
                   by the tour file's NAME header (line 1) and file name.
  nint_golden.npz  random s < 2^46 with nint(sqrt(s)) = (isqrt(4s)+1)//2 from
                   Python's math.isqrt (stdlib; a different formula from the
                   oracle's).
  tsplib/*.tsp.xz  TSPLIB data files, byte for byte (xz-compressed), for the
                   .tsp reader tests: berlin52 (EUC_2D), att532 (ATT),
                   pla7397 and pla33810 (CEIL_2D), pcb3038 (negative values,
                   exponent notation), vm1084 (non-integral values) and
                   rat195 (an HTML page in place of the data).
  monalisa_tour.json  the header lines, line ending, size and SHA-256 of
                   TSPLIB/monalisa_5757191.tour, so the tests can rebuild that
                   file byte for byte from monalisa.npz's tour.

Nothing here imports or runs reference code; the .tsp/.tour files are parsed
as text.
"""
import hashlib
import json
import lzma
import math
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# known optima, src/utils/cut.c:205-310 (EUC_2D files with integral coordinates)
KNOWN_OPT = {
    "berlin52": 7542, "kroA100": 21282, "kroB100": 22141, "kroC100": 20749,
    "kroD100": 21294, "kroE100": 22068, "kroA150": 26524, "kroA200": 29368,
    "a280": 2579, "pr226": 80369, "pr264": 49135, "pr299": 48191,
    "pr439": 107217, "ts225": 126643, "rat783": 8806, "pr2392": 378032,
    "pcb3038": 137694,
}


def read_tsp(path):
    with open(path) as f:
        lines = f.read().split("\n")
    start = [k for k, l in enumerate(lines) if l.strip().startswith("NODE_COORD_SECTION")][0]
    head = "\n".join(lines[:start])
    assert "EUC_2D" in head, path
    xs, ys = [], []
    for l in lines[start + 1:]:
        p = l.split()
        if len(p) < 3 or p[0] == "EOF":
            continue
        x, y = float(p[1]), float(p[2])
        assert x == int(x) and y == int(y), (path, l)
        xs.append(int(x))
        ys.append(int(y))
    return np.array(xs, np.int32), np.array(ys, np.int32)


def read_tour(path):
    with open(path) as f:
        lines = f.read().split("\n")
    name = lines[0]
    start = [k for k, l in enumerate(lines) if l.strip().startswith("TOUR_SECTION")][0]
    ids = []
    for l in lines[start + 1:]:
        l = l.strip()
        if not l or l in ("-1", "EOF"):
            continue
        ids.append(int(l) - 1)
    return name, np.array(ids, np.uint32)


TSPLIB_FILES = ["berlin52", "att532", "pla7397", "pla33810", "pcb3038", "vm1084", "rat195"]


def write_tsplib(tsplib):
    out = os.path.join(HERE, "tsplib")
    os.makedirs(out, exist_ok=True)
    for name in TSPLIB_FILES:
        with open(os.path.join(tsplib, name + ".tsp"), "rb") as f:
            data = f.read()
        with open(os.path.join(out, name + ".tsp.xz"), "wb") as f:
            f.write(lzma.compress(data, preset=9))
    with open(os.path.join(tsplib, "monalisa_5757191.tour"), "rb") as f:
        raw = f.read()
    eol = "\r\n" if b"\r\n" in raw else "\n"
    text = raw.decode()
    head = text[:text.index("TOUR_SECTION") + len("TOUR_SECTION") + len(eol)]
    meta = dict(header=head, eol=eol, bytes=len(raw), sha256=hashlib.sha256(raw).hexdigest())
    with open(os.path.join(HERE, "monalisa_tour.json"), "w") as f:
        json.dump(meta, f, indent=1)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    what = sys.argv[2] if len(sys.argv) > 2 else "all"
    tsplib = os.path.join(ref, "TSPLIB")
    if what in ("all", "tsplib"):
        write_tsplib(tsplib)
    if what not in ("all", "npz"):
        print("wrote fixtures to", HERE)
        return

    arrays = {}
    names = []
    for name, opt in KNOWN_OPT.items():
        x, y = read_tsp(os.path.join(tsplib, name + ".tsp"))
        arrays[name + "_x"] = x
        arrays[name + "_y"] = y
        arrays[name + "_opt"] = np.array([opt], np.int64)
        names.append(name)
    arrays["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "tsplib_int.npz"), **arrays)

    x, y = read_tsp(os.path.join(tsplib, "mona-lisa100K.tsp"))
    name, tour = read_tour(os.path.join(tsplib, "monalisa_5757191.tour"))
    stated = int(re.search(r"_(\d+)", name).group(1))
    assert stated == 5757191, name
    assert sorted(tour.tolist()) == list(range(len(x)))
    np.savez_compressed(os.path.join(HERE, "monalisa.npz"),
                        x=x.astype(np.uint16), y=y.astype(np.uint16), tour=tour,
                        stated_length=np.array([stated], np.int64),
                        known_opt=np.array([5757080], np.int64))

    rng = np.random.default_rng(20260101)
    s = np.concatenate([
        rng.integers(0, 1 << 24, 2000, dtype=np.int64),
        rng.integers(0, 1 << 46, 3000, dtype=np.int64),
    ])
    # values straddling the nint breakpoints k^2+k and k^2+k+1
    k = rng.integers(1, 1 << 23, 1500, dtype=np.int64)
    s = np.concatenate([s, k * k + k, k * k + k + 1, k * k, k * k - 1])
    # floor(sqrt(s) + 1/2) = floor((floor(2 sqrt(s)) + 1) / 2) = (isqrt(4s) + 1) // 2,
    # a different route from the oracle's k + [s > k^2 + k]
    nint = np.array([(math.isqrt(4 * int(v)) + 1) // 2 for v in s], np.int64)
    np.savez_compressed(os.path.join(HERE, "nint_golden.npz"), s=s, nint=nint)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
