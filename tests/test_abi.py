"""CPU-side checks of the product library: it loads, exports every entry point
include/logo2opt.h declares, its host-side workload code matches the oracle
bit for bit, and argument validation fails with the documented error codes
before any device work (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import logo2opt as L
import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "logo2opt.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(logo2opt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (logo2opt_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = L.load()
    for n in names:
        assert hasattr(lib, n)
    assert sorted(L.EXPORTS) == names


def test_struct_layouts_match_reference():
    # city_coords 8 B, best2_out 12 B (defs.h:173-187); process_time = 3 doubles + best2_out
    assert C.sizeof(L.CityCoords) == 8
    assert C.sizeof(L.Best2Out) == 12
    assert C.sizeof(L.ProcessTime) == 3 * 8 + 12 + 4  # padded to 8
    assert C.sizeof(L.Job) == 32 and C.sizeof(L.Result) == 40


def test_strerror_codes():
    for code in range(0, -9, -1):
        assert L.strerror(code)
    assert L.strerror(-100) == "unknown error"


@pytest.mark.parametrize("n,side,seed", [(1000, 10**4, 1), (33810, 10**6, 2), (5000, 10**4, 3)])
def test_uniform_matches_oracle(n, side, seed):
    assert (L.gen_uniform(n, side, seed) == O.gen_uniform(n, side, seed)).all()


def test_clustered_matches_oracle():
    assert (L.gen_clustered(33810, 10**6, 7) == O.gen_clustered(33810, 10**6, 7)).all()
    assert (L.gen_clustered(5000, 10**4, 8, 10, 0.05, 0.2) == O.gen_clustered(5000, 10**4, 8, 10, 0.05, 0.2)).all()


def test_random_perm_matches_oracle():
    for n, s in [(4, 1), (1000, 2), (100000, 3)]:
        assert (L.random_perm(n, s) == O.random_perm(n, s)).all()


@pytest.mark.parametrize("kind,n,side", [("uniform", 3000, 10**4), ("uniform", 2000, 60),
                                         ("clustered", 8000, 10**6), ("uniform", 500, 3)])
def test_nn_tour_matches_oracle(kind, n, side):
    xy = L.gen_uniform(n, side, 11) if kind == "uniform" else L.gen_clustered(n, side, 12)
    for start in (0, n // 2):
        assert (L.nn_tour(xy, start) == O.nn_tour(xy, start)).all()


def test_invalid_inputs_rejected_before_device_work():
    lib = L.load()
    h = C.c_void_p()
    # non-integral coordinates
    c = L.float_coords(np.array([0, 1.5, 2, 3], np.float32), np.zeros(4, np.float32))
    assert lib.logo2opt_create(0, 4, C.cast(c, C.c_void_p), C.byref(h)) == -2
    # span >= 2^22
    c = L.float_coords(np.array([0, 1 << 22, 2, 3], np.float32), np.zeros(4, np.float32))
    assert lib.logo2opt_create(0, 4, C.cast(c, C.c_void_p), C.byref(h)) == -2
    # n out of range
    c = L.float_coords(np.array([0, 1], np.float32), np.zeros(2, np.float32))
    assert lib.logo2opt_create(0, 2, C.cast(c, C.c_void_p), C.byref(h)) == -1
    assert lib.logo2opt_create(0, 4, None, C.byref(h)) == -1


def test_engine_bytes_for():
    assert L.Engine.bytes_for(1000) == 2 * 16128 + 2 * 4096
    assert L.Engine.bytes_for(1000, ils_rounds=10) > L.Engine.bytes_for(1000)


def test_header_is_c_and_cpp_adapter_compiles(tmp_path):
    pkg = os.path.dirname(L.LIB_PATH)
    inc = os.path.join(ROOT, "include")
    c_src = tmp_path / "t.c"
    c_src.write_text('#include "logo2opt.h"\n#include <string.h>\n'
                     'int main(void) { return strcmp(logo2opt_strerror(LOGO2OPT_ETOUR),\n'
                     '                               "route is not a permutation of 0..n-1") != 0; }\n')
    exe = tmp_path / "tc"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, str(c_src), "-o", str(exe), "-L", pkg,
                    "-llogo2opt", f"-Wl,-rpath,{pkg}"], check=True)
    assert subprocess.run([str(exe)]).returncode == 0
    cpp_src = tmp_path / "t.cpp"
    cpp_src.write_text('#include "hip_local_solver.h"\n#include <cstdio>\n'
                       'int main() { try { logo2opt_city_coords c[2] = {{0, 0}, {1, 0}};\n'
                       '  logo2opt::HIPSolver s(c, 2, 0); } catch (const std::runtime_error &e) {\n'
                       '  std::puts(e.what()); return 0; } return 1; }\n')
    exe2 = tmp_path / "tcpp"
    subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-I", inc, "-I", os.path.join(pkg, "csrc"),
                    str(cpp_src), "-o", str(exe2), "-L", pkg, "-llogo2opt", f"-Wl,-rpath,{pkg}"], check=True)
    r = subprocess.run([str(exe2)], capture_output=True, text=True)
    # n = 2 is rejected before any device work: the adapter throws, never exits
    assert r.returncode == 0 and "invalid argument" in r.stdout


def test_cpp_adapter_file_helpers(tmp_path):
    # readFileCoords / saveRouteToAFile / routeInitFromFile under the
    # reference's names (host only, no device work)
    pkg = os.path.dirname(L.LIB_PATH)
    inc = os.path.join(ROOT, "include")
    tsp = tmp_path / "sq.tsp"
    tsp.write_text("NAME : sq\nTYPE : TSP\nDIMENSION : 4\nEDGE_WEIGHT_TYPE : ATT\nNODE_COORD_SECTION\n"
                   "1 10 10\n2 40 10\n3 40 50\n4 10 50\nEOF\n")
    src = tmp_path / "f.cpp"
    src.write_text(
        '#include "hip_local_solver.h"\n#include <cstdio>\n'
        'int main(int, char **argv) {\n'
        '  int metric = -1;\n'
        '  std::vector<logo2opt_city_coords> c = logo2opt::readFileCoords(argv[1], &metric);\n'
        '  std::vector<uint32_t> r = {2, 0, 3, 1};\n'
        '  uint64_t len = 0;\n'
        '  if (logo2opt_host_route_length(4, c.data(), metric, r.data(), &len)) return 2;\n'
        '  logo2opt::saveRouteToAFile(r, argv[2], (unsigned long)len);\n'
        '  std::vector<uint32_t> back;\n'
        '  logo2opt::routeInitFromFile(back, 4, (std::string(argv[2]) + "_" + std::to_string(len) + ".tour").c_str());\n'
        '  std::printf("%zu %d %g %g %llu %d\\n", c.size(), metric, c[2].x, c[2].y, (unsigned long long)len,\n'
        '              (int)(back == r));\n'
        '  try { logo2opt::readFileCoords(argv[3]); } catch (const std::runtime_error &e) { std::puts(e.what()); }\n'
        '  return 0; }\n')
    exe = tmp_path / "f"
    subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-I", inc, "-I", os.path.join(pkg, "csrc"),
                    str(src), "-o", str(exe), "-L", pkg, "-llogo2opt", f"-Wl,-rpath,{pkg}"], check=True)
    out = subprocess.run([str(exe), str(tsp), str(tmp_path / "best"), str(tmp_path / "none.tsp")],
                         capture_output=True, text=True, check=True).stdout.split("\n")
    # tour 3-1-4-2 has Euclidean edges 50, 40, 50, 40; ATT is the smallest m
    # with 10 m^2 >= d^2: 16, 13, 16, 13
    att = sum(next(m for m in range(100) if 10 * m * m >= d * d) for d in (50, 40, 50, 40))
    assert att == 58
    assert out[0].split() == ["4", "2", "30", "40", str(att), "1"]
    assert (tmp_path / f"best_{att}.tour").read_text().split() == ["3", "1", "4", "2"]
    assert "could not be opened" in out[1]
