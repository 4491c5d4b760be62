"""The host construction code (logo-tsp-solver_amd/csrc/l2o_workload.cpp:
generators, grid nearest-neighbour and multiple-fragment tours) built with
AddressSanitizer and UndefinedBehaviorSanitizer (g++, no GPU code) and run
on uniform, clustered and degenerate instances (all points equal,
collinear, a square grid full of ties, duplicated points, n = 4, the full
2^22 coordinate span): no sanitizer report, every tour a permutation, and
every tour identical to the one the shipped library builds."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import logo2opt as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r'''
#include "logo2opt.h"
#include <cstdio>
#include <vector>
static unsigned long long fnv(const std::vector<uint32_t> &r)
{
    unsigned long long h = 1469598103934665603ull;
    for (uint32_t v : r)
        for (int b = 0; b < 4; ++b) {
            h ^= (v >> (8 * b)) & 0xFF;
            h *= 1099511628211ull;
        }
    return h;
}
static bool perm(const std::vector<uint32_t> &r)
{
    std::vector<char> s(r.size(), 0);
    for (uint32_t v : r) {
        if (v >= r.size() || s[v]) return false;
        s[v] = 1;
    }
    return true;
}
static void run(const char *name, std::vector<int32_t> xy, uint32_t start)
{
    const uint32_t n = (uint32_t)(xy.size() / 2);
    std::vector<uint32_t> a(n), b(n);
    const int ra = logo2opt_nn_tour(n, xy.data(), start, a.data());
    const int rb = logo2opt_mf_tour(n, xy.data(), b.data());
    std::printf("%s %d %d %llu %d %d %llu\n", name, ra, (int)perm(a), fnv(a), rb, (int)perm(b), fnv(b));
}
int main()
{
    std::vector<int32_t> xy;
    xy.assign(2 * 3000, 0); logo2opt_gen_uniform(3000, 10000, 11, xy.data()); run("uniform", xy, 0);
    xy.assign(2 * 5000, 0); logo2opt_gen_clustered(5000, 1000000, 60, 0.02, 0.10, 12, xy.data());
    run("clustered", xy, 4999);
    xy.assign(2 * 500, 0); logo2opt_gen_uniform(500, (1u << 22) - 1, 13, xy.data()); run("wide", xy, 0);
    xy.assign(2 * 50, 7); run("equal", xy, 0);
    xy.clear(); for (int k = 0; k < 100; ++k) { xy.push_back(3 * k); xy.push_back(0); } run("collinear", xy, 0);
    xy.clear(); for (int i = 0; i < 30; ++i) for (int j = 0; j < 30; ++j) { xy.push_back(10 * i); xy.push_back(10 * j); }
    run("grid", xy, 0);
    xy.clear(); for (int k = 0; k < 200; ++k) { xy.push_back(k / 2 * 5); xy.push_back(k / 2 * 3 % 97); } run("dups", xy, 0);
    xy = {0, 0, 1, 0, 1, 1, 0, 1}; run("tiny4", xy, 0);
    return 0;
}
'''


def fnv(r):
    h = 1469598103934665603
    for byte in np.asarray(r, "<u4").tobytes():
        h = ((h ^ byte) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def instances():
    yield "uniform", L.gen_uniform(3000, 10000, 11), 0
    yield "clustered", L.gen_clustered(5000, 1000000, 12), 4999
    yield "wide", L.gen_uniform(500, (1 << 22) - 1, 13), 0
    yield "equal", np.full(100, 7, np.int32), 0
    yield "collinear", np.array([[3 * k, 0] for k in range(100)], np.int32).ravel(), 0
    yield "grid", np.array([[10 * i, 10 * j] for i in range(30) for j in range(30)], np.int32).ravel(), 0
    yield "dups", np.array([[k // 2 * 5, k // 2 * 3 % 97] for k in range(200)], np.int32).ravel(), 0
    yield "tiny4", np.array([0, 0, 1, 0, 1, 1, 0, 1], np.int32), 0


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_construction_under_sanitizers(tmp_path):
    drv = tmp_path / "drv.cpp"
    drv.write_text(DRIVER)
    exe = tmp_path / "drv"
    subprocess.run(["g++", "-std=c++17", "-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "logo-tsp-solver_amd", "csrc", "l2o_workload.cpp"), str(drv), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300,
                         env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1"))
    assert out.returncode == 0, out.stderr[-4000:]
    got = {l.split()[0]: [int(v) for v in l.split()[1:]] for l in out.stdout.strip().split("\n")}
    for name, xy, start in instances():
        ra, pa, ha, rb, pb, hb = got[name]
        assert (ra, pa, rb, pb) == (0, 1, 0, 1), (name, got[name])
        assert ha == fnv(L.nn_tour(xy, start)), name
        assert hb == fnv(L.mf_tour(xy)), name
