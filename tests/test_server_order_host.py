"""The server's admission order on the CPU: csrc/l2o_server.cpp compiled with
g++ (AddressSanitizer and UndefinedBehaviorSanitizer, no GPU code) against a
host stand-in for the engine, in which a slot finishes at the first engine
step after its admission.  With one running request at a time, completions
then come in admission order, so the test reads the order itself: FIFO,
shortest-first, shortest-first with a bound on overtaking, the 1,024-request
look-ahead window, and the rejection of an unknown order.  The reference has
no serving layer; the same order is checked on the GPU, with real results, in
tests/test_gpu_serving.py."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# engine stand-in: every used slot is done after one step; EAGAIN when no slot is free
DRIVER = r"""
#include <cstdio>
#include <cstring>
#include <vector>
#include "logo2opt.h"

struct logo2opt_engine {
    std::vector<int> used, ns;
    std::vector<int32_t> done;
};

extern "C" {
int logo2opt_engine_create(int, uint64_t, uint32_t max_slots, logo2opt_engine **out)
{
    logo2opt_engine *e = new logo2opt_engine();
    e->used.assign(max_slots, 0);
    e->ns.assign(max_slots, 0);
    *out = e;
    return LOGO2OPT_OK;
}
int logo2opt_engine_destroy(logo2opt_engine *e) { delete e; return LOGO2OPT_OK; }
int logo2opt_engine_add(logo2opt_engine *e, uint32_t n, const logo2opt_city_coords *, const uint32_t *,
                        const logo2opt_job *, int32_t *slot)
{
    for (size_t k = 0; k < e->used.size(); ++k)
        if (!e->used[k]) {
            e->used[k] = 1;
            e->ns[k] = (int)n;
            *slot = (int32_t)k;
            return LOGO2OPT_OK;
        }
    return LOGO2OPT_EAGAIN;
}
int logo2opt_engine_run(logo2opt_engine *e, uint32_t, uint32_t, uint32_t *live)
{
    for (size_t k = 0; k < e->used.size(); ++k)
        if (e->used[k] == 1) {
            e->used[k] = 2;
            e->done.push_back((int32_t)k);
        }
    if (live) *live = 0;
    return LOGO2OPT_OK;
}
int logo2opt_engine_poll_done(logo2opt_engine *e, int32_t *out, uint32_t cap, uint32_t *count)
{
    uint32_t c = 0;
    while (c < cap && c < e->done.size()) { out[c] = e->done[c]; ++c; }
    e->done.erase(e->done.begin(), e->done.begin() + c);
    *count = c;
    return LOGO2OPT_OK;
}
int logo2opt_engine_result(logo2opt_engine *e, int32_t slot, uint32_t *route_out, logo2opt_result *res,
                           logo2opt_best2_out *)
{
    const int n = e->ns[slot];
    for (int k = 0; k < n; ++k) route_out[k] = (uint32_t)k;
    std::memset(res, 0, sizeof(*res));
    res->length = (uint64_t)n;
    return LOGO2OPT_OK;
}
int logo2opt_engine_remove(logo2opt_engine *e, int32_t slot) { e->used[slot] = 0; return LOGO2OPT_OK; }
int logo2opt_engine_set_pruning(logo2opt_engine *, int) { return LOGO2OPT_OK; }
int logo2opt_nn_tour(uint32_t n, const int32_t *, uint32_t, uint32_t *route)
{
    for (uint32_t k = 0; k < n; ++k) route[k] = k;
    return LOGO2OPT_OK;
}
}

// submit `sizes` in order, serve one request at a time; print the ids
// (1-based submission order) in completion order
static void run(const char *name, const std::vector<uint32_t> &sizes, int order, uint32_t bypass, size_t show)
{
    logo2opt_server *s = nullptr;
    if (logo2opt_server_create(0, 1u << 30, 8, 0, &s)) { std::printf("%s create-failed\n", name); return; }
    logo2opt_server_set_limits(s, 1, 0);
    if (order >= 0 && logo2opt_server_set_order(s, order, bypass)) { std::printf("%s order-failed\n", name); return; }
    std::vector<logo2opt_city_coords> xy;
    std::vector<uint32_t> route;
    logo2opt_job job;
    std::memset(&job, 0, sizeof(job));
    for (uint32_t n : sizes) {
        xy.assign(n, logo2opt_city_coords{0.f, 0.f});
        for (uint32_t k = 0; k < n; ++k) xy[k].x = (float)k;
        route.resize(n);
        for (uint32_t k = 0; k < n; ++k) route[k] = k;
        int64_t id = -1;
        if (logo2opt_server_submit(s, n, xy.data(), route.data(), &job, &id)) { std::printf("%s submit-failed\n", name); return; }
    }
    std::vector<int64_t> seen;
    for (;;) {
        uint32_t running = 0, queued = 0;
        if (logo2opt_server_pump(s, 1, 1, &running, &queued)) { std::printf("%s pump-failed\n", name); return; }
        if (running > 1) { std::printf("%s over-limit\n", name); return; }
        int64_t ids[64];
        uint32_t c = 0;
        logo2opt_server_poll(s, ids, 64, &c);
        for (uint32_t k = 0; k < c; ++k) seen.push_back(ids[k]);
        if (running == 0 && queued == 0) break;
    }
    std::printf("%s %zu", name, seen.size());
    for (size_t k = 0; k < seen.size() && k < show; ++k) std::printf(" %lld", (long long)seen[k]);
    std::printf("\n");
    logo2opt_server_destroy(s);
}

int main()
{
    const std::vector<uint32_t> five = {2400, 300, 1200, 150, 600};
    run("fifo", five, -1, 0, 5);
    run("fifo_explicit", five, LOGO2OPT_ORDER_FIFO, 0, 5);
    run("shortest", five, LOGO2OPT_ORDER_SHORTEST, 0, 5);
    run("bypass2", five, LOGO2OPT_ORDER_SHORTEST, 2, 5);
    run("bypass1", five, LOGO2OPT_ORDER_SHORTEST, 1, 5);
    // ties go to the earlier arrival
    run("ties", {500, 200, 200, 100, 100}, LOGO2OPT_ORDER_SHORTEST, 0, 5);
    // 1,025 requests of 1,000 cities, then one of 10: it enters the
    // 1,024-request window after two admissions
    std::vector<uint32_t> many(1025, 1000);
    many.push_back(10);
    run("window", many, LOGO2OPT_ORDER_SHORTEST, 0, 4);
    logo2opt_server *s = nullptr;
    logo2opt_server_create(0, 1u << 30, 8, 0, &s);
    std::printf("invalid %d %d\n", logo2opt_server_set_order(s, 7, 0), logo2opt_server_set_order(nullptr, 0, 0));
    logo2opt_server_destroy(s);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_server_admission_order_host(tmp_path):
    drv = tmp_path / "drv.cpp"
    drv.write_text(DRIVER)
    exe = tmp_path / "drv"
    subprocess.run(["g++", "-std=c++17", "-g", "-O1", "-pthread", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "logo-tsp-solver_amd", "csrc", "l2o_server.cpp"), str(drv), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1"))
    assert out.returncode == 0, out.stderr[-4000:]
    got = {l.split()[0]: [int(v) for v in l.split()[1:]] for l in out.stdout.strip().split("\n")}
    # ids are 1-based submission order; sizes 2400, 300, 1200, 150, 600
    assert got["fifo"] == [5, 1, 2, 3, 4, 5]
    assert got["fifo_explicit"] == [5, 1, 2, 3, 4, 5]
    assert got["shortest"] == [5, 4, 2, 5, 3, 1]
    # bound 2: 2,400 is overtaken by 150 and 300, then goes next; 1,200 is
    # overtaken by 150 and 600 and goes last
    assert got["bypass2"] == [5, 4, 2, 1, 5, 3]
    # bound 1: 150 overtakes everything before it once; 2,400 goes next, then
    # 300 (overtaken once), then 1,200 (overtaken once), then 600
    assert got["bypass1"] == [5, 4, 1, 2, 3, 5]
    assert got["ties"] == [5, 4, 5, 2, 3, 1]
    assert got["window"] == [1026, 1, 2, 1026, 3]
    assert got["invalid"] == [-1, -1]  # LOGO2OPT_EINVAL
