"""The oracle (oracle/logo_oracle.c, the parity checker) built with
AddressSanitizer and UndefinedBehaviorSanitizer (gcc) and run over ragged
sizes: the exact SIMD sweep (orc_step_fast, in its AVX-512, AVX2 and scalar
variants) must return the plain step's move for every size 4..70 and around
the 16-column block edges, under every metric and thread count, and the
local search, multi-start ILS and multiple-fragment construction must run
clean on small and degenerate instances."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r'''
#include "logo_oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
int main(void)
{
    static const uint32_t extra[] = {127, 128, 129, 255, 256, 257, 511, 1000, 1001};
    uint32_t sizes[80], ns = 0;
    for (uint32_t n = 4; n <= 70; ++n) sizes[ns++] = n;
    for (uint32_t k = 0; k < sizeof(extra) / sizeof(extra[0]); ++k) sizes[ns++] = extra[k];
    int mismatches = 0, checked = 0;
    for (uint32_t q = 0; q < ns; ++q) {
        const uint32_t n = sizes[q];
        int32_t *xy = malloc(sizeof(int32_t) * 2 * n);
        uint32_t *r = malloc(sizeof(uint32_t) * n);
        for (int v = 0; v < 2; ++v) {
            orc_gen_uniform(n, q % 3 == 0 ? 20 : (q % 3 == 1 ? 10000 : 4194303), 1000 + q, xy);
            if (v == 0) orc_random_perm(n, 2000 + q, r);
            else orc_nn_tour(n, xy, 0, r);
            for (int metric = 0; metric < 3; ++metric)
                for (int threads = 1; threads <= 3; threads += 2) {
                    orc_best2_out a, b;
                    const int oa = orc_step_m(n, xy, r, threads, metric, &a);
                    const int ob = orc_step_fast(n, xy, r, threads, metric, &b, NULL);
                    ++checked;
                    if (oa != ob || (oa && (a.i != b.i || a.j != b.j || a.minchange != b.minchange))) ++mismatches;
                }
        }
        if (n <= 300) {
            uint64_t sweeps = 0, len = 0;
            orc_best2_out log[4096];
            orc_local_search_m(n, xy, r, 2, (int)(n % 3), 0, log, 4096, &sweeps, &len);
            if (n >= 8) {
                uint64_t rl[4];
                orc_ils_multistart(n, xy, r, 5, 3, 3, (int)(n & 1), 2, (int)(n % 3), rl, NULL);
            }
        }
        orc_mf_tour(n, xy, r);
        free(xy);
        free(r);
    }
    {   /* degenerate: all points equal, collinear */
        int32_t xy[2 * 40];
        uint32_t r[40];
        for (int k = 0; k < 80; ++k) xy[k] = 5;
        orc_mf_tour(40, xy, r);
        orc_nn_tour(40, xy, 0, r);
        for (int k = 0; k < 40; ++k) { xy[2 * k] = 3 * k; xy[2 * k + 1] = 0; }
        orc_mf_tour(40, xy, r);
        orc_best2_out a, b;
        orc_random_perm(40, 9, r);
        const int oa = orc_step_m(40, xy, r, 1, 0, &a), ob = orc_step_fast(40, xy, r, 1, 0, &b, NULL);
        ++checked;
        if (oa != ob || (oa && (a.i != b.i || a.j != b.j || a.minchange != b.minchange))) ++mismatches;
    }
    printf("%s %d %d\n", orc_fast_isa(), checked, mismatches);
    return 0;
}
'''


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_under_sanitizers(tmp_path):
    drv = tmp_path / "drv.c"
    drv.write_text(DRIVER)
    exe = tmp_path / "drv"
    subprocess.run(["gcc", "-std=c11", "-g", "-O1", "-ffp-contract=off", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-I", os.path.join(ROOT, "oracle"),
                    os.path.join(ROOT, "oracle", "logo_oracle.c"), str(drv), "-o", str(exe), "-lpthread", "-lm"],
                   check=True)
    seen = set()
    for isa in ("avx512", "avx2", "scalar"):
        env = dict(os.environ, ORC_FAST_ISA=isa, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1")
        out = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
        assert out.returncode == 0, out.stderr[-4000:]
        name, checked, mismatches = out.stdout.split()
        assert int(checked) > 900 and int(mismatches) == 0, (isa, out.stdout)
        seen.add(name)
        flags = open("/proc/cpuinfo").read().split() if os.path.exists("/proc/cpuinfo") else []
        if isa == "scalar" or {"avx512": "avx512f", "avx2": "avx2"}[isa] in flags:
            assert name == isa, (isa, name)  # the variant asked for is the one that ran
    assert "scalar" in seen
