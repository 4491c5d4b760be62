"""The TSPLIB .tsp / .tour reader (logo-tsp-solver_amd/csrc/l2o_tsplib.cpp,
host code) built with AddressSanitizer and UndefinedBehaviorSanitizer (g++,
no GPU code) and run on the TSPLIB fixtures and on malformed files: every
fixture must parse, every malformed file must come back as an error code,
and no run may trip a sanitizer (a report aborts the program)."""
import lzma
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "tsplib")
ROCM_INC = "/opt/rocm/include"

DRIVER = r'''
#include "logo2opt.h"
#include <cstdio>
#include <vector>
int main(int argc, char **argv)
{
    for (int a = 1; a < argc; ++a) {
        logo2opt_tsp_info info;
        const int rc = logo2opt_tsp_read(argv[a], 0, &info, nullptr, 0);
        const int rr = logo2opt_tsp_read(argv[a], LOGO2OPT_TSP_ROUND, &info, nullptr, 0);
        int rc2 = 1;
        if (rr == 0 && info.n > 0 && info.n <= (1u << 20)) {
            std::vector<logo2opt_city_coords> c(info.n);
            rc2 = logo2opt_tsp_read(argv[a], LOGO2OPT_TSP_ROUND, &info, c.data(), info.n);
        }
        std::vector<uint32_t> r(64);
        const int rt = logo2opt_tour_read(argv[a], 64, r.data());
        std::printf("%d %d %d %u %d\n", rc, rr, rc2, info.n, rt);
    }
    return 0;
}
'''

H = "NAME: t\nTYPE: TSP\nDIMENSION: {d}\nEDGE_WEIGHT_TYPE: {w}\nNODE_COORD_SECTION\n"
BAD_TSP = {
    "empty": "",
    "short_nodes": H.format(d=5, w="EUC_2D") + "1 0 0\n2 1 1\nEOF\n",
    "dim0": H.format(d=0, w="EUC_2D") + "EOF\n",
    "dim_overflow": H.format(d="99999999999999999999", w="EUC_2D") + "1 0 0\nEOF\n",
    "dim_neg": H.format(d=-5, w="EUC_2D") + "1 0 0\nEOF\n",
    "dim_text": H.format(d="abc", w="EUC_2D") + "1 0 0\nEOF\n",
    "dim_max": H.format(d=(1 << 20) + 1, w="EUC_2D") + "1 0 0\nEOF\n",
    "coord_text": H.format(d=3, w="EUC_2D") + "1 a b\n2 1 1\n3 2 2\nEOF\n",
    "missing_coord": H.format(d=3, w="EUC_2D") + "1 0\n2 1 1\n3 2 2\nEOF\n",
    "extra_fields": H.format(d=3, w="EUC_2D") + "1 0 0 9\n2 1 1\n3 2 2\nEOF\n",
    "id_range": H.format(d=3, w="EUC_2D") + "1 0 0\n2 1 1\n7 2 2\nEOF\n",
    "id_dup": H.format(d=3, w="EUC_2D") + "1 0 0\n1 1 1\n3 2 2\nEOF\n",
    "geo": H.format(d=3, w="GEO") + "1 0 0\n2 1 1\n3 2 2\nEOF\n",
    "no_section": "NAME: t\nTYPE: TSP\nDIMENSION: 3\nEDGE_WEIGHT_TYPE: EUC_2D\n1 0 0\n2 1 1\n3 2 2\nEOF\n",
    "colon_only": "DIMENSION:\nEDGE_WEIGHT_TYPE:\nNODE_COORD_SECTION\nEOF\n",
    "nan": H.format(d=3, w="EUC_2D") + "1 nan 0\n2 1 1\n3 2 2\nEOF\n",
    "inf": H.format(d=3, w="EUC_2D") + "1 inf 0\n2 1 1\n3 2 2\nEOF\n",
    "huge": H.format(d=3, w="EUC_2D") + "1 1e30 0\n2 1 1\n3 2 2\nEOF\n",
    "huge_neg": H.format(d=3, w="EUC_2D") + "1 -1e300 0\n2 1 1\n3 2 2\nEOF\n",
    "long_line": "x" * (1 << 20) + "\n",
}
TOURS = {  # n = 64: expected logo2opt_tour_read code
    "tour_ok": ("".join(f"{k}\n" for k in range(1, 65)), 0),
    "tour_trailing": ("".join(f"{k}\n" for k in range(1, 65)) + "-1\nEOF\n", 0),
    "tour_section": ("TYPE: TOUR\nDIMENSION: 64\nTOUR_SECTION\n" + "".join(f"{k}\n" for k in range(64, 0, -1))
                     + "-1\nEOF\n", 0),
    "tour_dup": ("".join(f"{k}\n" for k in range(1, 64)) + "1\n", -3),
    "tour_range": ("".join(f"{k}\n" for k in range(1, 64)) + "99\n", -3),
    "tour_zero": ("".join(f"{k}\n" for k in range(0, 64)), -3),
    "tour_short": ("1\n2\n3\n-1\n", -9),
    "tour_section_long": ("TYPE: TOUR\nTOUR_SECTION\n" + "".join(f"{k}\n" for k in range(1, 66)) + "-1\n", -9),
}


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(ROCM_INC), reason="needs g++ and HIP headers")
def test_tsplib_reader_under_sanitizers(tmp_path):
    drv = tmp_path / "drv.cpp"
    drv.write_text(DRIVER)
    exe = tmp_path / "drv"
    # host code only: the reader includes the HIP runtime header for shared
    # declarations but calls nothing on a device
    subprocess.run(["g++", "-std=c++17", "-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"), "-I", ROCM_INC,
                    "-I", os.path.join(ROOT, "logo-tsp-solver_amd", "csrc"),
                    os.path.join(ROOT, "logo-tsp-solver_amd", "csrc", "l2o_tsplib.cpp"), str(drv), "-o", str(exe)],
                   check=True)
    files, expect = [], []
    for name in sorted(os.listdir(GOLD)):
        if name.endswith(".tsp.xz"):
            p = tmp_path / name[:-3]
            with open(os.path.join(GOLD, name), "rb") as f:
                p.write_bytes(lzma.decompress(f.read()))
            files.append(str(p))
            # rat195.tsp as the reference ships it is an HTML page
            expect.append(("tsp_bad" if name.startswith("rat195") else "tsp_ok", None))
    for name, text in BAD_TSP.items():
        p = tmp_path / name
        p.write_text(text)
        files.append(str(p))
        expect.append(("tsp_bad", None))
    p = tmp_path / "fraction"  # non-integral: ECOORD unless rounding is asked for
    p.write_text(H.format(d=3, w="EUC_2D") + "1 0.5 0\n2 1.5 1\n3 -2.5 2\nEOF\n")
    files.append(str(p))
    expect.append(("tsp_round", None))
    p = tmp_path / "binary"
    p.write_bytes(bytes(range(256)) * 64)
    files.append(str(p))
    expect.append(("tsp_bad", None))
    for name, (text, code) in TOURS.items():
        p = tmp_path / name
        p.write_text(text)
        files.append(str(p))
        expect.append(("tour", code))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([str(exe)] + files, capture_output=True, text=True, env=env, timeout=300)
    assert out.returncode == 0, out.stderr[-4000:]
    lines = out.stdout.strip().split("\n")
    assert len(lines) == len(files)
    for f, line, (kind, code) in zip(files, lines, expect):
        rc, rr, rc2, n, rt = (int(v) for v in line.split())
        if kind == "tsp_ok":  # (a file with non-integral coordinates needs rounding)
            assert rr == 0 and rc2 == 0 and n > 0 and rc in (0, -2), (f, line)
        elif kind == "tsp_round":
            assert rc == -2 and rr == 0 and rc2 == 0 and n == 3, (f, line)
        elif kind == "tsp_bad":
            assert rc != 0 and (rr != 0 or rc2 != 0), (f, line)
        else:
            assert rt == code, (f, line)
