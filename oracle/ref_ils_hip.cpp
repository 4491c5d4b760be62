// ref_ils_hip.cpp -- runs the reference's own ILS drivers with the library as
// their local solver, through the reference-side binding
// (logo-tsp-solver_amd/csrc/logo_hip_solver.h).
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference sources where they lie (src/solvers/ILSGlobalSolverSequential.cpp,
// ILSGlobalSolverMT.cpp, ILSGlobalSolver.cpp, cpuSolver.cpp, cpuMTSolver.cpp,
// twoOptLocalSolver.cpp, src/utils/common.c) into oracle/_ref/ref_ils_hip,
// linked with liblogo2opt.so.  Nothing of the reference is copied here; this
// file only derives from its classes.  tests/test_gpu_ref_ils.py runs it on
// the GPU and checks its output against the oracle.
//
//   ref_ils_hip <instance.bin> <srand seed> <target length>
//
// instance.bin: uint32 n, n float pairs (city_coords), n uint32 (start tour).
// Output lines (whitespace separated, first token is the tag):
//   once   <lengths...>    ILSGlobalSolverSequential::optimize(route, 1)
//   seq    <lengths...>    ILSGlobalSolverSequential::optimize(route) with
//                          srand(seed) and the stop rule "best <= target"
//   mt     <lengths...>    ILSGlobalSolverMT::optimize(route), one thread,
//                          comm = 1, same seed and target
//   <tag>_ok <0|1>         routeLength(bestGlobalMinima.back()) ==
//                          bestGlobalMinimaLengths.back() and the sizes agree
//   <tag>_tour <ids...>    bestGlobalMinima.back()
//   bench  <i> <j> <minchange> <kernel_s> <htd_s> <dth_s>   benchmark(start)
#include <headers.h>
#include <ILSGlobalSolverSequential.h>
#include <ILSGlobalSolverMT.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "logo_hip_solver.h"

namespace {

// The drivers build their local solver in init() from the solver type
// (ILSGlobalSolverSequential.cpp:29-58, ILSGlobalSolverMT.cpp:33-68); these
// subclasses put HIPSolver in its place, which is the one line a maintainer
// adds there.
class SeqHIP : public ILSGlobalSolverSequential {
public:
    SeqHIP(city_coords *c, cmdArguments *a, ROUTE_DATA_TYPE n, unsigned long target)
        : ILSGlobalSolverSequential(c, a, "hip", TYPE_GLOBAL_SOLVER_CPU, n, 0, target, 120.0f, 0.0f, 0)
    {
        delete localSolver;
        localSolver = new HIPSolver(c, a, "hip", 0, n);
    }
    static vector<unsigned long> &lengths() { return bestGlobalMinimaLengths; }
    static vector<vector<ROUTE_DATA_TYPE> > &tours() { return bestGlobalMinima; }
};

class MTHIP : public ILSGlobalSolverMT {
public:
    MTHIP(city_coords *c, cmdArguments *a, ROUTE_DATA_TYPE n, unsigned long target)
        : ILSGlobalSolverMT(c, a, "hip", TYPE_GLOBAL_SOLVER_CPU, n, 0, 0, target, 120.0f, 0.0f, 0, 1, 0, 1, 1)
    {
        delete localSolver;
        localSolver = new HIPSolver(c, a, "hip", 0, n, 0);
    }
};

void print_list(const char *tag, const vector<unsigned long> &v)
{
    std::printf("%s", tag);
    for (unsigned long x : v) std::printf(" %lu", x);
    std::printf("\n");
}

void report(const char *tag, city_coords *coords)
{
    vector<unsigned long> &L = SeqHIP::lengths();
    vector<vector<ROUTE_DATA_TYPE> > &T = SeqHIP::tours();
    print_list(tag, L);
    int ok = !L.empty() && L.size() == T.size() && routeLength(T.back(), coords) == L.back();
    std::printf("%s_ok %d\n%s_tour", tag, ok, tag);
    for (ROUTE_DATA_TYPE c : T.back()) std::printf(" %u", c);
    std::printf("\n");
    L.clear();
    T.clear();
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s instance.bin seed target\n", argv[0]);
        return 2;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t n = 0;
    if (std::fread(&n, 4, 1, f) != 1 || n < 8) return 2;
    std::vector<city_coords> coords(n);
    vector<ROUTE_DATA_TYPE> start(n);
    if (std::fread(coords.data(), sizeof(city_coords), n, f) != n || std::fread(start.data(), 4, n, f) != n)
        return 2;
    std::fclose(f);
    const unsigned seed = (unsigned)std::strtoul(argv[2], nullptr, 10);
    const unsigned long target = std::strtoul(argv[3], nullptr, 10);
    cmdArguments args;
    std::memset(&args, 0, sizeof(args));
    args.solution = target;
    args.pthreads = 1;
    initStartTime();
    try {
        {
            SeqHIP ils(coords.data(), &args, n, target);
            vector<ROUTE_DATA_TYPE> r = start;
            ils.optimize(r, 1);
            report("once", coords.data());
            process_time t = ils.benchmark(start);
            std::printf("bench %u %u %d %.9f %.9f %.9f\n", t.out.i, t.out.j, t.out.minchange, t.kernel_time,
                        t.HtD_time, t.DtH_time);
            std::srand(seed);
            r = start;
            ils.optimize(r, 0);
            report("seq", coords.data());
        }
        {
            // ILSGlobalSolverMT's constructor runs init() before it sets
            // setAffinity and maxCoresUsed (ILSGlobalSolverMT.h:56-64), and
            // init() reads both (ILSGlobalSolverMT.cpp:70-80): the object is
            // built in zeroed storage so that they read 0 (no affinity).
            alignas(MTHIP) static unsigned char mem[sizeof(MTHIP)];
            std::memset(mem, 0, sizeof(mem));
            MTHIP *ils = new (mem) MTHIP(coords.data(), &args, n, target);
            std::srand(seed);
            vector<ROUTE_DATA_TYPE> r = start;
            ils->optimize(r, 0);
            report("mt", coords.data());
            ils->~MTHIP();
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
