// ref_driver.cpp -- exposes the reference's own CPU 2-opt classes through a
// small C ABI so that the oracle's tests can compare against them.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference sources where they lie (src/solvers/cpuSolver.cpp,
// src/solvers/twoOptLocalSolver.cpp, src/solvers/ILSGlobalSolver.cpp,
// src/utils/common.c) into oracle/_ref/libref2opt.so.  Nothing of the
// reference is copied here; this file only calls its classes.
//
// The reference's scalar CPU gain (cpuSolver.cpp:133-138) equals the exact
// integer gain whenever every edge length is an integer (collinear points),
// which is the regime the tests use it in.
#include <headers.h>
#include <cpuSolver.h>
#include <ILSGlobalSolver.h>
#ifdef L2O_REF_MT
#include <cpuMTSolver.h>
#endif

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

// Records every best2_out that TwoOptLocalSolver::optimize receives
// (twoOptLocalSolver.cpp:35 calls the virtual optimizeStep).
class LoggingCPUSolver : public CPUSolver {
public:
    LoggingCPUSolver(city_coords* c, cmdArguments* a, ROUTE_DATA_TYPE n)
        : CPUSolver(c, a, std::string("ref"), 0, n) {}
    struct best2_out optimizeStep(const vector<ROUTE_DATA_TYPE>& r) override {
        struct best2_out o = CPUSolver::optimizeStep(r);
        log.push_back(o);
        return o;
    }
    struct best2_out step(const vector<ROUTE_DATA_TYPE>& r) { return CPUSolver::optimizeStep(r); }
    std::vector<struct best2_out> log;
};

// ILSGlobalSolver::kick is protected (ILSGlobalSolver.h:214).
class KickProbe : public ILSGlobalSolver {
public:
    void optimize(vector<ROUTE_DATA_TYPE>&, int) override {}
    void init() override {}
    void close() override {}
    void doKick(vector<ROUTE_DATA_TYPE>& r) { kick(r); }
};

cmdArguments quietArgs() {
    cmdArguments a;
    std::memset(&a, 0, sizeof(a));
    a.showLocalOptimizationInfo = 0;
    a.solution = 1;
    return a;
}

}  // namespace

extern "C" {

uint64_t ref_route_length(const float* xy, const uint32_t* route, uint32_t n) {
    std::vector<ROUTE_DATA_TYPE> r(route, route + n);
    return routeLength(r, (city_coords*)xy);
}

// one optimizeStep; out = {i, j, minchange}
void ref_step(const float* xy, const uint32_t* route, uint32_t n, int32_t* out) {
    cmdArguments a = quietArgs();
    LoggingCPUSolver s((city_coords*)xy, &a, n);
    std::vector<ROUTE_DATA_TYPE> r(route, route + n);
    struct best2_out o = s.step(r);
    out[0] = (int32_t)o.i;
    out[1] = (int32_t)o.j;
    out[2] = o.minchange;
}

// TwoOptLocalSolver::optimize; log receives {i, j, minchange} per sweep
// (the final, non-improving sweep included).  Returns the sweep count.
uint64_t ref_optimize(const float* xy, uint32_t* route, uint32_t n, int32_t* log,
                      uint64_t cap) {
    cmdArguments a = quietArgs();
    LoggingCPUSolver s((city_coords*)xy, &a, n);
    std::vector<ROUTE_DATA_TYPE> r(route, route + n);
    std::vector<unsigned long> best;
    s.optimize(r, best);
    for (uint32_t k = 0; k < n; k++) route[k] = r[k];
    for (uint64_t k = 0; k < s.log.size() && k < cap; k++) {
        log[3 * k] = (int32_t)s.log[k].i;
        log[3 * k + 1] = (int32_t)s.log[k].j;
        log[3 * k + 2] = s.log[k].minchange;
    }
    return s.log.size();
}

#ifdef L2O_REF_MT
// CPUMTSolver::optimizeStep is protected (cpuMTSolver.h)
class MTProbe : public CPUMTSolver {
public:
    MTProbe(city_coords* c, cmdArguments* a, ROUTE_DATA_TYPE n) : CPUMTSolver(c, a, std::string("ref-mt"), 0, n) {}
    struct best2_out step(const vector<ROUTE_DATA_TYPE>& r) { return CPUMTSolver::optimizeStep(r); }
};

// one CPUMTSolver::optimizeStep (cpuMTSolver.cpp:75-158) over `threads` pthreads
// (the reference's --cores); out = {i, j, minchange}.  Built only into the
// SIMD library, where it is the timed multi-core CPU baseline.
void ref_step_mt(const float* xy, const uint32_t* route, uint32_t n, int32_t threads, int32_t* out) {
    cmdArguments a = quietArgs();
    a.maxCoresUsed = threads;
    a.setAffinity = 0;
    MTProbe s((city_coords*)xy, &a, n);
    std::vector<ROUTE_DATA_TYPE> r(route, route + n);
    struct best2_out o = s.step(r);
    out[0] = (int32_t)o.i;
    out[1] = (int32_t)o.j;
    out[2] = o.minchange;
}
#endif

// srand(seed) then one reference kick
void ref_kick(uint32_t* route, uint32_t n, unsigned seed) {
    KickProbe k;
    std::vector<ROUTE_DATA_TYPE> r(route, route + n);
    srand(seed);
    k.doKick(r);
    for (uint32_t i = 0; i < n; i++) route[i] = r[i];
}

}  // extern "C"
