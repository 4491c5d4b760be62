// hip_local_solver.h -- C++ adapter with the reference's LocalSolver method
// names (src/solvers/localSolver.h:116-125, CUDASolver.h:33-120) over the C
// ABI in include/logo2opt.h.  Header only; link with -llogo2opt.
//
//   reference CUDASolver                  HIPSolver (this file)
//   ctor(coords, args, desc, dev, size)   HIPSolver(coords, size, dev)
//   init(dev)   CUDASolver.cpp:52-139      init()      -> logo2opt_create
//   close()     CUDASolver.cpp:141-144     close()     -> logo2opt_destroy
//   optimizeStep CUDASolver.cpp:285-386    optimizeStep -> logo2opt_step
//   benchmark   CUDASolver.cpp:146-283     benchmark   -> logo2opt_step (timed)
//   TwoOptLocalSolver::optimize            optimize    -> logo2opt_optimize
//     (twoOptLocalSolver.cpp:22-64; host apply per move)   (apply on the device)
//
// Errors: the reference calls exit(-1) (UtilsCUDA.cpp:318-328); the adapter
// throws std::runtime_error carrying logo2opt_strerror().
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "logo2opt.h"

namespace logo2opt {

class HIPSolver {
public:
    HIPSolver(const logo2opt_city_coords *coords, uint32_t size, int deviceId = 0, uint32_t syncEvery = 64)
        : coords_(coords), size_(size), device_(deviceId), sync_every_(syncEvery)
    {
        init();
    }
    ~HIPSolver() { close(); }
    HIPSolver(const HIPSolver &) = delete;
    HIPSolver &operator=(const HIPSolver &) = delete;

    void init()
    {
        if (!ctx_) check(logo2opt_create(device_, size_, coords_, &ctx_), "logo2opt_create");
    }

    void close()
    {
        if (ctx_) {
            logo2opt_destroy(ctx_);
            ctx_ = nullptr;
        }
    }

    // one sweep, no apply; minchange = 0 when the tour is 2-opt optimal
    logo2opt_best2_out optimizeStep(const std::vector<uint32_t> &route)
    {
        logo2opt_best2_out out;
        check(logo2opt_step(ctx_, route.data(), &out, nullptr), "logo2opt_step");
        return out;
    }

    // local search to a 2-opt local optimum; bestLength receives the final
    // length when it improves on its last entry (the reference only reads
    // bestLength for its progress trace, twoOptLocalSolver.cpp:52-61)
    void optimize(std::vector<uint32_t> &route, std::vector<unsigned long> &bestLength)
    {
        uint64_t length = 0, moves = 0;
        check(logo2opt_optimize(ctx_, route.data(), &length, &moves, nullptr, 0, sync_every_), "logo2opt_optimize");
        if (bestLength.empty() || length < bestLength.back()) bestLength.push_back((unsigned long)length);
    }

    // ILS on the device (ILSGlobalSolverSequential.cpp:60-100 with a fixed
    // round count and the seeded double-bridge kick)
    uint64_t ils(std::vector<uint32_t> &route, uint64_t seed, uint32_t rounds,
                 int kick = LOGO2OPT_KICK_DOUBLE_BRIDGE)
    {
        uint64_t best = 0;
        check(logo2opt_ils(ctx_, route.data(), seed, rounds, kick, &best, nullptr), "logo2opt_ils");
        return best;
    }

    logo2opt_process_time benchmark(std::vector<uint32_t> &route)
    {
        logo2opt_process_time t;
        logo2opt_best2_out out;
        check(logo2opt_step(ctx_, route.data(), &out, &t), "logo2opt_step");
        return t;
    }

    uint64_t routeLength(const std::vector<uint32_t> &route)
    {
        uint64_t L = 0;
        check(logo2opt_route_length(ctx_, route.data(), &L), "logo2opt_route_length");
        return L;
    }

private:
    static void check(int rc, const char *what)
    {
        if (rc != LOGO2OPT_OK) throw std::runtime_error(std::string(what) + ": " + logo2opt_strerror(rc));
    }

    const logo2opt_city_coords *coords_;
    uint32_t size_;
    int device_;
    uint32_t sync_every_;
    logo2opt_ctx *ctx_ = nullptr;
};

}  // namespace logo2opt
