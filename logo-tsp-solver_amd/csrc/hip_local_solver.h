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
//   reference file helpers                 free functions (this file)
//   readFileCoords  tsp_loader.c:49-162    readFileCoords    -> logo2opt_tsp_read
//   routeInitFromFile common.c:109-129     routeInitFromFile -> logo2opt_tour_read
//   saveRouteToAFile common.c:153-192      saveRouteToAFile  -> logo2opt_tour_write
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
    // metric: LOGO2OPT_METRIC_EUC_2D (the reference's), _CEIL_2D or _ATT
    HIPSolver(const logo2opt_city_coords *coords, uint32_t size, int deviceId = 0, uint32_t syncEvery = 64,
              int metric = LOGO2OPT_METRIC_EUC_2D)
        : coords_(coords), size_(size), device_(deviceId), sync_every_(syncEvery), metric_(metric)
    {
        init();
    }
    ~HIPSolver() { close(); }
    HIPSolver(const HIPSolver &) = delete;
    HIPSolver &operator=(const HIPSolver &) = delete;

    void init()
    {
        if (!ctx_) check(logo2opt_create_metric(device_, size_, coords_, metric_, &ctx_), "logo2opt_create_metric");
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

    // local search to a 2-opt local optimum.  bestLength is left untouched:
    // TwoOptLocalSolver::optimize only reads it for its progress trace
    // (twoOptLocalSolver.cpp:52-61), and the ILS drivers that pass their
    // global best list compare the new tour's length against its last entry
    // afterwards (ILSGlobalSolverSequential.cpp:74-92,
    // ILSGlobalSolverMT.cpp:137-178): a length pushed here would make every
    // improvement look like a tie and be reverted.  lastLength() and
    // lastMoves() give the result of the last call.
    void optimize(std::vector<uint32_t> &route, std::vector<unsigned long> &bestLength)
    {
        (void)bestLength;
        check(logo2opt_optimize(ctx_, route.data(), &last_length_, &last_moves_, nullptr, 0, sync_every_),
              "logo2opt_optimize");
    }
    uint64_t lastLength() const { return last_length_; }
    uint64_t lastMoves() const { return last_moves_; }

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

    // row pruning of the sweep (no counterpart in the reference): same moves,
    // faster local searches on large instances; see logo2opt_set_pruning
    void setPruning(bool on)
    {
        check(logo2opt_set_pruning(ctx_, on ? LOGO2OPT_PRUNE_ALL : LOGO2OPT_PRUNE_OFF), "logo2opt_set_pruning");
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
    int metric_;
    logo2opt_ctx *ctx_ = nullptr;
    uint64_t last_length_ = 0, last_moves_ = 0;
};

namespace detail {
inline void check(int rc, const std::string &what)
{
    if (rc != LOGO2OPT_OK) throw std::runtime_error(what + ": " + logo2opt_strerror(rc));
}
}  // namespace detail

// readFileCoords (tsp_loader.c:49-162): the coordinates of a TSPLIB .tsp file
// (translated so the minimum is 0), with its EDGE_WEIGHT_TYPE in *metric.
inline std::vector<logo2opt_city_coords> readFileCoords(const char *filename, int *metric = nullptr,
                                                        uint32_t flags = 0)
{
    logo2opt_tsp_info info;
    detail::check(logo2opt_tsp_read(filename, flags, &info, nullptr, 0), std::string("read ") + filename);
    std::vector<logo2opt_city_coords> c(info.n);
    detail::check(logo2opt_tsp_read(filename, flags, &info, c.data(), info.n), std::string("read ") + filename);
    if (metric) *metric = info.metric;
    return c;
}

// routeInitFromFile (common.c:109-129): a tour of `size` cities, TSPLIB or
// header-less form, appended to route as 0-based ids.
inline void routeInitFromFile(std::vector<uint32_t> &route, uint32_t size, const char *filename)
{
    std::vector<uint32_t> r(size);
    detail::check(logo2opt_tour_read(filename, size, r.data()), std::string("read ") + filename);
    route.insert(route.end(), r.begin(), r.end());
}

// saveRouteToAFile (common.c:153-165): writes <filename>_<length>.tour in the
// reference's header-less layout (1-based ids, one per line).
inline void saveRouteToAFile(const std::vector<uint32_t> &route, const char *filename, unsigned long length)
{
    const std::string fn = std::string(filename) + "_" + std::to_string(length) + ".tour";
    detail::check(logo2opt_tour_write(fn.c_str(), filename, (uint32_t)route.size(), route.data(), length, 1),
                  "write " + fn);
}

}  // namespace logo2opt
