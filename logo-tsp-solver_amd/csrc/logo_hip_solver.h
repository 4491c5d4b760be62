// logo_hip_solver.h -- the reference-side binding: a LocalSolver backend for
// krocki/LOGO-TSP-Solver that runs the 2-opt local search on this library.
//
// It belongs in the reference tree beside CUDASolver (src/solvers/), includes
// the reference's own twoOptLocalSolver.h and replaces CUDASolver
// (src/solvers/CUDASolver.{h,cpp}) method by method:
//
//   CUDASolver::init           CUDASolver.cpp:52-139    -> logo2opt_create
//   CUDASolver::close          CUDASolver.cpp:141-144   -> logo2opt_destroy
//   CUDASolver::optimizeStep   CUDASolver.cpp:285-386   -> logo2opt_step
//   CUDASolver::benchmark      CUDASolver.cpp:146-283   -> logo2opt_step (timed)
//   TwoOptLocalSolver::optimize twoOptLocalSolver.cpp:22-64
//                                                        -> logo2opt_optimize
//                                                           (apply on the device)
//
// A driver selects it with one line where it constructs its local solver
// (ILSGlobalSolverSequential.cpp:49, ILSGlobalSolverMT.cpp:59):
//     localSolver = new HIPSolver(coords, args, sstm.str(), deviceId, size, threadID);
// and links with -llogo2opt.  city_coords, best2_out and process_time have the
// library's layouts (defs.h:173-198), so the casts below are exact.
//
// Errors: CUDASolver calls exit(-1) on a CUDA error (UtilsCUDA.cpp:318-328);
// this binding throws std::runtime_error with logo2opt_strerror's text.
#ifndef LOGO_HIP_SOLVER_H
#define LOGO_HIP_SOLVER_H

#include <stdexcept>
#include <string>

#include <twoOptLocalSolver.h>

#include "logo2opt.h"

class HIPSolver : public TwoOptLocalSolver {
public:
    HIPSolver(city_coords *_coords, cmdArguments *_args, string _description, short _deviceId,
              ROUTE_DATA_TYPE _size, unsigned _tID = 0)
        : TwoOptLocalSolver(_coords, _args, _description, _deviceId, _size, _tID)
    {
        init();
    }
    ~HIPSolver() { close(); }

    void init()
    {
        if (!ctx) check(logo2opt_create((int)deviceId, size, (const logo2opt_city_coords *)coords, &ctx),
                        "logo2opt_create");
    }

    void close()
    {
        if (ctx) {
            logo2opt_destroy(ctx);
            ctx = NULL;
        }
    }

    // The whole local search on the device.  bestLength is only read by
    // TwoOptLocalSolver::optimize (for its progress trace) and is left as it
    // is: the ILS drivers compare the returned tour's length with its last
    // entry (ILSGlobalSolverSequential.cpp:74-92, ILSGlobalSolverMT.cpp:137-178).
    void optimize(vector<ROUTE_DATA_TYPE> &route, vector<unsigned long> &bestLength)
    {
        (void)bestLength;
        uint64_t len = 0, moves = 0;
        check(logo2opt_optimize(ctx, &route[0], &len, &moves, NULL, 0, 64), "logo2opt_optimize");
    }

    // one sweep of `route` with its times: kernel_time from HIP events,
    // HtD_time the host preparation and upload, DtH_time the result read back
    struct process_time benchmark(vector<ROUTE_DATA_TYPE> &route)
    {
        logo2opt_process_time t;
        logo2opt_best2_out o;
        check(logo2opt_step(ctx, &route[0], &o, &t), "logo2opt_step");
        struct process_time r;
        r.kernel_time = t.kernel_time;
        r.HtD_time = t.HtD_time;
        r.DtH_time = t.DtH_time;
        r.out.i = o.i;
        r.out.j = o.j;
        r.out.minchange = o.minchange;
        return r;
    }

protected:
    struct best2_out optimizeStep(const vector<ROUTE_DATA_TYPE> &route)
    {
        logo2opt_best2_out o;
        check(logo2opt_step(ctx, &route[0], &o, NULL), "logo2opt_step");
        struct best2_out r;
        r.i = o.i;
        r.j = o.j;
        r.minchange = o.minchange;
        return r;
    }

private:
    static void check(int rc, const char *what)
    {
        if (rc != LOGO2OPT_OK) throw std::runtime_error(std::string(what) + ": " + logo2opt_strerror(rc));
    }

    logo2opt_ctx *ctx = NULL;
};

#endif
