/*
 * logo2opt.h -- C ABI of the MI355X (gfx950) 2-opt local-search library.
 *
 * Drop-in boundary for the GPU 2-opt step of krocki/LOGO-TSP-Solver.  The
 * reference reaches its GPU backend through the LocalSolver virtual interface
 * (src/solvers/localSolver.h:116-125) implemented by CUDASolver
 * (src/solvers/CUDASolver.h:33-120, src/solvers/CUDASolver.cpp:48-434).
 * The entry points below replace that class method by method; the C++
 * adapter in logo-tsp-solver_amd/csrc/hip_local_solver.h maps them back onto
 * the reference's method names.
 *
 * Conventions
 *   - Plain C structs with the reference's layouts: logo2opt_city_coords is
 *     city_coords (src/defs.h:173-178), logo2opt_best2_out is best2_out
 *     (src/defs.h:181-187), logo2opt_process_time is process_time
 *     (src/defs.h:190-198).
 *   - Coordinates must be integral with x and y spans below 2^22 (they are
 *     translated so the minimum is 0); edge lengths are exact TSPLIB lengths:
 *     EUC_2D nint(sqrt(dx^2+dy^2)) by default (the reference's metric), or
 *     CEIL_2D / ATT chosen per context or job (LOGO2OPT_METRIC_*).  Tours are
 *     permutations of 0..n-1, 3 <= n <= 2^20.
 *   - A move (i,J) names tour positions, 1 <= i, i+2 <= J <= n-1.  It is
 *     applied as rotate+reverse: r' = r[i..J-1] ++ r[i-1..0] ++ r[n-1..J]
 *     (src/solvers/twoOptLocalSolver.cpp:41-48).  Among the minimal gains the
 *     lexicographically lowest (i,J) wins (src/solvers/cpuSolver.cpp:140-145).
 *   - Every call returns 0 or a negative LOGO2OPT_E* code; nothing exits or
 *     throws across the ABI (the reference calls exit(-1) on CUDA errors,
 *     src/utils/UtilsCUDA.cpp:318-328).
 *   - A context or engine is confined to one host thread at a time, owns its
 *     own HIP stream and keeps its device buffers resident between calls.
 *     Host arrays stay owned by the caller; the library copies them.
 */
#ifndef LOGO2OPT_H
#define LOGO2OPT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference data layouts ------------------------------------------- */
typedef struct { float x; float y; } logo2opt_city_coords;           /* defs.h:173-178 */
typedef struct { uint32_t i; uint32_t j; int32_t minchange; } logo2opt_best2_out; /* defs.h:181-187 */
typedef struct {                                                      /* defs.h:190-198 */
    double kernel_time;
    double HtD_time;
    double DtH_time;
    logo2opt_best2_out out;
} logo2opt_process_time;

/* ---- error codes ------------------------------------------------------- */
#define LOGO2OPT_OK          0
#define LOGO2OPT_EINVAL     -1  /* bad argument (null pointer, n out of range) */
#define LOGO2OPT_ECOORD     -2  /* coordinates not integral or span >= 2^22    */
#define LOGO2OPT_ETOUR      -3  /* route is not a permutation of 0..n-1        */
#define LOGO2OPT_ENOMEM     -4  /* host or device allocation failed           */
#define LOGO2OPT_EDEVICE    -5  /* a HIP runtime call failed                   */
#define LOGO2OPT_EAGAIN     -6  /* engine full: HBM budget or slot limit       */
#define LOGO2OPT_ESTATE     -7  /* call not valid in the current state         */
#define LOGO2OPT_ENOTFOUND  -8  /* unknown slot / request id                   */
#define LOGO2OPT_EFORMAT    -9  /* not a supported TSPLIB .tsp / .tour file    */
#define LOGO2OPT_EIO       -10  /* file could not be opened, read or written   */
#define LOGO2OPT_ESIZE     -11  /* n above LOGO2OPT_NOALLOC_MAX_N (no-alloc call) */
const char *logo2opt_strerror(int code);

#define LOGO2OPT_MAX_N (1u << 20)
#define LOGO2OPT_KICK_DOUBLE_BRIDGE 0 /* build-defined seeded double-bridge        */
#define LOGO2OPT_KICK_REFERENCE     1 /* rotate+reverse, ILSGlobalSolver.cpp:37-48 */

/* Edge lengths (TSPLIB EDGE_WEIGHT_TYPE) for s = dx^2 + dy^2:
 *   EUC_2D  nint(sqrt(s))      the reference's metric (headers.h:126-142,
 *                              exact here)
 *   CEIL_2D ceil(sqrt(s))      e.g. pla7397, pla33810, pla85900
 *   ATT     ceil(sqrt(s/10))   TSPLIB's pseudo-Euclidean distance (att532)
 * The reference reads every file as EUC_2D (tsp_loader.c:49-162 ignores
 * EDGE_WEIGHT_TYPE); CEIL_2D and ATT have no counterpart there. */
#define LOGO2OPT_METRIC_EUC_2D  0
#define LOGO2OPT_METRIC_CEIL_2D 1
#define LOGO2OPT_METRIC_ATT     2

/* ======================================================================== */
/* Single-instance context: the CUDASolver replacement                      */
/* ======================================================================== */
typedef struct logo2opt_ctx logo2opt_ctx;

/* CUDASolver::init(dev) (CUDASolver.cpp:52-139): select the device, allocate
 * resident buffers and copy the coordinates once. */
int logo2opt_create(int device, uint32_t n, const logo2opt_city_coords *coords,
                    logo2opt_ctx **out);
/* logo2opt_create with the edge-length metric (LOGO2OPT_METRIC_*) that every
 * call on the context uses; logo2opt_create is the EUC_2D case. */
int logo2opt_create_metric(int device, uint32_t n, const logo2opt_city_coords *coords,
                           int metric, logo2opt_ctx **out);
/* CUDASolver::close / cleanCUDA (CUDASolver.cpp:141-144,388-401). */
int logo2opt_destroy(logo2opt_ctx *ctx);

/* CUDASolver::optimizeStep (CUDASolver.cpp:285-386): one best-improvement
 * sweep of `route`, no apply.  out->minchange = 0 (and i = j = 0) when no
 * improving move exists; otherwise out->i < out->j.  t (optional) receives
 * times in seconds like CUDASolver::benchmark (CUDASolver.cpp:146-283):
 * kernel_time is the sweep kernel's duration from HIP events; HtD_time is the
 * host wall time of building the tour-ordered points (with their exact edge
 * lengths) and uploading them and the work list, synchronously; DtH_time the
 * host wall time of reading the result back. */
int logo2opt_step(logo2opt_ctx *ctx, const uint32_t *route, logo2opt_best2_out *out,
                  logo2opt_process_time *t);

/* TwoOptLocalSolver::optimize (twoOptLocalSolver.cpp:22-64) run on the
 * device: sweep, apply on the device, repeat to a local optimum.
 * route_inout: start tour in, local optimum out.  length (optional): final
 * uint64 length.  moves (optional): applied moves.  log (optional): the first
 * min(moves, log_cap) applied moves (i, J, gain) in order, log_cap < 2^32
 * (LOGO2OPT_EINVAL above); the context grows its device buffers to fit.
 * sync_every = 1 reproduces the reference's one-round-trip-per-move loop;
 * sync_every = K > 1 replays K sweep+apply steps as one captured HIP graph
 * between host checks. */
int logo2opt_optimize(logo2opt_ctx *ctx, uint32_t *route_inout, uint64_t *length,
                      uint64_t *moves, logo2opt_best2_out *log, uint64_t log_cap,
                      uint32_t sync_every);

/* ILSGlobalSolverSequential::optimize (ILSGlobalSolverSequential.cpp:60-100)
 * with a fixed round count: LO, then `rounds` times kick(best) -> LO ->
 * accept iff strictly shorter.  Everything stays on the device.  route_inout
 * receives the best tour; round_lengths (optional, rounds+1 entries) the best
 * length after the first LO and after every round. */
int logo2opt_ils(logo2opt_ctx *ctx, uint32_t *route_inout, uint64_t seed, uint32_t rounds,
                 int kick, uint64_t *best_length, uint64_t *round_lengths);

/* ILSGlobalSolverSequential::optimize with the reference's stop rules
 * (ILSGlobalSolverSequential.cpp:60-100; solution, error and timelimit of
 * ILSGlobalSolver.h:160-190).  After the first LO and after every round the
 * search ends once
 *   - the best length L satisfies !(solution * (1.0f + err) < (float)L),
 *     evaluated in float as the reference's loop condition is;
 *   - time_limit_s > 0 seconds have passed since the call began (checked by
 *     the host every 64 device steps; the round in progress then completes,
 *     as the reference checks its clock after each round); or
 *   - max_rounds rounds have run.
 * route_inout receives the best tour, round_lengths (optional, max_rounds+1
 * entries) the best length after the first LO and each round run,
 * rounds_done (optional) the rounds run. */
int logo2opt_ils_until(logo2opt_ctx *ctx, uint32_t *route_inout, uint64_t seed, uint32_t max_rounds,
                       int kick, uint64_t solution, float err, double time_limit_s,
                       uint64_t *best_length, uint64_t *round_lengths, uint32_t *rounds_done);

/* The job's stop_below for the reference's target rule: one more than the
 * largest L with !(solution * (1.0f + err) < (float)L) (no counterpart in
 * the reference: the conversion of its loop condition to an integer bound;
 * 0 for an unreachable target). */
uint64_t logo2opt_stop_below(uint64_t solution, float err);

/* ILSGlobalSolverMT::optimize with comm = 1 and `chains` threads
 * (ILSGlobalSolverMT.cpp:86-195): chains of ILS on the same instance that
 * share a global best, here race free.  All chains start from route_inout and
 * advance round by round on the device; after every round the group best is
 * the smallest best length (lowest chain on ties); a chain whose local search
 * improved its own best kicks that tour next, any other chain kicks the group
 * best (the reference's "route = bestGlobalMinima.back()").  Chain c draws
 * its kicks from seed + c.  route_inout receives the group best tour,
 * round_lengths (optional, rounds+1 entries) the group best after the first
 * local search and after every round.  The result does not depend on timing.
 * rounds >= 1, 1 <= chains <= 4096. */
int logo2opt_ils_multistart(logo2opt_ctx *ctx, uint32_t *route_inout, uint64_t seed, uint32_t rounds,
                            uint32_t chains, int kick, uint64_t *best_length, uint64_t *round_lengths);

/* logo2opt_step without a context and without allocating memory.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same inputs, checks and result as logo2opt_create_metric +
 * logo2opt_step; the device work uses one fixed array of the library (about
 * 3 MiB per device, held until the process ends) instead of a context's
 * buffers, so calls that overlap on one device race: serialise them.
 * n <= LOGO2OPT_NOALLOC_MAX_N, else LOGO2OPT_ESIZE.  Synchronous (returns
 * when *out is written). */
#define LOGO2OPT_NOALLOC_MAX_N 65536u
int logo2opt_step_noalloc(int device, uint32_t n, const logo2opt_city_coords *coords, int metric,
                          const uint32_t *route, logo2opt_best2_out *out);

/* routeLength (src/headers.h:170-189), exact, computed on the device. */
int logo2opt_route_length(logo2opt_ctx *ctx, const uint32_t *route, uint64_t *length);

typedef struct {
    uint64_t sweeps;        /* sweeps since create (moves + 1 per local search)  */
    uint64_t moves;         /* applied 2-opt moves                                */
    uint64_t evals;         /* candidate moves evaluated, (n-2)(n-3)/2 per sweep  */
    uint64_t fallbacks;     /* candidates re-evaluated on the exact FP64 path     */
    double   sweep_ms;      /* device time of timed sweep kernels (HIP events)    */
    double   wall_ms;       /* host wall time inside optimize/ils/run             */
    uint64_t sweep_launches;/* sweep kernel launches (timed mode: all are timed)  */
} logo2opt_stats;
int logo2opt_get_stats(logo2opt_ctx *ctx, logo2opt_stats *st);

/* Timing mode (on != 0): every sweep launch of optimize/ils is bracketed by
 * HIP events on the context's stream and summed into stats.sweep_ms; HIP
 * graphs are not used while it is on. */
int logo2opt_set_timing(logo2opt_ctx *ctx, int on);

/* Row pruning for this context's sweeps, including its multi-start ILS
 * chains; mode as for logo2opt_engine_set_pruning. */
int logo2opt_set_pruning(logo2opt_ctx *ctx, int mode);

/* Times `sweeps` back-to-back launches of the sweep kernel on `route`
 * (no apply) with HIP events on the context's stream: avg_ms receives the
 * average kernel duration.  Used for the roofline figure. */
int logo2opt_time_sweep(logo2opt_ctx *ctx, const uint32_t *route, uint32_t sweeps,
                        double *avg_ms);

/* Accuracy of the FP32 square root the sweep kernel uses (v_sqrt_f32): the
 * largest distance in ulps from the correctly rounded result over the float
 * bit patterns [lo_bits, hi_bits).  The sweep's exactness bound assumes at
 * most 3 ulps; the test suite checks it over [1, 2^46).  No reference
 * counterpart (the reference relies on --use_fast_math sqrtf, Makefile:84). */
int logo2opt_sqrt_ulp_check(int device, uint32_t lo_bits, uint32_t hi_bits, uint32_t *max_ulp);

/* ======================================================================== */
/* Engine: many resident instances, batched sweeps, device ILS               */
/* ======================================================================== */
typedef struct logo2opt_engine logo2opt_engine;

typedef struct {
    uint32_t ils_rounds;   /* 0: local search only; >0: ILS with this many kicks */
    uint32_t kick;         /* LOGO2OPT_KICK_*                                     */
    uint64_t seed;         /* kick generator seed                                 */
    uint32_t log_cap;      /* moves to log (0 = none)                             */
    uint32_t metric;       /* LOGO2OPT_METRIC_* (0 = EUC_2D)                      */
    uint64_t stop_below;   /* ILS: stop after the first local search or round     */
                           /* whose best length is below this (0 = never)         */
} logo2opt_job;

typedef struct {
    uint64_t length;       /* final length (best length for ILS)                 */
    uint64_t moves;        /* applied moves (all local searches)                 */
    uint64_t sweeps;
    uint64_t fallbacks;
    uint32_t rounds_done;  /* ILS rounds completed                                */
    uint32_t done;         /* 1 when finished                                     */
} logo2opt_result;

/* hbm_budget_bytes bounds the resident instance memory (0 = 1 GiB). */
int logo2opt_engine_create(int device, uint64_t hbm_budget_bytes, uint32_t max_slots,
                           logo2opt_engine **out);
int logo2opt_engine_destroy(logo2opt_engine *e);
/* Bytes of device memory an instance of n cities with `job` occupies. */
uint64_t logo2opt_engine_bytes_for(uint32_t n, const logo2opt_job *job);
/* Admit an instance: copies coordinates and start tour to the device.
 * Returns LOGO2OPT_EAGAIN (nothing changed) when the HBM budget or the slot
 * count is exhausted for now, and LOGO2OPT_ENOMEM when the instance needs
 * more than the whole budget (logo2opt_engine_bytes_for), so it can never be
 * admitted.  The new slot joins the next batched step. */
int logo2opt_engine_add(logo2opt_engine *e, uint32_t n, const logo2opt_city_coords *coords,
                        const uint32_t *route, const logo2opt_job *job, int32_t *slot);
/* Run up to max_steps batched steps (each: one sweep of every active
 * instance, the applies, the ILS bookkeeping), checking with the host every
 * `sync_every` steps.  Returns when every slot is done or max_steps ran;
 * *live receives the number of unfinished slots. */
int logo2opt_engine_run(logo2opt_engine *e, uint32_t max_steps, uint32_t sync_every,
                        uint32_t *live);
/* Non-blocking status for a slot (reads device state). */
int logo2opt_engine_result(logo2opt_engine *e, int32_t slot, uint32_t *route_out,
                           logo2opt_result *res, logo2opt_best2_out *log);
/* ILS slots: best length after the first local search and after each round
 * completed so far (at most cap values). */
int logo2opt_engine_round_lengths(logo2opt_engine *e, int32_t slot, uint64_t *out, uint32_t cap);
/* Free a slot's device memory (finished or not). */
int logo2opt_engine_remove(logo2opt_engine *e, int32_t slot);
/* Multi-start ILS group (see logo2opt_ils_multistart): `chains` chains of one
 * instance in consecutive slots first_slot .. first_slot+chains-1; job must
 * ask for ILS (ils_rounds >= 1); chain c uses seed job->seed + c.  Each slot
 * reports its own chain through engine_result / round_lengths (the group best
 * per round); removing any chain removes the whole group.  EAGAIN when no
 * run of `chains` free slots or not enough HBM budget is left. */
int logo2opt_engine_add_group(logo2opt_engine *e, uint32_t n, const logo2opt_city_coords *coords,
                              const uint32_t *route, const logo2opt_job *job, uint32_t chains,
                              int32_t *first_slot);
/* The group best (route_out: its tour; res: its length, moves/sweeps/
 * fallbacks summed over the chains, done when every chain is) and, in
 * best_chain, which chain holds it. */
int logo2opt_engine_group_result(logo2opt_engine *e, int32_t first_slot, uint32_t *route_out,
                                 logo2opt_result *res, int32_t *best_chain);
/* Slots finished since the last call (ids written to out, at most cap). */
int logo2opt_engine_poll_done(logo2opt_engine *e, int32_t *out, uint32_t cap, uint32_t *count);
int logo2opt_engine_stats(logo2opt_engine *e, logo2opt_stats *st);
int logo2opt_engine_set_timing(logo2opt_engine *e, int on);
/* on = 0: issue every step as direct launches instead of replaying captured
 * HIP graphs (also LOGO2OPT_GRAPHS=0 in the environment at engine creation). */
int logo2opt_engine_set_graphs(logo2opt_engine *e, int on);
/* Row pruning: sweeps skip the blocks of 4 rows in which no candidate can
 * pass the integer bound (a row's L1 distance to the bounding box of a wave's
 * columns already exceeds it).  The selected moves, tours and lengths are
 * identical; fewer candidates are evaluated, so the sweep's evals/s figure
 * does not apply.  mode: LOGO2OPT_PRUNE_OFF (the default), LOGO2OPT_PRUNE_ALL
 * (every instance), or LOGO2OPT_PRUNE_AUTO (instances of 16,000 cities or
 * more).  On MI355X a lone instance's local search gains from about 16,000
 * cities (4-9% slower below: AUTO); batches of many instances gain at every
 * size measured (ALL).  No counterpart in the reference (its kernels evaluate
 * every pair, opt_kernel.cu:34-262). */
#define LOGO2OPT_PRUNE_OFF  0
#define LOGO2OPT_PRUNE_ALL  1
#define LOGO2OPT_PRUNE_AUTO 2
int logo2opt_engine_set_pruning(logo2opt_engine *e, int mode);

/* ======================================================================== */
/* Server: continuous batching of local-search / ILS requests                */
/* ======================================================================== */
/* No counterpart in the reference (its multi-start driver runs one solver per
 * pthread, ILSGlobalSolverMT.cpp:86-195).  Requests queue FIFO on the host;
 * each pump admits them into engine slots while the HBM budget and slot count
 * allow, advances every admitted request by a batch of steps, and harvests
 * the finished ones, freeing their HBM for the next admissions. */
typedef struct logo2opt_server logo2opt_server;
int logo2opt_server_create(int device, uint64_t hbm_budget_bytes, uint32_t max_slots, uint32_t queue_cap,
                           logo2opt_server **out);
int logo2opt_server_destroy(logo2opt_server *s);
/* Queue a request (copied).  LOGO2OPT_EAGAIN (id = -1) when the queue is full.
 * route may be NULL: the server then builds the greedy nearest-neighbour tour
 * from city 0 (logo2opt_nn_tour) on a host thread of its own while the GPU
 * runs the admitted requests; the request is admitted once its tour is
 * built, so that time is part of its latency. */
int logo2opt_server_submit(logo2opt_server *s, uint32_t n, const logo2opt_city_coords *coords,
                           const uint32_t *route, const logo2opt_job *job, int64_t *id);
/* Admit what fits, run up to `steps` batched steps (host check every
 * sync_every), harvest finished requests. */
int logo2opt_server_pump(logo2opt_server *s, uint32_t steps, uint32_t sync_every, uint32_t *running,
                         uint32_t *queued);
/* Ids completed since the last poll, including requests rejected at
 * admission (their result reports the error). */
int logo2opt_server_poll(logo2opt_server *s, int64_t *ids, uint32_t cap, uint32_t *count);
/* Result of a completed request; latency_s = submit -> done, queue_s = submit -> admission.
 * A request rejected at admission returns its error code here (for example
 * LOGO2OPT_ENOMEM: it needs more than the server's whole HBM budget, or
 * LOGO2OPT_ECOORD / LOGO2OPT_ETOUR for invalid input), with latency_s set;
 * LOGO2OPT_ESTATE while a request is queued or running. */
int logo2opt_server_result(logo2opt_server *s, int64_t id, uint32_t *route_out, logo2opt_result *res,
                           double *latency_s, double *queue_s);
/* Forget a request (cancels it if queued or running). */
int logo2opt_server_release(logo2opt_server *s, int64_t id);
/* Admission limits besides HBM and slots (0 = none): at most max_running
 * requests at once, and at most max_inflight_pairs candidate moves per sweep
 * summed over running requests (the head of the queue is always admitted
 * when nothing runs). */
int logo2opt_server_set_limits(logo2opt_server *s, uint32_t max_running, uint64_t max_inflight_pairs);
/* Admission order.  LOGO2OPT_ORDER_FIFO (the default) admits in arrival
 * order, and a head that does not fit holds back everything behind it.
 * LOGO2OPT_ORDER_SHORTEST admits the request with the fewest cities whose
 * start tour is ready among the first 1,024 queued (earlier arrival on ties,
 * so an admission costs at most 1,024 lookups); a request that later
 * arrivals have overtaken max_bypass times is admitted next (0 = no bound: a
 * steady stream of smaller requests can then hold a large one back for as
 * long as it lasts).  The order matters only when admission is limited (by
 * logo2opt_server_set_limits, the slots or the HBM budget); results do not
 * depend on it.  The reference has no serving layer and no counterpart. */
#define LOGO2OPT_ORDER_FIFO 0
#define LOGO2OPT_ORDER_SHORTEST 1
int logo2opt_server_set_order(logo2opt_server *s, int order, uint32_t max_bypass);
/* Row pruning for the requests' sweeps; mode as for
 * logo2opt_engine_set_pruning (LOGO2OPT_PRUNE_AUTO suits mixed sizes). */
int logo2opt_server_set_pruning(logo2opt_server *s, int mode);
/* counts[6] = {submitted, rejected, admitted, done, queued now, running now} */
int logo2opt_server_counts(logo2opt_server *s, uint64_t *counts);
/* Open-loop replay of `count` requests arriving at arrival_s[q] seconds
 * (ascending) after the start; runs until all finished or horizon_s.
 * latency_out[q]: arrival -> completion (-1 if rejected or unfinished);
 * queue_out[q] (optional): arrival -> admission. */
int logo2opt_server_replay(logo2opt_server *s, uint32_t count, const double *arrival_s, const uint32_t *n,
                           const logo2opt_city_coords *const *coords, const uint32_t *const *routes,
                           const logo2opt_job *jobs, double horizon_s, uint32_t steps, uint32_t sync_every,
                           double *latency_out, double *queue_out, double *elapsed_s);

/* ======================================================================== */
/* Sweep tile shapes (autotuning)                                            */
/* ======================================================================== */
/* A shape is {k, t, waves}: a workgroup of `waves` wave64s whose lanes each
 * own k adjacent candidate columns of the triangle for t rows.  The
 * reference hard-codes <<<64,1024>>> and 2048-point tiles
 * (CUDASolver.cpp:122-128, cuda_defs.h:24-37); here the shape is chosen per
 * n bucket from a table written by the autotuner (tune.py). */
/* Shapes compiled into the library (3 ints per row); count = total rows. */
int logo2opt_tuning_shapes(int32_t *out, uint32_t cap, uint32_t *count);
/* Two tables: table 0 is used while an engine holds fewer than 8 live
 * instances (latency), table 1 for larger batches (throughput).  Replace a
 * table: row r applies to n <= max_n[r] (the last row to all larger n); work
 * lists built afterwards use it. */
int logo2opt_tuning_set(int table, const uint32_t *max_n, const int32_t *shapes, uint32_t rows);
int logo2opt_tuning_get(int table, uint32_t *max_n, int32_t *shapes, uint32_t cap, uint32_t *rows);
/* Load a table file ("max_n k t waves" per line).  On first use the library
 * loads $LOGO2OPT_TUNING / $LOGO2OPT_TUNING_BATCH, else tuning_gfx950.txt /
 * tuning_batch_gfx950.txt next to itself. */
int logo2opt_tuning_load(int table, const char *path);
/* Times `sweeps` launches of the sweep kernels over every live instance of
 * the engine (no apply), HIP events on the engine stream. */
int logo2opt_engine_time_sweep(logo2opt_engine *e, uint32_t sweeps, double *avg_ms);

/* ======================================================================== */
/* Synthetic workloads and construction (host, deterministic)               */
/* ======================================================================== */
/* SplitMix64-seeded xoshiro256** streams; coordinates are interleaved int32. */
void logo2opt_gen_uniform(uint32_t n, uint32_t side, uint64_t seed, int32_t *xy);
void logo2opt_gen_clustered(uint32_t n, uint32_t side, uint32_t nclusters, double sigma_frac,
                            double noise_frac, uint64_t seed, int32_t *xy);
void logo2opt_random_perm(uint32_t n, uint64_t seed, uint32_t *route);
/* Greedy nearest-neighbour tour from `start` (exact squared distances,
 * lowest id on ties) on a uniform grid; replaces routeInit's NN option
 * (construction.c:139-142, kdtree_utils.c:279-301). */
int  logo2opt_nn_tour(uint32_t n, const int32_t *xy, uint32_t start, uint32_t *route);
/* Multiple-fragment (greedy edge) tour, the reference's default construction
 * (defs.h:70 "mf"; multipleFragment_KD, kdtree_utils.c:56-248): edges in
 * increasing (squared length, lower id, higher id) order join cities of degree
 * < 2 in different fragments; the final path is closed.  Exact integer
 * arithmetic, duplicates kept (the reference's kd-tree drops them).
 * route[0] = 0, route[1] = the lower-id neighbour of city 0. */
int  logo2opt_mf_tour(uint32_t n, const int32_t *xy, uint32_t *route);
/* int32 xy -> reference city_coords */
void logo2opt_xy_to_coords(uint32_t n, const int32_t *xy, logo2opt_city_coords *coords);

/* ======================================================================== */
/* TSPLIB files (host)                                                       */
/* ======================================================================== */
typedef struct {
    char     name[64];     /* NAME (truncated)                                    */
    uint32_t n;            /* DIMENSION                                           */
    int32_t  metric;       /* LOGO2OPT_METRIC_* from EDGE_WEIGHT_TYPE             */
    double   offset_x;     /* subtracted from every x (the minimum, rounded when  */
    double   offset_y;     /*   LOGO2OPT_TSP_ROUND is given), and from every y    */
    uint32_t integral;     /* 1 when every coordinate in the file is an integer  */
    uint32_t reserved;
} logo2opt_tsp_info;

/* Round non-integral coordinates to the nearest integer (halves away from
 * zero) instead of failing with LOGO2OPT_ECOORD; lengths are then those of
 * the rounded instance. */
#define LOGO2OPT_TSP_ROUND 1u

/* readFileCoords (src/utils/tsp_loader.c:49-162): reads a TSPLIB .tsp file.
 * The reference skips a fixed number of header lines and always uses EUC_2D;
 * here the header is parsed by keyword in any order ("KEY : value", spaces
 * optional), TYPE must be TSP, EDGE_WEIGHT_TYPE EUC_2D, CEIL_2D or ATT, and
 * NODE_COORD_SECTION must list every id 1..DIMENSION once (any order; decimal
 * or exponent notation).  info is filled whenever the header parses.  When
 * coords is not NULL it receives the n coordinates (by id - 1) translated by
 * (offset_x, offset_y) so the minimum is 0, which leaves every edge length of
 * the three metrics unchanged; cap must be >= n.  Errors: LOGO2OPT_EIO (open
 * or read), LOGO2OPT_EFORMAT (not TSPLIB, unsupported TYPE /
 * EDGE_WEIGHT_TYPE / NODE_COORD_TYPE, bad or missing entries),
 * LOGO2OPT_ECOORD (non-integral coordinates without LOGO2OPT_TSP_ROUND, or a
 * span >= 2^22), LOGO2OPT_EINVAL (null path/info, cap < n with coords). */
int logo2opt_tsp_read(const char *path, uint32_t flags, logo2opt_tsp_info *info,
                      logo2opt_city_coords *coords, uint32_t cap);

/* routeInitFromFile (src/utils/common.c:109-129): reads a tour of n cities
 * into route (0-based ids).  Accepts the TSPLIB form (header, TOUR_SECTION,
 * 1-based ids, terminated by -1 or EOF; a DIMENSION must equal n) and the
 * reference's own header-less form (n 1-based ids, one per line, as
 * saveRouteToAFile writes them; like routeInitFromFile, only the first n ids
 * are read and anything after them is ignored).  Errors: LOGO2OPT_EIO,
 * LOGO2OPT_EFORMAT (too few ids, or a TOUR_SECTION with more than n before
 * its -1; non-numeric entries; TYPE other than TOUR),
 * LOGO2OPT_ETOUR (not a permutation). */
int logo2opt_tour_read(const char *path, uint32_t n, uint32_t *route);

/* saveRouteToAFile (src/utils/common.c:153-192): writes route (0-based ids)
 * as a TSPLIB tour: NAME, COMMENT "length <length>" (when length > 0),
 * TYPE : TOUR, DIMENSION, TOUR_SECTION, the 1-based ids, -1, EOF.  With
 * headerless != 0 only the 1-based ids are written, one per line, like the
 * reference.  The caller chooses the path (the reference appends
 * "_<length>.tour" itself). */
int logo2opt_tour_write(const char *path, const char *name, uint32_t n, const uint32_t *route,
                        uint64_t length, int headerless);

/* Length of route under a metric on the host (exact, uint64); the same
 * function as the device's.  For checking files and tours without a GPU. */
int logo2opt_host_route_length(uint32_t n, const logo2opt_city_coords *coords, int metric,
                               const uint32_t *route, uint64_t *length);

#ifdef __cplusplus
}
#endif
#endif
