// l2o_state.h -- device-resident instance state shared by the host engine and
// the gfx950 kernels.
//
// Per instance (n cities) HBM holds, 16-byte aligned:
//   pts[2][n]  l2o_pt {x, y, D, W}   tour-ordered points (ping-pong): integer
//                                     coordinates shifted into [0, 2^22),
//                                     D[k] = exact metric length of the edge
//                                     P[k-1] -> P[k] (EUC_2D, CEIL_2D or ATT),
//                                     D[0] the closing edge,
//                                     W[k] = ceil(sqrt(g) D[k]), g = 2 or 20
//   tour[2][n] uint32                 city ids in the same order (ping-pong)
//   bpts[n], btour[n]                 best tour of an ILS run
//   log[log_cap]                      applied moves {i, J, gain}
// The reference keeps the tour on the host and re-uploads ordered
// coordinates every move (CUDASolver.cpp:316-329); here both stay resident and
// the move is applied by a device gather.
#pragma once
#include <stdint.h>

#define L2O_SENTINEL 0xFFFFFFFFFFFFFFFFull

// key = ((gain + 2^24) << 40) | (i << 20) | J   for gain < 0, i < J < 2^20.
// Unsigned order of keys is the lexicographic order of (gain, i, J), so the
// minimum key is the reference's strict-'<' scan winner
// (cpuSolver.cpp:140-145) and an atomicMin reduction is order independent.
#define L2O_KEY_GAIN_BIAS (1ll << 24)

enum l2o_phase : uint32_t {
    L2O_FREE = 0,     // slot unused
    L2O_LO = 1,       // local search running: sweep + apply every step
    L2O_LO_DONE = 2,  // local optimum reached, ILS bookkeeping pending
    L2O_DONE = 3,     // finished (result ready)
    L2O_WAIT = 4,     // chain of a group at a local optimum, waiting for the others
    L2O_KICK = 5,     // chain of a group: kick from `kick_src`'s best, then local search
};

#define L2O_NO_GROUP 0xFFFFFFFFu

struct __attribute__((aligned(16))) l2o_pt {
    int32_t x, y;  // coordinates, 0 <= x, y < 2^22
    int32_t d;     // exact metric length of the edge into this point
    int32_t w;     // ceil(sqrt(g) * d), the edge's tier-1 sweep bound
};

struct l2o_move {
    uint32_t i;
    uint32_t j;
    int32_t gain;
};

struct l2o_inst {
    // ---- fixed at admission
    l2o_pt *pts[2];
    uint32_t *tour[2];
    l2o_pt *bpts;
    uint32_t *btour;
    l2o_move *log;
    unsigned long long *rlog;  // ILS: best length after the first LO and each round
    uint32_t n;
    uint32_t log_cap;
    float slack;          // bound on |exact gain - FP32 estimate| (see l2o_kernels.hip)
    uint32_t ils_rounds;  // 0 = local search only
    uint32_t kick;
    uint32_t have_best;
    uint32_t metric;      // edge length: L2O_EUC_2D, L2O_CEIL_2D or L2O_ATT
    uint32_t group_lead;  // multi-start ILS: first slot of the group (L2O_NO_GROUP: none)
    uint32_t group_size;  // chains in the group (slots group_lead .. group_lead+size-1)
    uint32_t improved;    // the chain's last local search improved its own best
    uint32_t kick_src;    // slot whose best tour the next kick starts from
    float est_scale;      // FP32 tier: D ~ sqrt(s) * est_scale + est_add
    float est_add;
    uint32_t qshift;      // tier 0: coordinates >> qshift fit in 8 bits
    unsigned long long stop_below;  // ILS: done once best_length < stop_below (0: never)
    // ---- dynamic
    uint32_t phase;
    uint32_t parity;
    uint32_t apply_done;  // last-workgroup counter of the apply kernel
    uint32_t rounds_done;
    unsigned long long key;
    unsigned long long length;
    unsigned long long best_length;
    unsigned long long moves;
    unsigned long long sweeps;
    unsigned long long fallbacks;
    unsigned long long rng[4];
};

// One workgroup of the sweep kernel: rows a in [a0, a0 + rows) and columns b
// in [b0, b0 + 64 * waves * k) of the candidate triangle a + 2 <= b <= n - 2;
// rows <= T of the launch's shape (the tail of a launch uses shorter items).
struct l2o_sweep_item {
    uint32_t inst;
    uint32_t a0;
    uint32_t b0;
    uint32_t rows;
};

// One workgroup of the apply kernel: elements [chunk*CHUNK, (chunk+1)*CHUNK).
struct l2o_apply_item {
    uint32_t inst;
    uint32_t chunk;
};

#define L2O_MAX_GROUPS 32  // distinct sweep tile shapes in one engine step
#define L2O_APPLY_CHUNK 1024
#define L2O_APPLY_THREADS 256
#define L2O_ILS_THREADS 1024
