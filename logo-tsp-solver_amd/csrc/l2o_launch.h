// l2o_launch.h -- host-side launch wrappers for the kernels in l2o_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "l2o_state.h"

namespace l2o {

// Sweep tile shape: each workgroup has `waves` wave64s, each lane walks `k`
// adjacent diagonals for `t` steps.  A workgroup covers 64*waves*k diagonals.
struct sweep_cfg {
    int k;
    int t;
    int waves;
    int diag() const { return 64 * waves * k; }
};

bool sweep_cfg_supported(const sweep_cfg &c);
// compiled shapes: fills up to cap rows {k, t, waves}, returns the total count
int sweep_shapes(int (*out)[3], int cap);
// one workgroup per item
hipError_t launch_sweep(const sweep_cfg &cfg, const l2o_sweep_item *items, uint32_t nitems,
                        l2o_inst *insts, hipStream_t stream);
hipError_t launch_apply(const l2o_apply_item *items, uint32_t nitems, l2o_inst *insts, unsigned *live,
                        hipStream_t stream);
hipError_t launch_ils(l2o_inst *insts, uint32_t nslots, unsigned *live, hipStream_t stream);
hipError_t launch_length(const l2o_inst *insts, uint32_t slot, uint32_t n, unsigned long long *out,
                         hipStream_t stream);
hipError_t launch_sqrt_ulp(uint32_t lo, uint32_t hi, unsigned *max_ulp, hipStream_t stream);

}  // namespace l2o
