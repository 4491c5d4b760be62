// l2o_launch.h -- host-side launch wrappers for the kernels in l2o_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "l2o_state.h"

namespace l2o {

// ceil(sqrt(2) * d) for 0 <= d < 2^24.  For d >= 1, sqrt(2) d is irrational,
// so the ceiling is isqrt(2 d^2) + 1; isqrt is exact after the correction.
__host__ __device__ inline int32_t ceil_sqrt2(int32_t d)
{
    if (d <= 0) return 0;
    const long long s = 2ll * d * d;
    long long k = (long long)sqrt((double)s);
    if (k * k > s) k -= 1;
    else if ((k + 1) * (k + 1) <= s) k += 1;
    return (int32_t)(k + 1);
}

// Sweep tile shape: each workgroup has `waves` wave64s, each lane owns `k`
// adjacent candidate columns for `t` rows.  A workgroup covers 64*waves*k
// columns.
struct sweep_cfg {
    int k;
    int t;
    int waves;
    int cols() const { return 64 * waves * k; }
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
