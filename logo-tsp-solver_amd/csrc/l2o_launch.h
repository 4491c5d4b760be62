// l2o_launch.h -- host-side launch wrappers for the kernels in l2o_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "l2o_state.h"

namespace l2o {

// Edge-length metrics (TSPLIB EDGE_WEIGHT_TYPE), for integer s = dx^2 + dy^2:
//   EUC_2D  nint(sqrt(s))
//   CEIL_2D ceil(sqrt(s))
//   ATT     ceil(sqrt(s / 10)): TSPLIB's t = nint(r), r = sqrt(s/10), plus 1
//           when t < r, equals ceil(r) for every s (when r is not an integer,
//           nint(r) < r exactly when the fraction of r is below 1/2)
enum : uint32_t { L2O_EUC_2D = 0, L2O_CEIL_2D = 1, L2O_ATT = 2 };

// Exact metric length.  With f = 10 for ATT and 1 otherwise,
// m = floor(sqrt(s / f)) satisfies f m^2 <= s < f (m+1)^2, and
//   EUC_2D  nint(sqrt(s)) = m + [s > m^2 + m]  (sqrt(s) is never m + 1/2)
//   CEIL_2D               = m + [s > m^2]
//   ATT                   = m + [s > 10 m^2]
// The FP64 square root starts m within one of the floor; the integer
// comparisons make it exact (s < 2^45 here).
__host__ __device__ inline int32_t metric_dist(long long s, uint32_t metric)
{
    const long long f = metric == L2O_ATT ? 10 : 1;
    const long long e = metric == L2O_EUC_2D ? 1 : 0;
    long long m = (long long)sqrt((double)s * (metric == L2O_ATT ? 0.1 : 1.0));
    if (f * m * m > s) m -= 1;
    else if (f * (m + 1) * (m + 1) <= s) m += 1;
    return (int32_t)(m + (s > f * m * m + e * m ? 1 : 0));
}

// Tier-1 factor g of the sweep's integer bound.  |dx| + |dy| <= sqrt(2) d for
// the Euclidean length d, and d < D + 1/2 (EUC_2D), d <= D (CEIL_2D),
// d <= sqrt(10) D (ATT), so a candidate with gain <= cap has
//   L1(a,b) + L1(a+1,b+1) <= sqrt(g) (cap + 1 + D[a+1] + D[b+1])
// with g = 2, 2, 20.
__host__ __device__ inline int32_t tier1_g(uint32_t metric) { return metric == L2O_ATT ? 20 : 2; }

// ceil(sqrt(g) * d) for 0 <= d < 2^24, g in {2, 20}.  Neither is a square, so
// for d >= 1 sqrt(g) d is irrational and the ceiling is isqrt(g d^2) + 1;
// isqrt is exact after the correction.
__host__ __device__ inline int32_t ceil_sqrtg(int32_t d, int32_t g)
{
    if (d <= 0) return 0;
    const long long s = (long long)g * d * d;
    long long k = (long long)sqrt((double)s);
    if (k * k > s) k -= 1;
    else if ((k + 1) * (k + 1) <= s) k += 1;
    return (int32_t)(k + 1);
}

__host__ __device__ inline int32_t ceil_sqrt2(int32_t d) { return ceil_sqrtg(d, 2); }

// The sweep's coarse test (tier 0) works on coordinates >> qshift, which must
// fit in 8 bits: the smallest shift that brings the largest translated
// coordinate (below 2^22) under 256.
__host__ __device__ inline uint32_t tier0_shift(uint32_t max_coord)
{
    uint32_t s = 0;
    while ((max_coord >> s) > 255u) ++s;
    return s;
}

// Sweep tile shape: each workgroup has `waves` wave64s, each lane owns `k`
// adjacent candidate columns for `t` rows (the last lane of a wave one fewer:
// its last column is the next wave's first).  A workgroup covers
// waves*(64k-1) columns.
struct sweep_cfg {
    int k;
    int t;
    int waves;
    int cols() const { return waves * (64 * k - 1); }
};

bool sweep_cfg_supported(const sweep_cfg &c);
// compiled shapes: fills up to cap rows {k, t, waves}, returns the total count
int sweep_shapes(int (*out)[3], int cap);
// Sweep variants (same results): the full sweep (every block of rows through
// tier 0), and row pruning, which first skips the blocks of rows a
// bounding-box bound rules out and sends the rest through tier 0
// (L2O_SWEEP_PRUNE) or straight to tier 1 (L2O_SWEEP_PRUNE_T1)
enum { L2O_SWEEP_FULL = 0, L2O_SWEEP_PRUNE = 1, L2O_SWEEP_PRUNE_T1 = 2 };
// one workgroup per item, of the variant `mode`
hipError_t launch_sweep(const sweep_cfg &cfg, const l2o_sweep_item *items, uint32_t nitems,
                        l2o_inst *insts, hipStream_t stream, int mode = L2O_SWEEP_FULL);
hipError_t launch_apply(const l2o_apply_item *items, uint32_t nitems, l2o_inst *insts, unsigned *live,
                        hipStream_t stream);
hipError_t launch_ils(l2o_inst *insts, uint32_t nslots, unsigned *live, hipStream_t stream);
hipError_t launch_group(l2o_inst *insts, uint32_t nslots, unsigned *live, hipStream_t stream);
hipError_t launch_length(const l2o_inst *insts, uint32_t slot, uint32_t n, unsigned long long *out,
                         hipStream_t stream);
hipError_t launch_sqrt_ulp(uint32_t lo, uint32_t hi, unsigned *max_ulp, hipStream_t stream);

}  // namespace l2o
