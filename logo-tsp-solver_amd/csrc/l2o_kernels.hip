// l2o_kernels.hip -- gfx950 kernels of the 2-opt local-search path.
//
// Replaces the reference's kernel2opt / kernel2opt_extended
// (src/cuda/opt_kernel.cu:34-262, 273-627) and the host-side apply
// (src/solvers/twoOptLocalSolver.cpp:40-50).  Wave64, no MFMA: the sweep is
// integer VALU work (v_sad_u32) with rare FP32/FP64 re-evaluation.
//
// Candidates.  With a = i-1 and b = J-1 a candidate is a pair
// 0 <= a, a+2 <= b <= n-2 with exact gain
//     G = n(a,b) + n(a+1,b+1) - D[a+1] - D[b+1],
// n(x,y) the exact metric length of P[x] -> P[y] (TSPLIB EUC_2D nint(d), or
// CEIL_2D / ATT, per instance),
// D[k] = n(k-1,k)  (cpuSolver.cpp:133-138 in tour-position form).  The sweep
// returns the lexicographic minimum of (G, i, J) over G < 0, the move the
// reference's strict-'<' scan keeps.
//
// Filter tiers, each sound on its own, make a candidate cost about two VALU
// operations:
//  0. Coarse bound on 8-bit coordinates: one v_sad_u8 per candidate decides
//     whether a block of four rows runs tier 1 at all (see the Tier 0 notes
//     before sweep_item).
//  1. Integer L1 bound.  |dx|+|dy| <= sqrt(2) d and d < D + 1/2 (EUC_2D;
//     for CEIL_2D d <= D, for ATT d <= sqrt(10) D), so G <= best-1 implies
//         L1(a,b) + L1(a+1,b+1) <= ceil(sqrt(g)*best) + W[a+1] + W[b+1],
//     W[k] = ceil(sqrt(g) D[k]) (stored with each point), g = 2 (EUC_2D,
//     CEIL_2D) or 20 (ATT); see tier1_g in l2o_launch.h.  All terms are
//     integers: the test is exact, no rounding margin.  A lane owns K+1
//     consecutive columns b and keeps per-column offsets o with
//     o_k + o_{k+1} = -W[b_k + 1]; then V(a,k) = L1(a, b_k) + o_k is two
//     v_sad_u32 (|x-x'| + acc) and the test of candidate (a, b_k) is
//     V(a,k) + V(a+1,k+1) <= ceil(sqrt(g)*best) + W[a+1], whose right side is
//     wave-uniform.  Each V is computed once and used by two candidates.
//  2. FP32 estimate q(a,b) + q(a+1,b+1) - D[a+1] - D[b+1] <= best-1+slack,
//     |exact - estimate| <= slack (per instance, set by the host from the
//     coordinate span), only for lanes flagged by tier 1.
//  3. Exact FP64 sqrt + integer fix-up for the survivors (metric_dist).
// Tiers 1-3 only drop candidates with G > cap (tier 0 only ones that tier 1
// would drop as well), where cap is the smaller of
// (the wave's best exact gain of EARLIER steps) - 1 -- a later step has a
// larger i, so a tie with that best can never win -- and the best gain
// published so far by any wave of the sweep (any i, so taken inclusively);
// candidates of the steps being processed are re-evaluated together.  Every
// candidate that can still be the lexicographic minimum therefore reaches the
// exact pass, and the atomicMin over 64-bit keys makes the grid result exact
// and deterministic whatever the order of waves (the reference's grid
// tie-break is a racy lock, opt_kernel.cu:239-261).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

#include "l2o_launch.h"
#include "l2o_state.h"

namespace l2o {

// ---------------------------------------------------------------------------
// exact metric length of the edge P1 -> P2 (metric_dist in l2o_launch.h)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int exact_dist(int x1, int y1, int x2, int y2, uint32_t metric)
{
    // coordinates lie in [0, 2^22): the differences and s are exact
    const long long dx = (long long)(x1 - x2);
    const long long dy = (long long)(y1 - y2);
    return metric_dist(dx * dx + dy * dy, metric);
}

// FP32 estimate of the metric length: sqrt(s) * scale + add, with scale = 1,
// add = 0 for EUC_2D (D = nint(d)), scale = 1, add = 1/2 for CEIL_2D (D in
// [d, d+1)) and scale = 1/sqrt(10), add = 1/2 for ATT; |D - estimate| <= 1/2
// plus the FP32 error, which the instance's slack bounds.  The integer
// differences convert to FP32 exactly.
__device__ __forceinline__ float est_dist(int x1, int y1, int x2, int y2, float scale, float add)
{
    const float dx = (float)(x1 - x2), dy = (float)(y1 - y2);
    return __builtin_fmaf(__builtin_amdgcn_sqrtf(__builtin_fmaf(dx, dx, dy * dy)), scale, add);
}

// |a - b| + c: one v_sad_u32
__device__ __forceinline__ uint32_t sad(uint32_t a, uint32_t b, uint32_t c)
{
    return (a > b ? a - b : b - a) + c;
}

// ceil(sqrt(g) * best) for best <= 0 (sqrt(g) * best is irrational unless 0)
__device__ __forceinline__ int tier1_base(int best, int g)
{
    return best == 0 ? 0 : 1 - ceil_sqrtg(-best, g);
}

__device__ __forceinline__ unsigned long long make_key(int gain, uint32_t i, uint32_t j)
{
    return ((unsigned long long)(gain + L2O_KEY_GAIN_BIAS) << 40) |
           ((unsigned long long)i << 20) | (unsigned long long)j;
}

__device__ __forceinline__ int key_gain(unsigned long long key)
{
    return (int)((long long)(key >> 40) - L2O_KEY_GAIN_BIAS);
}

// one step of a 64-bit minimum scan over the lanes: DPP CTRL moves both
// halves; lanes the move leaves without a source keep their own value
template <int CTRL, int ROWS>
__device__ __forceinline__ unsigned long long dpp_min_u64(unsigned long long v)
{
    const int lo = (int)(uint32_t)v, hi = (int)(uint32_t)(v >> 32);
    const uint32_t l2 = (uint32_t)__builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWS, 0xF, false);
    const uint32_t h2 = (uint32_t)__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWS, 0xF, false);
    const unsigned long long o = ((unsigned long long)h2 << 32) | l2;
    return o < v ? o : v;
}

// minimum over the 64 lanes, returned wave-uniform (in SGPRs): row_shr 1, 2,
// 4, 8 within each row of 16 lanes, then row_bcast 15 and 31 carry the rows
// below upwards, so lane 63 ends with the minimum (VALU only, no LDS)
__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v)
{
    v = dpp_min_u64<0x111, 0xF>(v);
    v = dpp_min_u64<0x112, 0xF>(v);
    v = dpp_min_u64<0x114, 0xF>(v);
    v = dpp_min_u64<0x118, 0xF>(v);
    v = dpp_min_u64<0x142, 0xA>(v);
    v = dpp_min_u64<0x143, 0xC>(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}

// a wave-uniform 64-bit value moved to SGPRs (readfirstlane returns int: each
// half is taken as unsigned before it is widened)
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// point records in global memory (address space 1: the pointers come from
// the instance record, so the space is stated for global_* instructions)
typedef __attribute__((address_space(1))) l2o_pt gpt;
typedef __attribute__((address_space(1))) uint32_t gu32;

__device__ __forceinline__ l2o_pt ld_pt(const gpt *P, uint32_t k)
{
    l2o_pt r;
    r.x = P[k].x;
    r.y = P[k].y;
    r.d = P[k].d;
    r.w = P[k].w;
    return r;
}

__device__ __forceinline__ void st_pt(gpt *P, uint32_t k, const l2o_pt r)
{
    P[k].x = r.x;
    P[k].y = r.y;
    P[k].d = r.d;
    P[k].w = r.w;
}

__device__ __forceinline__ unsigned wave_sum_u32(unsigned v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// ---------------------------------------------------------------------------
// sweep
// ---------------------------------------------------------------------------
// A work item covers rows a in [a0, a0+T) and columns b in
// [b0, b0 + WAVES (64K - 1)).  Wave w starts at bw = b0 + w (64K - 1); lane l
// keeps the K+1 points P[bl .. bl+K], bl = bw + lK, in registers and owns the
// candidate columns [bl, bl+K) -- except lane 63's last one, which is the
// next wave's first.  Rows stream through the scalar cache (s_load, uniform
// addresses), so the loop reads no LDS and no vector memory.
//
// The offsets form one alternating chain over the wave's columns
// (o_i + o_{i+1} = -W of column i+1), so V(a, k+1) of lane l's column K is
// V(a, 0) of lane l+1: a candidate's second term for k = K-1 comes from the
// next lane through DPP (v_add_u32_dpp wave_shl:1) instead of two more
// v_sad_u32.  Per step and lane: K values V (2K v_sad_u32), K adds (one with
// DPP), a min tree (v_min3) and one compare, ballotted into an SGPR mask.
// Lane 63 has no next lane (DPP reads 0 there); the chain's free start is set
// so that its o of column K-1 is 2^30, which keeps that candidate -- owned by
// the next wave -- out of tier 1.  Waves are independent (no LDS, no
// barrier): each reduces its own key.
//
// Invalid candidates never pass tier 1: a column b > n-2 gets +2^30 in its
// offset pair, and in a wave that meets the diagonal (some b < a+2) each
// candidate is masked by its own step limit.
#define L2O_T1_BIG 0x40000000u

// DPP wave_shl:1 -- lane l reads lane l+1; lane 63 reads 0 (bound_ctrl)
__device__ __forceinline__ uint32_t from_next_lane(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xF, 0xF, true);
}

// inclusive prefix sum over the 64 lanes (mod 2^32) with DPP: shifts by 1, 2,
// 4, 8 within each row of 16 lanes (zero-filled at the row start), then the
// last lane of row 0 / rows 0-1 broadcast into the rows above
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false); // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false); // row_bcast:31
    return v;
}

// maximum over the 64 lanes, returned wave-uniform: the scan pattern above
// with max, where a lane that a shift leaves without a source keeps its value
__device__ __forceinline__ int wave_max_i32(int v)
{
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x111, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x112, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x114, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x118, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xA, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xC, 0xF, false));
    return __builtin_amdgcn_readlane(v, 63);
}

// Row pruning (PRUNE, opt-in).  Every candidate (a, b) of a wave has P[b] and
// P[b+1] inside the wave's bounding box B of its column points and
// W[b+1] <= Wmax, the wave's largest column W, so its tier-1 sum is at least
//   lb(P[a]) + lb(P[a+1]) - Wmax,   lb(p) = L1 distance from p to B.
// A row whose bound exceeds tier 1's threshold tb + W[a+1] has no tier-1 hit;
// tb only falls during a sweep, so rows ruled out with the threshold at the
// wave's start stay ruled out.  Before its row loop each wave tests its rows
// with one lane per block of 4 rows and skips the blocks with no row left.
// A skipped block would have had no tier-1 hit, so the candidates that reach
// the exact pass, and the selected move, are those of the full sweep.  After
// a skipped run of blocks V is recomputed from the next block's first row.
//
// K = 6 and 8 need 86-90 VGPRs unconstrained (5 waves per SIMD); asking for
// 6 waves makes the allocator keep 80 and spill 16-68 bytes per lane outside
// the hot loops instead (the config-2 local search 1.7% faster, the batch
// 1.1%, on one box A/B)
// Tier 0 (the full sweep, and the live blocks of the L2O_SWEEP_PRUNE
// variant).  With X = x >> q (2^q = Q), |x - x'| >= Q |X - X'| -
// (Q - 1), so a candidate's tier-1 sum L1(a,b) + L1(a+1,b+1) - W[b+1] is at
// least Q Sq - 4 (Q - 1) - W[b+1], where Sq is the same L1 sum of the shifted
// coordinates (one v_sad_u8 of four bytes).  With wq = the lane's largest
// ceil(W[b+1] / Q) and wa = ceil(W[a+1] / Q), a tier-1 hit has
//   Sq - wq - wa <= floor((tb + 4 (Q - 1)) / Q),
// so a block of steps in which no candidate meets this holds no tier-1 hit
// and is skipped.  A block is tested with its largest wa for each of its
// rows (a looser, valid test: one subtract per block, then per step K
// v_sad_u8 and K/2 v_min3; with each row's own wa the test costs a subtract
// per step, and re-testing the blocks that pass with them measured slower).
// Each wave stages its rows' packed byte pairs and the blocks' wa in LDS once
// (one lane per block) and reads them back as broadcasts, so the test is VALU
// work only.  Sq <= 1,020 and wq, wa < 2^10; the sum carries a bias of 2^16
// so that it stays positive.
#ifndef L2O_TIER0
#define L2O_TIER0 1
#endif
#ifndef L2O_COOP
#define L2O_COOP 1
#endif
#define L2O_T0_BIAS 0x10000u
#ifndef L2O_WPE6_MIN_K
#define L2O_WPE6_MIN_K 6
#endif
//
// sweep_item evaluates one work item; k_sweep runs one item per workgroup and
// the hardware dispatcher balances the items over the CUs.
template <int K, int T, int WAVES, int MODE>
__device__ __forceinline__ void sweep_item(const l2o_sweep_item it, l2o_inst *__restrict__ insts)
{
    constexpr int UF = 4;  // steps per unrolled iteration
    constexpr bool PRUNE = MODE != L2O_SWEEP_FULL;
    constexpr bool TIER0 = L2O_TIER0 && MODE != L2O_SWEEP_PRUNE_T1;
    static_assert(T % UF == 0, "T must be a multiple of the unroll factor");
    static_assert(!(PRUNE || TIER0) || T <= 64 * UF, "pruning and tier 0 keep one bit per block of rows");
    l2o_inst *in = insts + it.inst;
    // The fields this kernel reads but no wave writes come in one batch of
    // scalar loads before any branch (constant address space: other kernels
    // and the host write them, and the scalar cache starts clean), so the
    // prologue is four dependent loads deep: item, instance, data.
    const __attribute__((address_space(4))) l2o_inst *ci = (const __attribute__((address_space(4))) l2o_inst *)in;
    const uint32_t phase = ci->phase, par = ci->parity;
    const int n = (int)ci->n;
    const float slack = ci->slack;
    const uint32_t metric = ci->metric;
    const float es = ci->est_scale, ea = ci->est_add;
    const uint32_t qs = ci->qshift;
    l2o_pt *const p0 = ci->pts[0], *const p1 = ci->pts[1];
    // (an empty asm that consumes them keeps the loads above the branches)
    asm volatile("" ::"s"(phase), "s"(par), "s"(n), "s"(slack), "s"(p0), "s"(p1), "s"(it.a0), "s"(it.b0),
                 "s"(it.rows), "s"(metric), "s"(es), "s"(ea), "s"(qs));
    if (phase != L2O_LO) return;
    const gpt *__restrict__ P = (const gpt *)(par ? p1 : p0);
    const int a0 = (int)it.a0;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    constexpr int WC = 64 * K - 1;             // candidate columns of a wave
    const int bw = (int)it.b0 + w * WC;       // first column of this wave
    if (bw > n - 2) return;                   // wave-uniform: no candidate column
    const int bl = bw + lane * K;
    // rows a <= n-4 have candidates
    int t_end = min((int)it.rows, n - 3 - a0);
    // a wave whose first column reaches past the last row + 2 never meets the
    // diagonal b = a+1; otherwise rows past its last column - 2 are empty
    const bool diag = bw < a0 + t_end + 1;
    if (diag) t_end = min(t_end, min(bw + WC - 1, n - 2) - 1 - a0);

    // PRUNE, TIER0: the points of block l of the rows for lane l (rows
    // a0+4l .. a0+4l+4), for the row test and the tier-0 staging below,
    // requested before the columns (their addresses do not depend on them);
    // indices are clamped to a0 + t_end <= n - 3
    constexpr bool QROWS = PRUNE || TIER0;
    // COOP (the pruned sweep): the wave's points arrive as contiguous 16-byte
    // loads (one cache line serves 8 lanes, instead of a lane stride of UF or
    // K points) into its LDS stage, and each lane reads its own points back
    // from there: the rows a0 + min(i, tc) for i < 64 MR, then the columns
    // min(bw + i, n-1) for i < 64 MC.  Both segments are requested at once.
    // The stage belongs to this wave alone and later holds its tier-0 block
    // records.  Most pruned waves end after the row test, so the prologue is
    // most of their life: the config-2 pruned search took 0.220 s instead of
    // 0.245 s.  The full sweep keeps per-lane loads (0.296 s with the stage
    // against 0.284 s: its waves live long, and the stage takes 10 KB of LDS
    // per two-wave workgroup instead of 4 KB).  K <= 4 only: the stage grows
    // with 64 K.
    constexpr bool COOP = L2O_COOP && K <= 4 && PRUNE;
    constexpr int MR = (T + 1 + 63) / 64, MC = (64 * K + 1 + 63) / 64;
    constexpr int STG = COOP ? 64 * (MR > MC ? MR : MC) : 0;  // uint4 per wave
    constexpr int T0W = 2 * T;                                  // tier 0: words per wave (8 per block)
    constexpr int T0U = TIER0 ? T0W / 4 : 0;
    constexpr int LW = STG > T0U ? STG : (T0U > 0 ? T0U : 1);
    __shared__ __attribute__((aligned(16))) uint4 wlds[WAVES * LW];
    uint4 *const wst = wlds + w * LW;
    const int tc = max(t_end, 0);
    // Lanes read slots that other lanes wrote: to the compiler, which sees one
    // lane, those addresses are unrelated, so each phase change of the stage
    // is a wave barrier plus a memory clobber that keeps the LDS operations
    // in program order (the hardware then completes them in order).
    auto stage_fence = [&]() __attribute__((always_inline)) {
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
    };
    auto ld4 = [&](int k) __attribute__((always_inline)) {
        return make_uint4((uint32_t)P[k].x, (uint32_t)P[k].y, (uint32_t)P[k].d, (uint32_t)P[k].w);
    };
    uint4 gr[COOP && QROWS ? MR : 1], gc[COOP ? MC : 1];
    if constexpr (COOP) {
        if constexpr (QROWS) {
#pragma unroll
            for (int m = 0; m < MR; ++m) gr[m] = ld4(a0 + min(lane + 64 * m, tc));
        }
#pragma unroll
        for (int m = 0; m < MC; ++m) gc[m] = ld4(min(bw + lane + 64 * m, n - 1));
    }
    int qx[QROWS ? UF + 1 : 1], qy[QROWS ? UF + 1 : 1], qw[QROWS ? UF + 1 : 1];
    if constexpr (QROWS) {
        if constexpr (COOP) {
#pragma unroll
            for (int m = 0; m < MR; ++m) wst[lane + 64 * m] = gr[m];
            stage_fence();
#pragma unroll
            for (int s = 0; s <= UF; ++s) {
                const uint4 v = wst[min(lane * UF + s, tc)];
                qx[s] = (int)v.x;
                qy[s] = (int)v.y;
                qw[s] = (int)v.w;
            }
            stage_fence();  // (the column stage below reuses these slots)
        } else {
#pragma unroll
            for (int s = 0; s <= UF; ++s) {
                const uint32_t q = (uint32_t)min(a0 + lane * UF + s, a0 + max(t_end, 0));
                qx[s] = P[q].x;
                qy[s] = P[q].y;
                qw[s] = P[q].w;
            }
        }
    }
    uint4 cp[COOP ? K + 1 : 1];
    if constexpr (COOP) {
#pragma unroll
        for (int m = 0; m < MC; ++m) wst[lane + 64 * m] = gc[m];
        stage_fence();
#pragma unroll
        for (int j = 0; j <= K; ++j) cp[j] = wst[lane * K + j];
        stage_fence();  // (the tier-0 block records below reuse these slots)
    }

    // lane columns and offsets.  Wn[j] = W of column bl+j, minus 2^30 when
    // the candidate column bl+j-1 lies beyond n-2 (its pair sum gets +2^30).
    // (W first, coordinates after the offsets: fewer registers live at once)
    uint32_t cx[K + 1], cy[K + 1], o[K + 1];
    int wmax = 0;  // PRUNE: the lane's, then the wave's largest W of a column b+1
    int wq = 0;    // TIER0: the lane's largest ceil(W / 2^qs) of a column b+1
    {
        uint32_t Wn[K + 1];
#pragma unroll
        for (int j = 0; j <= K; ++j)
            Wn[j] = (COOP ? cp[j].w : (uint32_t)P[min(bl + j, n - 1)].w) - (bl + j - 1 > n - 2 ? L2O_T1_BIG : 0u);
        // (beyond n-2 Wn is negative as an int and drops out)
        if constexpr (PRUNE) {
#pragma unroll
            for (int j = 1; j <= K; ++j) wmax = max(wmax, (int)Wn[j]);
        }
        if constexpr (TIER0) {
#pragma unroll
            for (int j = 1; j <= K; ++j) wq = max(wq, (max((int)Wn[j], 0) + (1 << qs) - 1) >> qs);
        }
        // The chain o_i = (-1)^i (o_0 - A_i), A_i = sum_{t=1..i} (-1)^t W_t
        // over the wave's columns i = lK + j: A at each lane's first column is
        // an exclusive scan of the lanes' signed sums; o_0 makes o of column
        // 64K-1 (lane 63, k = K-1) equal 2^30.
        const uint32_t sgn_lane = ((lane * K) & 1) ? 0xFFFFFFFFu : 1u;  // (-1)^{lK}
        uint32_t csum = 0;
#pragma unroll
        for (int j = 1; j <= K; ++j) csum += (j & 1) ? 0u - Wn[j] : Wn[j];
        csum *= sgn_lane;
        const uint32_t incl = wave_inclusive_scan(csum);
        const uint32_t excl = incl - csum;                                                // A_{lK}
        const uint32_t a_all = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);       // A_{64K}
        const uint32_t w_last = (uint32_t)__builtin_amdgcn_readlane((int)Wn[K], 63);     // W_{64K}, sign +
        const uint32_t o_first = a_all - w_last - L2O_T1_BIG;               // o_0 of the wave
        o[0] = (o_first - excl) * sgn_lane;
#pragma unroll
        for (int j = 1; j <= K; ++j) o[j] = 0u - Wn[j] - o[j - 1];
        // opaque from here: otherwise the compiler cancels o_k + o_{k+1}
        // inside a candidate's sum and rewrites v_sad_u32 as max/min/sub
#pragma unroll
        for (int j = 0; j <= K; ++j) asm volatile("" : "+v"(o[j]));
    }
#pragma unroll
    for (int j = 0; j <= K; ++j) {
        if constexpr (COOP) {
            cx[j] = cp[j].x;
            cy[j] = cp[j].y;
        } else {
            const uint32_t c = (uint32_t)min(bl + j, n - 1);
            cx[j] = (uint32_t)P[c].x;
            cy[j] = (uint32_t)P[c].y;
        }
    }
    // PRUNE: the wave's bounding box [bx0, bx1] x [by0, by1] of its column points
    int bx0 = 0, bx1 = 0, by0 = 0, by1 = 0;
    if constexpr (PRUNE) {
        int mx = (int)cx[0], Mx = mx, my = (int)cy[0], My = my;
#pragma unroll
        for (int j = 1; j <= K; ++j) {
            mx = min(mx, (int)cx[j]);
            Mx = max(Mx, (int)cx[j]);
            my = min(my, (int)cy[j]);
            My = max(My, (int)cy[j]);
        }
        bx0 = -wave_max_i32(-mx);
        bx1 = wave_max_i32(Mx);
        by0 = -wave_max_i32(-my);
        by1 = wave_max_i32(My);
        wmax = wave_max_i32(wmax);
    }
    // TIER0: column pairs (b, b+1) of the lane as bytes of coordinates >> qs,
    // and the wave's blocks of rows in LDS, 32 bytes each: for block q (rows
    // 4q .. 4q+3) the same bytes of rows (a, a+1), a = a0+4q+s, s = 0..3, then
    // the block's largest wa = ceil(W[a+1] / 2^qs), for the whole blocks (a+1
    // <= a0 + t_end <= n-3).  Only this wave reads its blocks, and LDS
    // operations of one wave complete in order: no barrier.
    uint32_t cpk[TIER0 ? K : 1];
    uint32_t c0 = 0;
    uint4 *const t0blk = wst;  // 2 per block
    auto t0_pack = [&](uint32_t xa, uint32_t ya, uint32_t xb, uint32_t yb) __attribute__((always_inline)) {
        return (xa >> qs) | ((ya >> qs) << 8) | ((xb >> qs) << 16) | ((yb >> qs) << 24);
    };
    // block q from its five points (rows 4q .. 4q+4)
    auto t0_stage = [&](int q, const int *x, const int *y, const int *wr) __attribute__((always_inline)) {
        uint4 v;
        v.x = t0_pack(x[0], y[0], x[1], y[1]);
        v.y = t0_pack(x[1], y[1], x[2], y[2]);
        v.z = t0_pack(x[2], y[2], x[3], y[3]);
        v.w = t0_pack(x[3], y[3], x[4], y[4]);
        t0blk[2 * q] = v;
        const uint32_t mw = (uint32_t)max(max(wr[1], wr[2]), max(wr[3], wr[4]));
        t0blk[2 * q + 1].x = (mw + (1u << qs) - 1) >> qs;
    };
    if constexpr (TIER0) {
#pragma unroll
        for (int k = 0; k < K; ++k) cpk[k] = t0_pack(cx[k], cy[k], cx[k + 1], cy[k + 1]);
        c0 = L2O_T0_BIAS - (uint32_t)wq;
        if constexpr (!PRUNE) {
            // one lane per whole block (PRUNE: after the row test, below)
            if (lane < max(t_end, 0) / UF) t0_stage(lane, qx, qy, qw);
            __builtin_amdgcn_wave_barrier();
        }
    }
    // diagonal waves: candidate (a0+t, bl+k) exists iff t - lim0 <= k
    const int lim0 = bl - 2 - a0;

    // rows: row a0+t at prow[4t .. 4t+3] = {x, y, D, W}
    const __attribute__((address_space(4))) int *__restrict__ prow =
        (const __attribute__((address_space(4))) int *)(P + a0);

    uint32_t V[K];  // V(a, k) of the current row a
    {
        const uint32_t x0 = (uint32_t)prow[0], y0 = (uint32_t)prow[1];
#pragma unroll
        for (int k = 0; k < K; ++k) V[k] = sad(y0, cy[k], sad(x0, cx[k], o[k]));
    }
    // cap = the largest gain still worth an exact evaluation.  It starts at
    // -1 (improving moves only) or at the instance's best key so far, which
    // other waves of this sweep publish as they go; that key may come from a
    // larger i, so it is taken inclusively (a tie at a smaller i wins).  The
    // wave's own best of earlier rows (smaller i) is taken exclusively.
    auto key_now = [&]() __attribute__((always_inline)) -> unsigned long long {
        return uniform_u64(__hip_atomic_load(&in->key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    };
    int cap = -1;
    const int g = tier1_g(metric);
    int tb = 0;                // tier 1: ceil(sqrt(g) * (cap + 1))
    float thr = slack - 1.0f;  // tier 2: cap + slack
    // TIER0: BIAS + floor((tb + 4 (2^qs - 1)) / 2^qs)
    auto t0_of = [&](int b) __attribute__((always_inline)) {
        return (int)L2O_T0_BIAS + ((b + 4 * ((1 << qs) - 1)) >> qs);
    };
    int t0h = t0_of(0);
    auto set_cap = [&](int c) __attribute__((always_inline)) {
        cap = c;
        tb = __builtin_amdgcn_readfirstlane(tier1_base(c + 1, g));
        thr = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int((float)c + slack)));
        if constexpr (TIER0) t0h = t0_of(tb);
    };
    // (PRUNE: read after the row test below, by waves with a live block only;
    // the row test then uses cap = -1, a valid, looser threshold)
    if constexpr (!PRUNE) {
        const unsigned long long g = key_now();
        if (g != L2O_SENTINEL && key_gain(g) < cap) set_cap(key_gain(g));
    }
    // PRUNE: bit q set when block q (rows a0+4q .. a0+4q+3, the last block
    // cut at t_end) holds a row that the bound does not rule out
    unsigned long long live = 0;
    if constexpr (PRUNE) {
        auto lb = [&](int x, int y) __attribute__((always_inline)) {
            return max(max(bx0 - x, x - bx1), 0) + max(max(by0 - y, y - by1), 0);
        };
        const int r0 = lane * UF;
        bool keep = false;
        int lp = lb(qx[0], qy[0]);
#pragma unroll
        for (int s = 0; s < UF; ++s) {
            const int lq = lb(qx[s + 1], qy[s + 1]);
            keep = keep || (r0 + s < t_end && lp + lq - wmax <= tb + qw[s + 1]);
            lp = lq;
        }
        live = __ballot(keep);
        if (!live) return;  // nothing to evaluate, nothing published
        if constexpr (TIER0) {
            // lane l holds the points of block l (rows 4l .. 4l+4), which are
            // unclamped for a whole block
            if (keep && r0 + UF <= t_end) t0_stage(lane, qx, qy, qw);
            __builtin_amdgcn_wave_barrier();
        }
        const unsigned long long g = key_now();
        if (g != L2O_SENTINEL && key_gain(g) < cap) set_cap(key_gain(g));
    }
    unsigned long long lbest = L2O_SENTINEL;
    unsigned long long published = L2O_SENTINEL;
    unsigned fb = 0;

    // Steps t0 .. t0+ns-1 hold a tier-1 hit somewhere in the wave: every lane
    // re-runs tier 1 on its candidates of those steps (same integer values as
    // the main loop), then tiers 2 and 3 on the survivors; then the wave's
    // thresholds are refreshed.
    auto slow = [&](int t0, int ns) __attribute__((always_inline)) {
        for (int s = 0; s < ns; ++s) {
            const int t = t0 + s;
            const int xa = prow[4 * t], ya = prow[4 * t + 1];
            const int xb = prow[4 * t + 4], yb = prow[4 * t + 5], db = prow[4 * t + 6], wr = prow[4 * t + 7];
            const int u = tb + wr;
            const float u2 = thr + (float)db;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint32_t va = sad((uint32_t)ya, cy[k], sad((uint32_t)xa, cx[k], o[k]));
                const uint32_t vb = sad((uint32_t)yb, cy[k + 1], sad((uint32_t)xb, cx[k + 1], o[k + 1]));
                // lane 63's last column belongs to the next wave
                if ((int)(va + vb) <= u && (!diag || t - lim0 <= k) && (k + 1 < K || lane != 63)) {
                    const int dc = P[bl + k + 1].d;
                    const float w2 = (est_dist(xa, ya, (int)cx[k], (int)cy[k], es, ea) +
                                      est_dist(xb, yb, (int)cx[k + 1], (int)cy[k + 1], es, ea)) - (float)dc;
                    if (w2 <= u2) {
                        ++fb;
                        const int gain = exact_dist(xa, ya, (int)cx[k], (int)cy[k], metric) +
                                         exact_dist(xb, yb, (int)cx[k + 1], (int)cy[k + 1], metric) - db -
                                         dc;
                        if (gain < 0) {
                            const uint32_t i = (uint32_t)(a0 + t + 1);
                            const uint32_t j = (uint32_t)(bl + k + 1);
                            const unsigned long long key = make_key(gain, i, j);
                            lbest = key < lbest ? key : lbest;
                        }
                    }
                }
            }
        }
        // a lane that found a key below the wave's last published one makes
        // the wave reduce and publish its minimum at once (so that other
        // waves can tighten their cap); the minimum and the thresholds are
        // uniform and stay in SGPRs
        int c = cap;
        if (__ballot(lbest < published)) {
            const unsigned long long wb = wave_min_u64(lbest);
            published = wb;
            c = min(c, key_gain(wb) - 1);
            if (lane == 0) atomicMin(&in->key, wb);
        }
        const unsigned long long g = key_now();
        if (g != L2O_SENTINEL) c = min(c, key_gain(g));
        if (c != cap) set_cap(c);
    };

    // NS steps from t with rows a0+t+1 .. a0+t+NS given in r (x, y, W per
    // row, wave-uniform): V advances by NS rows; returns the wave's tier-1 hits
    auto block = [&](int t, const int (*r)[3], auto ns_tag, auto mask_tag)
                     __attribute__((always_inline)) -> unsigned long long {
        constexpr int NS = decltype(ns_tag)::value;
        constexpr bool MASK = decltype(mask_tag)::value;
        unsigned long long any = 0;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const uint32_t rx = (uint32_t)r[s][0], ry = (uint32_t)r[s][1];
            const int u = tb + r[s][2];
            uint32_t Vn[K];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                Vn[j] = sad(ry, cy[j], sad(rx, cx[j], o[j]));
                // keeps each V a pair of v_sad_u32 (the last one would
                // otherwise be reassociated into max/min/sub/add3)
                asm volatile("" : "+v"(Vn[j]));
            }
            int m = 0x7FFFFFFF;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                // column K of this lane is column 0 of the next lane
                int S = (int)(V[k] + (k + 1 < K ? Vn[k + 1] : from_next_lane(Vn[0])));
                if constexpr (MASK) S = (t + s - lim0 <= k) ? S : 0x7FFFFFFF;
                m = min(m, S);
            }
            any |= __ballot(m <= u);
#pragma unroll
            for (int k = 0; k < K; ++k) V[k] = Vn[k];
        }
        return any;
    };
    auto load_rows = [&](int t, int (*r)[3], auto ns_tag) __attribute__((always_inline)) {
        constexpr int NS = decltype(ns_tag)::value;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            r[s][0] = prow[4 * (t + s + 1)];
            r[s][1] = prow[4 * (t + s + 1) + 1];
            r[s][2] = prow[4 * (t + s + 1) + 3];
        }
    };

    // the rows of the next iteration are requested before this one computes
    // (the last iteration re-reads its own rows: every address stays < n)
    auto run = [&](auto mask_tag) __attribute__((always_inline)) {
        using UFc = std::integral_constant<int, UF>;
        int t = 0;
        const int t_main = t_end - t_end % UF;
        if (t_main > 0) {
            // two row buffers used in turn (no register copies between blocks)
            int ra[UF][3], rb[UF][3];
            load_rows(0, ra, UFc{});
            for (; t < t_main; t += 2 * UF) {
                load_rows(min(t + UF, t_main - UF), rb, UFc{});
                if (block(t, ra, UFc{}, mask_tag)) slow(t, UF);
                if (t + UF >= t_main) {
                    t += UF;
                    break;
                }
                load_rows(min(t + 2 * UF, t_main - UF), ra, UFc{});
                if (block(t + UF, rb, UFc{}, mask_tag)) slow(t + UF, UF);
            }
        }
        for (; t < t_end; ++t) {
            int r1[1][3];
            load_rows(t, r1, std::integral_constant<int, 1>{});
            if (block(t, r1, std::integral_constant<int, 1>{}, mask_tag)) slow(t, 1);
        }
    };
    // TIER0: every block of UF steps first gets the coarse test of every
    // candidate; only a block in which some candidate passes it loads its
    // rows, runs tier 1 (block) and, on a tier-1 hit, the slow path.  V is
    // carried from block to block while they run and recomputed from the
    // block's first row after a skipped run.
    auto run_tier0 = [&](auto mask_tag) __attribute__((always_inline)) {
        using UFc = std::integral_constant<int, UF>;
        using UF1c = std::integral_constant<int, UF + 1>;
        const int nmain = max(t_end, 0) / UF;  // whole blocks
        int tv = 0;                    // V holds row a0+tv
        auto process = [&](int q, const uint4 rp, const uint32_t wb) __attribute__((always_inline)) {
            int m = 0x7FFFFFFF;
            const uint32_t rs[UF] = {rp.x, rp.y, rp.z, rp.w};
            const uint32_t c = c0 - wb;
#pragma unroll
            for (int s = 0; s < UF; ++s) {
#pragma unroll
                for (int k = 0; k < K; ++k) m = min(m, (int)__builtin_amdgcn_sad_u8(cpk[k], rs[s], c));
            }
            if (!__ballot(m <= t0h)) return;
            const int t = q * UF;
            int r[UF + 1][3];
            load_rows(t - 1, r, UF1c{});  // rows t .. t+UF
            if (t != tv) {
#pragma unroll
                for (int k = 0; k < K; ++k)
                    V[k] = sad((uint32_t)r[0][1], cy[k], sad((uint32_t)r[0][0], cx[k], o[k]));
            }
            if (block(t, r + 1, UFc{}, mask_tag)) slow(t, UF);
            tv = t + UF;
        };
        // (no software prefetch: a second buffer of LDS values spills the
        // column registers; other waves on the SIMD cover the LDS latency)
        if constexpr (PRUNE) {
            // live blocks only
            unsigned long long todo = live & (nmain >= 64 ? ~0ull : (1ull << nmain) - 1);
            while (todo) {
                const int q = __builtin_ctzll(todo);
                todo &= todo - 1;
                process(q, t0blk[2 * q], t0blk[2 * q + 1].x);
            }
        } else {
            for (int q = 0; q < nmain; ++q) process(q, t0blk[2 * q], t0blk[2 * q + 1].x);
        }
        // the rows after the last whole block (PRUNE: if the bound left any):
        // tier 1 directly
        int t = nmain * UF;
        if (t < t_end && (!PRUNE || ((live >> nmain) & 1))) {
            if (t != tv) {
                const uint32_t x = (uint32_t)prow[4 * t], y = (uint32_t)prow[4 * t + 1];
#pragma unroll
                for (int k = 0; k < K; ++k) V[k] = sad(y, cy[k], sad(x, cx[k], o[k]));
            }
            for (; t < t_end; ++t) {
                int r1[1][3];
                load_rows(t, r1, std::integral_constant<int, 1>{});
                if (block(t, r1, std::integral_constant<int, 1>{}, mask_tag)) slow(t, 1);
            }
        }
    };
    // PRUNE: live blocks only.  Each buffer holds a block's first row and its
    // four step rows; the next live block's rows are requested before this
    // one computes.
    auto run_pruned = [&](auto mask_tag) __attribute__((always_inline)) {
        using UFc = std::integral_constant<int, UF>;
        using UF1c = std::integral_constant<int, UF + 1>;
        const int nmain = max(t_end, 0) / UF;  // whole blocks
        unsigned long long todo = live & (nmain >= 64 ? ~0ull : (1ull << nmain) - 1);
        int tv = 0;  // V holds row a0+tv
        auto reseat = [&](int t, int x, int y) __attribute__((always_inline)) {
            if (t != tv) {
#pragma unroll
                for (int k = 0; k < K; ++k) V[k] = sad((uint32_t)y, cy[k], sad((uint32_t)x, cx[k], o[k]));
            }
        };
        auto process = [&](int q, const int (*r)[3]) __attribute__((always_inline)) {
            const int t = q * UF;
            reseat(t, r[0][0], r[0][1]);
            if (block(t, r + 1, UFc{}, mask_tag)) slow(t, UF);
            tv = t + UF;
        };
        auto next = [&]() __attribute__((always_inline)) -> int {
            if (!todo) return -1;
            const int q = __builtin_ctzll(todo);
            todo &= todo - 1;
            return q;
        };
        int ra[UF + 1][3], rb[UF + 1][3];
        int qa = next();
        if (qa >= 0) load_rows(qa * UF - 1, ra, UF1c{});  // rows 4qa .. 4qa+4
        while (qa >= 0) {
            const int qb = next();
            if (qb >= 0) load_rows(qb * UF - 1, rb, UF1c{});
            process(qa, ra);
            if (qb < 0) break;
            qa = next();
            if (qa >= 0) load_rows(qa * UF - 1, ra, UF1c{});
            process(qb, rb);
        }
        // the rows after the last whole block, if the bound left any
        int t = nmain * UF;
        if (t < t_end && ((live >> nmain) & 1)) {
            reseat(t, prow[4 * t], prow[4 * t + 1]);
            for (; t < t_end; ++t) {
                int r1[1][3];
                load_rows(t, r1, std::integral_constant<int, 1>{});
                if (block(t, r1, std::integral_constant<int, 1>{}, mask_tag)) slow(t, 1);
            }
        }
    };
    if constexpr (TIER0) {
        if (diag) run_tier0(std::true_type{});
        else run_tier0(std::false_type{});
    } else if constexpr (PRUNE) {
        if (diag) run_pruned(std::true_type{});
        else run_pruned(std::false_type{});
    } else {
        if (diag) run(std::true_type{});
        else run(std::false_type{});
    }

    // every improvement was published by the slow path
    const unsigned wfb = wave_sum_u32(fb);
    if (lane == 0 && wfb) atomicAdd(&in->fallbacks, (unsigned long long)wfb);
}

template <int K, int T, int WAVES, int MODE>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(K >= L2O_WPE6_MIN_K ? 6 : 1))) void k_sweep(
    const l2o_sweep_item *__restrict__ items, l2o_inst *__restrict__ insts)
{
    sweep_item<K, T, WAVES, MODE>(items[blockIdx.x], insts);
}

// ---------------------------------------------------------------------------
// apply: rotate+reverse gather of points, edge lengths and city ids
// ---------------------------------------------------------------------------
// twoOptLocalSolver.cpp:41-48 gives r' = r[i..J-1] ++ r[i-1..0] ++ r[n-1..J].
// With L = J-i:  r'[k] = r[i+k] (k < L), r[i-1-(k-L)] (L <= k < L+i),
// r[n-1+L+i-k] (k >= L+i).  Edge lengths move with their edges; only
// D'[0] = d(P[i],P[J]) and D'[L] = d(P[J-1],P[i-1]) are new.
__global__ __launch_bounds__(L2O_APPLY_THREADS) void k_apply(const l2o_apply_item *__restrict__ items,
                                                             l2o_inst *__restrict__ insts,
                                                             unsigned *__restrict__ live)
{
    const l2o_apply_item it = items[blockIdx.x];
    l2o_inst *in = insts + it.inst;
    if (in->phase != L2O_LO) return;
    const unsigned long long key = in->key;
    const uint32_t n = in->n;
    if (key == L2O_SENTINEL) {
        // no improving move: the local search has converged (sweeps = moves + 1)
        if (it.chunk == 0 && threadIdx.x == 0) {
            in->sweeps += 1;
            if (in->ils_rounds == 0) {
                in->phase = L2O_DONE;
                atomicSub(live, 1u);
            } else {
                in->phase = L2O_LO_DONE;
            }
        }
        return;
    }
    const int gain = key_gain(key);
    const uint32_t i = (uint32_t)(key >> 20) & 0xFFFFFu;
    const uint32_t J = (uint32_t)key & 0xFFFFFu;
    const uint32_t L = J - i;
    const uint32_t par = in->parity;
    const uint32_t metric = in->metric;
    const gpt *__restrict__ P = (const gpt *)in->pts[par];
    const gu32 *__restrict__ R = (const gu32 *)in->tour[par];
    gpt *__restrict__ Pn = (gpt *)in->pts[par ^ 1];
    gu32 *__restrict__ Rn = (gu32 *)in->tour[par ^ 1];
    const uint32_t k0 = it.chunk * L2O_APPLY_CHUNK;
    const uint32_t k1 = min(n, k0 + L2O_APPLY_CHUNK);
    // A thread's CHUNK/THREADS elements are independent: their sources are
    // computed first and all their loads issued before the first store.
    constexpr int PER = L2O_APPLY_CHUNK / L2O_APPLY_THREADS;
    uint32_t src[PER], de[PER];  // source point; position whose edge (d, w) lands here
    bool ok[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const uint32_t k = k0 + threadIdx.x + e * L2O_APPLY_THREADS;
        ok[e] = k < k1;
        uint32_t sr, dr;
        if (k < L) {
            sr = i + k;
            dr = sr;
        } else if (k == L) {
            sr = i - 1;
            dr = sr;
        } else if (k < L + i) {
            sr = i - 1 - (k - L);
            dr = sr + 1;
        } else {
            sr = n - 1 + L + i - k;
            dr = sr + 1 == n ? 0 : sr + 1;
        }
        src[e] = ok[e] ? sr : 0;
        de[e] = ok[e] ? dr : 0;
    }
    l2o_pt p[PER], q[PER];
    uint32_t rid[PER];
    // the far ends of the two new edges, loaded with the rest (one round trip)
    const l2o_pt endJ = ld_pt(P, J), endJ1 = ld_pt(P, J - 1);
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        p[e] = ld_pt(P, src[e]);
        q[e] = ld_pt(P, de[e]);
        rid[e] = R[src[e]];
    }
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const uint32_t k = k0 + threadIdx.x + e * L2O_APPLY_THREADS;
        if (!ok[e]) continue;
        l2o_pt r = p[e];
        if (k == 0 || k == L) {
            // the two new edges: P[i] -> P[J] at position 0, P[J-1] -> P[i-1] at L
            const l2o_pt o = k == 0 ? endJ : endJ1;
            r.d = exact_dist(r.x, r.y, o.x, o.y, metric);
            r.w = ceil_sqrtg(r.d, tier1_g(metric));
        } else {
            r.d = q[e].d;
            r.w = q[e].w;
        }
        st_pt(Pn, k, r);
        Rn[k] = rid[e];
    }
    // Every wave of the workgroup has read key, parity and its sources once
    // all are past the barrier (each loaded value feeds a store issued before
    // it), so the last chunk to count may commit the move.  No memory fence:
    // nothing in this kernel reads the new buffers, and the next kernel sees
    // them at the kernel boundary; the empty asm keeps the compiler from
    // moving memory operations across the barrier.
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (threadIdx.x == 0) {
        const uint32_t nch = (n + L2O_APPLY_CHUNK - 1) / L2O_APPLY_CHUNK;
        const uint32_t prev = atomicAdd(&in->apply_done, 1u);
        if (prev == nch - 1) {  // last chunk of this instance: commit the move
            in->apply_done = 0;
            in->length += (unsigned long long)(long long)gain;
            if (in->moves < in->log_cap) {
                l2o_move m;
                m.i = i;
                m.j = J;
                m.gain = gain;
                in->log[in->moves] = m;
            }
            in->moves += 1;
            in->sweeps += 1;
            in->parity = par ^ 1;
            in->key = L2O_SENTINEL;
        }
    }
}

// ---------------------------------------------------------------------------
// ILS bookkeeping: accept/revert, kick, restart (ILSGlobalSolverSequential.cpp:73-99)
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long rotl64(unsigned long long v, int k)
{
    return (v << k) | (v >> (64 - k));
}

__device__ unsigned long long xoshiro_next(unsigned long long *s)
{
    const unsigned long long result = rotl64(s[1] * 5ull, 7) * 9ull;
    const unsigned long long t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

// A plain ILS instance at L2O_LO_DONE: accept iff strictly shorter (copy to
// the best buffers), stop after `rounds` or once the best length is below
// stop_below, else kick its best and restart.
// A chain of a multi-start group at L2O_LO_DONE only accepts into its own best
// and waits (L2O_WAIT) for k_group; at L2O_KICK it kicks the best tour of
// `kick_src` that k_group chose.  While any chain of a group is at L2O_KICK,
// all of them are, so no best buffer is written while a kick reads it.
__global__ __launch_bounds__(L2O_ILS_THREADS) void k_ils(l2o_inst *__restrict__ insts, unsigned *__restrict__ live)
{
    l2o_inst *in = insts + blockIdx.x;
    const uint32_t phase = in->phase;
    if (phase != L2O_LO_DONE && phase != L2O_KICK) return;
    const bool grouped = in->group_lead != L2O_NO_GROUP;
    __shared__ int s_accept, s_done;
    __shared__ uint32_t s_a, s_b, s_c;
    __shared__ unsigned long long s_sum[L2O_ILS_THREADS / 64];
    const uint32_t n = in->n;
    const uint32_t par = in->parity;
    if (threadIdx.x == 0) {
        int acc = 0, done = 0;
        if (phase == L2O_LO_DONE) {
            acc = !in->have_best || in->length < in->best_length;
            if (acc) {
                in->best_length = in->length;
                in->have_best = 1;
            }
            if (grouped) {
                in->improved = acc;
            } else {
                if (in->rlog) in->rlog[in->rounds_done] = in->best_length;
                done = in->rounds_done >= in->ils_rounds || in->best_length < in->stop_below;
            }
        }
        s_accept = acc;
        s_done = done;
        if (!done && (phase == L2O_KICK || !grouped)) {
            unsigned long long *rng = in->rng;
            if (in->kick == 1) {  // reference kick: a, b = rand() % n, sorted (ILSGlobalSolver.cpp:39-44)
                uint32_t a = (uint32_t)(xoshiro_next(rng) % n);
                uint32_t b = (uint32_t)(xoshiro_next(rng) % n);
                if (b < a) {
                    const uint32_t tmp = a;
                    a = b;
                    b = tmp;
                }
                s_a = a;
                s_b = b;
                s_c = 0;
            } else {  // double-bridge cut points 1 <= p1 < p2 < p3 <= n-1
                uint32_t p[3];
                int have = 0;
                while (have < 3) {
                    const uint32_t v = 1u + (uint32_t)(xoshiro_next(rng) % (unsigned long long)(n - 1));
                    bool dup = false;
                    for (int q = 0; q < have; ++q) dup |= (p[q] == v);
                    if (!dup) p[have++] = v;
                }
                if (p[0] > p[1]) { const uint32_t tmp = p[0]; p[0] = p[1]; p[1] = tmp; }
                if (p[1] > p[2]) { const uint32_t tmp = p[1]; p[1] = p[2]; p[2] = tmp; }
                if (p[0] > p[1]) { const uint32_t tmp = p[0]; p[0] = p[1]; p[1] = tmp; }
                s_a = p[0];
                s_b = p[1];
                s_c = p[2];
            }
        }
    }
    __syncthreads();
    if (s_accept) {
        const l2o_pt *__restrict__ P = in->pts[par];
        const uint32_t *__restrict__ R = in->tour[par];
        l2o_pt *__restrict__ BP = in->bpts;
        uint32_t *__restrict__ BR = in->btour;
        for (uint32_t k = threadIdx.x; k < n; k += L2O_ILS_THREADS) {
            BP[k] = P[k];
            BR[k] = R[k];
        }
    }
    __syncthreads();
    if (grouped && phase == L2O_LO_DONE) {
        if (threadIdx.x == 0) {
            __threadfence();
            in->phase = L2O_WAIT;
        }
        return;
    }
    if (s_done) {
        if (threadIdx.x == 0) {
            in->phase = L2O_DONE;
            atomicSub(live, 1u);
        }
        return;
    }
    // new current tour = kick(best of the source), written to buffer 0
    const l2o_inst *src = grouped ? insts + in->kick_src : in;
    const l2o_pt *__restrict__ BP = src->bpts;
    const uint32_t *__restrict__ BR = src->btour;
    const uint32_t a = s_a, b = s_b, c = s_c;
    const bool ref = in->kick == 1;
    auto src_of = [&](uint32_t k) -> uint32_t {
        if (ref) {  // rotate(a) then reverse from b-a (ILSGlobalSolver.cpp:46-47)
            return k < b - a ? (a + k) % n : (n - 1 - k + b) % n;
        }
        // A C B D: A=[0,a) B=[a,b) C=[b,c) D=[c,n)
        if (k < a) return k;
        if (k < a + (c - b)) return b + (k - a);
        if (k < c) return a + (k - a - (c - b));
        return k;
    };
    l2o_pt *__restrict__ Pn = in->pts[0];
    uint32_t *__restrict__ Rn = in->tour[0];
    const uint32_t metric = in->metric;
    const int g = tier1_g(metric);
    unsigned long long sum = 0;
    for (uint32_t k = threadIdx.x; k < n; k += L2O_ILS_THREADS) {
        const uint32_t s = src_of(k);
        const uint32_t sp = src_of(k == 0 ? n - 1 : k - 1);
        const l2o_pt p = BP[s], q = BP[sp];
        const int d = exact_dist(q.x, q.y, p.x, p.y, metric);
        sum += (unsigned long long)d;
        l2o_pt r;
        r.x = p.x;
        r.y = p.y;
        r.d = d;
        r.w = ceil_sqrtg(d, g);
        Pn[k] = r;
        Rn[k] = BR[s];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned lo = __shfl_xor((unsigned)sum, off);
        const unsigned hi = __shfl_xor((unsigned)(sum >> 32), off);
        sum += ((unsigned long long)hi << 32) | lo;
    }
    if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long L = 0;
        for (int k = 0; k < L2O_ILS_THREADS / 64; ++k) L += s_sum[k];
        in->length = L;
        in->parity = 0;
        in->rounds_done += 1;
        in->key = L2O_SENTINEL;
        in->apply_done = 0;
        __threadfence();
        in->phase = L2O_LO;
    }
}

// ---------------------------------------------------------------------------
// Multi-start ILS round barrier: ILSGlobalSolverMT::optimize with comm = 1
// (ILSGlobalSolverMT.cpp:86-195) shares a global best between threads under a
// mutex, so its result depends on timing.  Here the chains of a group advance
// round by round: once every chain has reached its local optimum (L2O_WAIT),
// one thread per group takes the group best (smallest best length, lowest
// chain on ties), logs it, and sends each chain on with a kick of its own
// tour if that improved its best this round, else of the group best (the
// reference's "route = bestGlobalMinima.back()"); after `rounds` kicks every
// chain stops.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_group(l2o_inst *__restrict__ insts, uint32_t nslots,
                                              unsigned *__restrict__ live)
{
    const uint32_t id = blockIdx.x * 64 + threadIdx.x;
    if (id >= nslots) return;
    l2o_inst *lead = insts + id;
    if (lead->group_lead != id || lead->phase != L2O_WAIT) return;
    const uint32_t m = lead->group_size;
    uint32_t best = id;
    for (uint32_t c = 0; c < m; ++c) {
        const l2o_inst *in = insts + id + c;
        if (in->phase != L2O_WAIT) return;  // not all chains are done with this round
        if (in->best_length < insts[best].best_length) best = id + c;
    }
    const unsigned long long G = insts[best].best_length;
    const uint32_t rounds = lead->rounds_done;
    const bool done = rounds >= lead->ils_rounds || G < lead->stop_below;
    for (uint32_t c = 0; c < m; ++c) {
        l2o_inst *in = insts + id + c;
        if (in->rlog) in->rlog[rounds] = G;
        in->kick_src = in->improved ? id + c : best;
        in->phase = done ? L2O_DONE : L2O_KICK;
    }
    if (done) atomicSub(live, m);
}

// ---------------------------------------------------------------------------
// exact tour length of the current buffer: sum of D (headers.h:170-189)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_length(const l2o_inst *__restrict__ insts, uint32_t slot,
                                                unsigned long long *__restrict__ out)
{
    const l2o_inst *in = insts + slot;
    const l2o_pt *P = in->pts[in->parity];
    const uint32_t n = in->n;
    unsigned long long sum = 0;
    for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
        const l2o_pt p = P[k], q = P[k == 0 ? n - 1 : k - 1];
        sum += (unsigned long long)exact_dist(q.x, q.y, p.x, p.y, in->metric);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned lo = __shfl_xor((unsigned)sum, off);
        const unsigned hi = __shfl_xor((unsigned)(sum >> 32), off);
        sum += ((unsigned long long)hi << 32) | lo;
    }
    if ((threadIdx.x & 63) == 0 && sum) atomicAdd(out, sum);
}

// ---------------------------------------------------------------------------
// v_sqrt_f32 accuracy: max |ulp| distance from the correctly rounded sqrt
// over float bit patterns [lo, hi)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sqrt_ulp(uint32_t lo, uint32_t hi, unsigned *__restrict__ max_ulp)
{
    unsigned m = 0;
    for (unsigned long long b = (unsigned long long)lo + blockIdx.x * 256ull + threadIdx.x; b < hi;
         b += (unsigned long long)gridDim.x * 256ull) {
        const float x = __uint_as_float((uint32_t)b);
        const float hw = __builtin_amdgcn_sqrtf(x);
        const float cr = (float)sqrt((double)x);  // correctly rounded (sqrt is double-rounding safe)
        const int d = (int)__float_as_uint(hw) - (int)__float_as_uint(cr);
        const unsigned ad = (unsigned)(d < 0 ? -d : d);
        m = ad > m ? ad : m;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned o = __shfl_xor(m, off);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(max_ulp, m);
}

}  // namespace l2o

// ---------------------------------------------------------------------------
// launch wrappers used by the host engine
// ---------------------------------------------------------------------------
#include "l2o_launch.h"

namespace l2o {

// tile shapes compiled into the library (k diagonals per lane, t steps,
// waves per workgroup); the autotuner chooses among these per n bucket
#define L2O_SWEEP_SHAPES(X)                                                                          \
    X(1, 32, 1) X(1, 64, 1) X(2, 64, 1) X(2, 128, 1) X(4, 64, 1) X(4, 128, 1) X(4, 256, 1)           \
    X(6, 64, 1) X(6, 128, 1) X(8, 64, 1) X(8, 128, 1) X(8, 256, 1) X(4, 64, 2) X(4, 128, 2)           \
    X(6, 64, 2) X(8, 64, 2) X(8, 128, 2) X(8, 256, 2) X(2, 64, 4) X(4, 64, 4) X(6, 64, 4)             \
    X(8, 64, 4) X(8, 128, 4) X(6, 128, 2) X(6, 160, 1) X(6, 128, 4) X(4, 256, 2) X(4, 256, 4)

#define L2O_SWEEP_CASE(K_, T_, W_)                                                                   \
    if (cfg.k == K_ && cfg.t == T_ && cfg.waves == W_) {                                            \
        if (mode == L2O_SWEEP_PRUNE)                                                                 \
            hipLaunchKernelGGL((k_sweep<K_, T_, W_, L2O_SWEEP_PRUNE>), dim3(nitems), dim3(64 * W_), 0, stream, items, insts); \
        else if (mode == L2O_SWEEP_PRUNE_T1)                                                         \
            hipLaunchKernelGGL((k_sweep<K_, T_, W_, L2O_SWEEP_PRUNE_T1>), dim3(nitems), dim3(64 * W_), 0, stream, items, insts); \
        else                                                                                         \
            hipLaunchKernelGGL((k_sweep<K_, T_, W_, L2O_SWEEP_FULL>), dim3(nitems), dim3(64 * W_), 0, stream, items, insts); \
        return hipGetLastError();                                                                    \
    }

hipError_t launch_sweep(const sweep_cfg &cfg, const l2o_sweep_item *items, uint32_t nitems,
                        l2o_inst *insts, hipStream_t stream, int mode)
{
    if (nitems == 0) return hipSuccess;
    if (mode < L2O_SWEEP_FULL || mode > L2O_SWEEP_PRUNE_T1) return hipErrorInvalidValue;
    L2O_SWEEP_SHAPES(L2O_SWEEP_CASE)
    return hipErrorInvalidValue;
}

#define L2O_SWEEP_ROW(K_, T_, W_) {K_, T_, W_},
static const int g_shapes[][3] = {L2O_SWEEP_SHAPES(L2O_SWEEP_ROW)};

bool sweep_cfg_supported(const sweep_cfg &c)
{
    for (const auto &r : g_shapes)
        if (r[0] == c.k && r[1] == c.t && r[2] == c.waves) return true;
    return false;
}

int sweep_shapes(int (*out)[3], int cap)
{
    const int m = (int)(sizeof(g_shapes) / sizeof(g_shapes[0]));
    for (int q = 0; q < m && q < cap; ++q) {
        out[q][0] = g_shapes[q][0];
        out[q][1] = g_shapes[q][1];
        out[q][2] = g_shapes[q][2];
    }
    return m;
}

hipError_t launch_apply(const l2o_apply_item *items, uint32_t nitems, l2o_inst *insts, unsigned *live,
                        hipStream_t stream)
{
    if (nitems == 0) return hipSuccess;
    hipLaunchKernelGGL(k_apply, dim3(nitems), dim3(L2O_APPLY_THREADS), 0, stream, items, insts, live);
    return hipGetLastError();
}

hipError_t launch_ils(l2o_inst *insts, uint32_t nslots, unsigned *live, hipStream_t stream)
{
    if (nslots == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ils, dim3(nslots), dim3(L2O_ILS_THREADS), 0, stream, insts, live);
    return hipGetLastError();
}

hipError_t launch_group(l2o_inst *insts, uint32_t nslots, unsigned *live, hipStream_t stream)
{
    if (nslots == 0) return hipSuccess;
    hipLaunchKernelGGL(k_group, dim3((nslots + 63) / 64), dim3(64), 0, stream, insts, nslots, live);
    return hipGetLastError();
}

hipError_t launch_length(const l2o_inst *insts, uint32_t slot, uint32_t n, unsigned long long *out,
                         hipStream_t stream)
{
    const uint32_t blocks = min(1024u, (n + 255) / 256);
    hipLaunchKernelGGL(k_length, dim3(blocks), dim3(256), 0, stream, insts, slot, out);
    return hipGetLastError();
}

hipError_t launch_sqrt_ulp(uint32_t lo, uint32_t hi, unsigned *max_ulp, hipStream_t stream)
{
    hipLaunchKernelGGL(k_sqrt_ulp, dim3(4096), dim3(256), 0, stream, lo, hi, max_ulp);
    return hipGetLastError();
}

}  // namespace l2o
