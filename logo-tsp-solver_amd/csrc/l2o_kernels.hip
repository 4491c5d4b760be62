// l2o_kernels.hip -- gfx950 kernels of the 2-opt local-search path.
//
// Replaces the reference's kernel2opt / kernel2opt_extended
// (src/cuda/opt_kernel.cu:34-262, 273-627) and the host-side apply
// (src/solvers/twoOptLocalSolver.cpp:40-50).  Wave64, no MFMA: the sweep is
// FP32 VALU + transcendental work.
//
// Sweep geometry.  With a = i-1 and b = J-1 a candidate is a pair
// 0 <= a, a+2 <= b <= n-2 with gain
//     M(a,b) + M(a+1,b+1) - D[a+1] - D[b+1],   M(x,y) = d(P[x],P[y]),
// D[k] = d(P[k-1],P[k])  (cpuSolver.cpp:133-138 in tour-position form).
// A lane walks diagonals g = b-a: the distance M(a+1,b+1) computed at step a
// is the first term of step a+1, so each candidate costs ONE new distance
// (one v_sqrt_f32) instead of two.  A workgroup owns DIAG consecutive
// diagonals and T consecutive steps; the ordered points it touches are staged
// once in LDS (rows: uniform per step, a broadcast read; columns: SoA words
// read per lane without bank conflicts).
//
// Exactness.  The FP32 estimate g~ = q(a,b) + q(a+1,b+1) - D[a+1] - D[b+1]
// differs from the exact integer gain by at most `slack` (per instance, set by
// the host from the coordinate span: 1 for the two nint roundings plus the
// FP32 error of s, of v_sqrt_f32 and of the three adds).  A wave keeps the
// best exact key it has seen; a candidate is re-evaluated exactly (FP64 sqrt
// + integer fix-up) only if g~ <= best_gain - 1 + slack, i.e. only if it
// could still beat that key.  Candidates of later steps have larger i, so a
// tie with the current best can never win; candidates of the same step are
// all re-evaluated together.  The wave's result is therefore the exact
// lexicographic minimum of (gain, i, J) over its candidates, and the global
// atomicMin over 64-bit keys makes the grid result exact and deterministic
// (the reference's grid tie-break is a racy lock, opt_kernel.cu:239-261).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

#include "l2o_state.h"

namespace l2o {

// ---------------------------------------------------------------------------
// exact TSPLIB EUC_2D distance: nint(sqrt(s)) = k + [s > k^2 + k], k = isqrt(s)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int exact_dist(float x1, float y1, float x2, float y2)
{
    // coordinates are integers below 2^22: int32 conversion is exact
    const long long dx = (long long)((int)x1 - (int)x2);
    const long long dy = (long long)((int)y1 - (int)y2);
    const long long s = dx * dx + dy * dy;
    long long k = (long long)sqrt((double)s);
    if (k * k > s) k -= 1;
    else if ((k + 1) * (k + 1) <= s) k += 1;
    return (int)(k + (s > k * k + k ? 1 : 0));
}

// FP32 estimate of d(P1,P2): the second filter tier (error bounded by slack)
__device__ __forceinline__ float est_dist(float x1, float y1, float x2, float y2)
{
    const float dx = x1 - x2, dy = y1 - y2;
    return __builtin_amdgcn_sqrtf(__builtin_fmaf(dx, dx, dy * dy));
}

// Chebyshev lower bound max(|dx|, |dy|) <= d(P1,P2): exact in FP32 for integer
// coordinates below 2^22 (the differences are exact), and no square root
__device__ __forceinline__ float cheb_dist(float x1, float y1, float x2, float y2)
{
    return __builtin_fmaxf(__builtin_fabsf(x1 - x2), __builtin_fabsf(y1 - y2));
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

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned lo = __shfl_xor((unsigned)v, off);
        const unsigned hi = __shfl_xor((unsigned)(v >> 32), off);
        const unsigned long long o = ((unsigned long long)hi << 32) | lo;
        v = o < v ? o : v;
    }
    return v;
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
// LDS layout of one workgroup (all float):
//   rx, ry, rd [T+1]   rows, index 0 = P[a0], index t+1 = P[a0+t+1] (SoA so
//                      that four consecutive steps are one broadcast read)
//   cx, cy, cd [T+DIAG] columns P[a0+g0+c], SoA: a lane reads 4-byte words at
//                      a K-word stride, conflict free (an AoS float4 array read
//                      as ds_read_b96 at a 16-byte stride measured 65% bank-
//                      conflict cycles and kept LDS busy 86% of the time)
//   u[WAVES][T]        per-wave thr + D[a0+t+1]: the whole per-step
//                      threshold comes from LDS, so a candidate costs
//                      2 sub + mul + fma + sqrt + add + sub + cmp on the VALU
//   With K = 2 a lane's words sit 2 apart, so lanes l and l+16 would share a
//   bank: the columns are stored twice, the second copy an odd number of words
//   later, and lane l reads copy bit4(l) ^ bit5(l).  Every 32 consecutive lanes
//   then hit 32 distinct banks of 32, and all 64 lanes 64 distinct banks of 64.
template <int K, int T, int WAVES>
struct sweep_smem {
    static constexpr int NC = T + 64 * WAVES * K + 4;  // columns per copy
    static constexpr int CP = K == 2 ? 2 : 1;         // copies
    static constexpr int SC = NC + 2;                 // copy stride, even
    float rx[T + 4], ry[T + 4], rd[T + 4];
    float cx[CP * SC], cy[CP * SC], cd[CP * SC];
    float u[WAVES][T];
    unsigned long long key[WAVES];
    unsigned fb[WAVES];
};

// one work item: all candidates of diagonals [g0, g0+DIAG) x steps [a0, a0+T)
template <int K, int T, int WAVES>
__device__ __forceinline__ void sweep_item(const l2o_sweep_item it, l2o_inst *__restrict__ insts,
                                           sweep_smem<K, T, WAVES> &sm)
{
    constexpr int NT = 64 * WAVES;
    constexpr int DIAG = NT * K;
    constexpr int NCOL = T + DIAG;
    constexpr int UF = 4;  // steps per unrolled iteration
    static_assert(T % UF == 0, "T must be a multiple of the unroll factor");

    l2o_inst *in = insts + it.inst;
    if (in->phase != L2O_LO) return;
    const int n = (int)in->n;
    const float4 *__restrict__ P = in->pts[in->parity];
    const int a0 = (int)it.a0, g0 = (int)it.g0;
    const float slack = in->slack;

    using SM = sweep_smem<K, T, WAVES>;
    for (int c = threadIdx.x; c < NCOL; c += NT) {
        const float4 p = P[min(a0 + g0 + c, n - 1)];
#pragma unroll
        for (int q = 0; q < SM::CP; ++q) {
            sm.cx[q * SM::SC + q + c] = p.x;
            sm.cy[q * SM::SC + q + c] = p.y;
            sm.cd[q * SM::SC + q + c] = p.z;
        }
    }
    // this lane's copy of the columns (see sweep_smem)
    const int coff = SM::CP == 2 ? ((((threadIdx.x >> 4) ^ (threadIdx.x >> 5)) & 1) * (SM::SC + 1)) : 0;
    const float *__restrict__ ccx = sm.cx + coff;
    const float *__restrict__ ccy = sm.cy + coff;
    const float *__restrict__ ccd = sm.cd + coff;
    auto col = [&](int c) { return make_float4(ccx[c], ccy[c], ccd[c], 0.0f); };
    for (int r = threadIdx.x; r <= T; r += NT) {
        const float4 p = P[min(a0 + r, n - 1)];
        sm.rx[r] = p.x;
        sm.ry[r] = p.y;
        sm.rd[r] = p.z;
    }
    __syncthreads();

    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int cb = w * 64 * K + lane * K;  // this lane's first diagonal, relative to g0
    const int gmin = g0 + w * 64 * K, gmax = gmin + 64 * K - 1;
    // candidate (a, a+g) exists iff a + g <= n-2, i.e. step t <= n-2-g-a0
    const int t_full = max(0, min(T, n - 1 - gmax - a0));  // every lane valid below
    const int t_end = max(0, min(T, n - 1 - gmin - a0));   // no lane valid from here
    float *__restrict__ su = sm.u[w];

    float thr = slack - 1.0f;  // (best_gain - 1) + slack with best_gain = 0: only gain <= -1 matters
    for (int t = lane; t < T; t += 64) su[t] = thr + sm.rd[t + 1];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    // Two filter tiers.  Tier 1 (every candidate) bounds the gain from below
    // with Chebyshev distances: nint(d) >= d - 1/2 >= cheb - 1/2, so an exact
    // gain <= best-1 implies cheb(a,b) + cheb(a+1,b+1) - D[a+1] - D[b+1] <=
    // best; that sum is an exact FP32 integer and the test against
    // thr + D[a+1] (thr = best-1+slack) keeps slack-1 of margin for the FP32
    // rounding of the threshold.  Tier 2 (lanes flagged by tier 1) is the FP32
    // distance estimate, |exact - estimate| <= slack.  Both tiers only drop
    // candidates that cannot reach best-1, so the exact pass sees every move
    // that can win.  A wave ballot of tier 1 is rarely nonzero: the Chebyshev
    // sum exceeds the two tour edges unless both pairs are spatial neighbours.
    float qp[K];  // tier-1 bound of the previous step, per diagonal
    {
        const float x0 = sm.rx[0], y0 = sm.ry[0];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const float4 C = col(cb + k);
            qp[k] = cheb_dist(x0, y0, C.x, C.y);
        }
    }
    unsigned long long lbest = L2O_SENTINEL;
    unsigned fb = 0;

    // tier 2 and exact re-evaluation of tier-1 hits (steps t0 .. t0+ns-1);
    // then the wave's threshold and its per-step thresholds in LDS are
    // refreshed.  Hit masks are wave ballots (SGPR pairs): lane bits are only
    // extracted here.
    auto slow = [&](int t0, int ns, const unsigned long long *hit) {
        for (int s = 0; s < ns; ++s) {
            const int t = t0 + s;
            const float xa = sm.rx[t], ya = sm.ry[t], xb = sm.rx[t + 1], yb = sm.ry[t + 1], db = sm.rd[t + 1];
            const float u = su[t];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if ((hit[s * K + k] >> lane) & 1ull) {
                    const float4 Ca = col(t + cb + k), Cb = col(t + 1 + cb + k);
                    const float w2 = (est_dist(xa, ya, Ca.x, Ca.y) + est_dist(xb, yb, Cb.x, Cb.y)) - Cb.z;
                    if (w2 <= u) {
                        ++fb;
                        const int gain = exact_dist(xa, ya, Ca.x, Ca.y) + exact_dist(xb, yb, Cb.x, Cb.y) -
                                         (int)db - (int)Cb.z;
                        if (gain < 0) {
                            const uint32_t i = (uint32_t)(a0 + t + 1);
                            const uint32_t j = (uint32_t)(a0 + t + 1 + g0 + cb + k);
                            const unsigned long long key = make_key(gain, i, j);
                            lbest = key < lbest ? key : lbest;
                        }
                    }
                }
            }
        }
        const unsigned long long wb = wave_min_u64(lbest);
        if (wb != L2O_SENTINEL) {
            const float nthr = (float)(key_gain(wb) - 1) + slack;
            if (nthr != thr) {
                thr = nthr;
                __builtin_amdgcn_wave_barrier();
                for (int t = lane; t < T; t += 64) su[t] = thr + sm.rd[t + 1];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
    };

    // Main loop.  Row points are wave-uniform: they come from global memory
    // through the scalar cache (below t_full every row index a0+t+1+s is
    // < n).  Column words are read through volatile pointers so that each is
    // one ds_read_b32 off a single per-lane base with an immediate offset:
    // merged ds_read2 pairs would need a fresh base VGPR per array (their
    // 8-bit offsets cannot span the column arrays), one VALU add each.
    const __attribute__((address_space(4))) float *__restrict__ prow =
        (const __attribute__((address_space(4))) float *)(P + a0 + 1);  // row t at prow[4t]
    typedef const volatile __attribute__((address_space(3))) float lds_vf;
    lds_vf *vcx = (lds_vf *)(ccx + cb + 1), *vcy = (lds_vf *)(ccy + cb + 1), *vcd = (lds_vf *)(ccd + cb + 1);
    int t = 0;
    const int t_main = t_full - t_full % UF;
    for (; t < t_main; t += UF) {
        float x[UF], y[UF], U[UF];
#pragma unroll
        for (int s = 0; s < UF; ++s) {
            x[s] = prow[4 * (t + s)];
            y[s] = prow[4 * (t + s) + 1];
            U[s] = su[t + s];
        }
        float4 C[UF + K - 1];
#pragma unroll
        for (int j = 0; j < UF + K - 1; ++j) C[j] = make_float4(vcx[t + j], vcy[t + j], vcd[t + j], 0.0f);
        unsigned long long hit[UF * K];
        unsigned long long any = 0;
#pragma unroll
        for (int s = 0; s < UF; ++s) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const float q = cheb_dist(x[s], y[s], C[s + k].x, C[s + k].y);
                const float w2 = (qp[k] + q) - C[s + k].z;
                qp[k] = q;
                hit[s * K + k] = __ballot(w2 <= U[s]);
                any |= hit[s * K + k];
            }
        }
        if (any) slow(t, UF, hit);
    }
    int lim[K];
#pragma unroll
    for (int k = 0; k < K; ++k) lim[k] = n - 2 - a0 - (g0 + cb + k);
    for (; t < t_end; ++t) {
        const float x = sm.rx[t + 1], y = sm.ry[t + 1], U = su[t];
        unsigned long long hit[K];
        unsigned long long any = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const float4 C = col(t + 1 + cb + k);
            const float q = cheb_dist(x, y, C.x, C.y);
            const float w2 = (qp[k] + q) - C.z;
            qp[k] = q;
            hit[k] = __ballot((w2 <= U) && (t <= lim[k]));
            any |= hit[k];
        }
        if (any) slow(t, 1, hit);
    }

    const unsigned long long wb = wave_min_u64(lbest);
    const unsigned wfb = wave_sum_u32(fb);
    if (lane == 0) {
        sm.key[w] = wb;
        sm.fb[w] = wfb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long key = sm.key[0];
        unsigned f = sm.fb[0];
#pragma unroll
        for (int k = 1; k < WAVES; ++k) {
            key = sm.key[k] < key ? sm.key[k] : key;
            f += sm.fb[k];
        }
        if (key != L2O_SENTINEL) {
            const unsigned long long cur = __hip_atomic_load(&in->key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (key < cur) atomicMin(&in->key, key);
        }
        if (f) atomicAdd(&in->fallbacks, (unsigned long long)f);
    }
}

// One workgroup per work item.  (A persistent variant that loops over items
// fetched from an atomic ticket is NOT used: with a lane-0 ticket fetch and
// lane-0 key atomics inside the item loop, the AMDGPU control-flow
// structurizer emits an exec-masked loop nest in which lanes 1..63 re-enter
// the item body on a stale ticket while lane 0 waits outside -- a livelock.
// Dispatch-round quantization is handled instead by short items, t = 64.)
template <int K, int T, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_sweep(const l2o_sweep_item *__restrict__ items,
                                                      l2o_inst *__restrict__ insts)
{
    __shared__ sweep_smem<K, T, WAVES> sm;
    sweep_item<K, T, WAVES>(items[blockIdx.x], insts, sm);
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
    const float4 *__restrict__ P = in->pts[par];
    const uint32_t *__restrict__ R = in->tour[par];
    float4 *__restrict__ Pn = in->pts[par ^ 1];
    uint32_t *__restrict__ Rn = in->tour[par ^ 1];
    const uint32_t k0 = it.chunk * L2O_APPLY_CHUNK;
    const uint32_t k1 = min(n, k0 + L2O_APPLY_CHUNK);
    for (uint32_t k = k0 + threadIdx.x; k < k1; k += L2O_APPLY_THREADS) {
        uint32_t src;
        float d;
        if (k < L) {
            src = i + k;
            if (k == 0) {
                const float4 a = P[i], b = P[J];
                d = (float)exact_dist(a.x, a.y, b.x, b.y);
            } else {
                d = P[src].z;
            }
        } else if (k == L) {
            src = i - 1;
            const float4 a = P[J - 1], b = P[i - 1];
            d = (float)exact_dist(a.x, a.y, b.x, b.y);
        } else if (k < L + i) {
            src = i - 1 - (k - L);
            d = P[src + 1].z;
        } else {
            src = n - 1 + L + i - k;
            d = P[src + 1 == n ? 0 : src + 1].z;
        }
        const float4 p = P[src];
        Pn[k] = make_float4(p.x, p.y, d, 0.0f);
        Rn[k] = R[src];
    }
    __threadfence();
    __syncthreads();
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
            __threadfence();
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

__global__ __launch_bounds__(L2O_ILS_THREADS) void k_ils(l2o_inst *__restrict__ insts, unsigned *__restrict__ live)
{
    l2o_inst *in = insts + blockIdx.x;
    if (in->phase != L2O_LO_DONE) return;
    __shared__ int s_accept, s_done;
    __shared__ uint32_t s_a, s_b, s_c;
    __shared__ unsigned long long s_sum[L2O_ILS_THREADS / 64];
    const uint32_t n = in->n;
    const uint32_t par = in->parity;
    if (threadIdx.x == 0) {
        const int acc = !in->have_best || in->length < in->best_length;
        s_accept = acc;
        if (acc) {
            in->best_length = in->length;
            in->have_best = 1;
        }
        if (in->rlog) in->rlog[in->rounds_done] = in->best_length;
        s_done = in->rounds_done >= in->ils_rounds;
        if (!s_done) {
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
    const float4 *__restrict__ P = in->pts[par];
    const uint32_t *__restrict__ R = in->tour[par];
    float4 *__restrict__ BP = in->bpts;
    uint32_t *__restrict__ BR = in->btour;
    if (s_accept) {
        for (uint32_t k = threadIdx.x; k < n; k += L2O_ILS_THREADS) {
            BP[k] = P[k];
            BR[k] = R[k];
        }
    }
    __syncthreads();
    if (s_done) {
        if (threadIdx.x == 0) {
            in->phase = L2O_DONE;
            atomicSub(live, 1u);
        }
        return;
    }
    // new current tour = kick(best), written to buffer 0
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
    float4 *__restrict__ Pn = in->pts[0];
    uint32_t *__restrict__ Rn = in->tour[0];
    unsigned long long sum = 0;
    for (uint32_t k = threadIdx.x; k < n; k += L2O_ILS_THREADS) {
        const uint32_t s = src_of(k);
        const uint32_t sp = src_of(k == 0 ? n - 1 : k - 1);
        const float4 p = BP[s], q = BP[sp];
        const int d = exact_dist(q.x, q.y, p.x, p.y);
        sum += (unsigned long long)d;
        Pn[k] = make_float4(p.x, p.y, (float)d, 0.0f);
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
// exact tour length of the current buffer: sum of D (headers.h:170-189)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_length(const l2o_inst *__restrict__ insts, uint32_t slot,
                                                unsigned long long *__restrict__ out)
{
    const l2o_inst *in = insts + slot;
    const float4 *P = in->pts[in->parity];
    const uint32_t n = in->n;
    unsigned long long sum = 0;
    for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
        const float4 p = P[k], q = P[k == 0 ? n - 1 : k - 1];
        sum += (unsigned long long)exact_dist(q.x, q.y, p.x, p.y);
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
    X(1, 64, 1) X(1, 128, 1) X(2, 64, 1) X(2, 128, 1) X(4, 64, 1) X(4, 128, 1) X(1, 128, 2)          \
    X(2, 64, 2) X(2, 128, 2) X(4, 64, 2) X(4, 128, 2) X(1, 64, 4) X(1, 128, 4) X(1, 256, 4)            \
    X(2, 64, 4) X(2, 128, 4) X(2, 256, 4) X(4, 64, 4) X(4, 128, 4) X(4, 256, 4)

#define L2O_SWEEP_CASE(K_, T_, W_)                                                                   \
    if (cfg.k == K_ && cfg.t == T_ && cfg.waves == W_) {                                            \
        hipLaunchKernelGGL((k_sweep<K_, T_, W_>), dim3(nitems), dim3(64 * W_), 0, stream, items, insts); \
        return hipGetLastError();                                                                    \
    }

hipError_t launch_sweep(const sweep_cfg &cfg, const l2o_sweep_item *items, uint32_t nitems,
                        l2o_inst *insts, hipStream_t stream)
{
    if (nitems == 0) return hipSuccess;
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
