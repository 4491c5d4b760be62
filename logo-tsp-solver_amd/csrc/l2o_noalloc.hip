// l2o_noalloc.hip -- logo2opt_step_noalloc: one best-improvement sweep
// (the logo2opt_step / CUDASolver::optimizeStep result) that allocates no
// memory, for callers that cannot allocate (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// Everything the sweep needs beyond the caller's inputs and output lives in
// one fixed device array, `scratch`, of NA_BYTES (about 3 MiB): the copied
// coordinates and tour, the tour-ordered points with their exact edge
// lengths, the instance record and the sweep's work list.  Its device address
// is taken once per device with hipGetSymbolAddress (the array's name is a
// host-side stand-in that must never reach a kernel); every call on that
// device reuses the same memory and rewrites all of it that the sweep reads,
// so calls that overlap on one device race and must be serialised by the
// caller.  Inputs are checked on the host with no allocation (the tour's
// permutation test uses an 8 KiB bitset on the stack).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>

#include "../../include/logo2opt.h"
#include "l2o_launch.h"
#include "l2o_state.h"

namespace {

constexpr uint32_t NA_MAX_N = LOGO2OPT_NOALLOC_MAX_N;
// one fixed tile shape (4 columns per lane, 128 rows, one wave): the work
// list's length then has the closed bound below, and no tuning table is read
constexpr int NA_K = 4, NA_T = 128, NA_W = 1;
constexpr uint32_t NA_COLS = NA_W * (64 * NA_K - 1);

// items of a sweep of n cities: row blocks a0 = 0, T, 2T, ... (a0 + 4 <= n),
// each with floor((n - 4 - a0) / COLS) + 1 column tiles
constexpr uint32_t na_items(uint32_t n)
{
    uint32_t c = 0;
    for (uint32_t a0 = 0; a0 + 4 <= n; a0 += NA_T) c += (n - 4 - a0) / NA_COLS + 1;
    return c;
}
constexpr uint32_t NA_MAX_ITEMS = na_items(NA_MAX_N);  // the count grows with n

constexpr size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t NA_OFF_COORDS = 0;                                                      // float x, y
constexpr size_t NA_OFF_TOUR = NA_OFF_COORDS + al256(sizeof(logo2opt_city_coords) * NA_MAX_N);
constexpr size_t NA_OFF_PTS = NA_OFF_TOUR + al256(sizeof(uint32_t) * NA_MAX_N);
constexpr size_t NA_OFF_ITEMS = NA_OFF_PTS + al256(sizeof(l2o_pt) * NA_MAX_N);
constexpr size_t NA_OFF_INST = NA_OFF_ITEMS + al256(sizeof(l2o_sweep_item) * NA_MAX_ITEMS);
constexpr size_t NA_BYTES = NA_OFF_INST + al256(sizeof(l2o_inst));

}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {

// tour-ordered points: P[k] = city route[k] translated by (-mx, -my), with the
// exact metric length of the edge route[k-1] -> route[k] and its tier-1 bound
__global__ __launch_bounds__(256) void k_na_points(unsigned char *__restrict__ base, uint32_t n, uint32_t metric,
                                                   double mx, double my)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const logo2opt_city_coords *C = (const logo2opt_city_coords *)(base + NA_OFF_COORDS);
    const uint32_t *R = (const uint32_t *)(base + NA_OFF_TOUR);
    l2o_pt *P = (l2o_pt *)(base + NA_OFF_PTS);
    const uint32_t c = R[k], pc = R[k == 0 ? n - 1 : k - 1];
    // integral coordinates with spans below 2^22: the differences are exact
    const int32_t x = (int32_t)((double)C[c].x - mx), y = (int32_t)((double)C[c].y - my);
    const int32_t px = (int32_t)((double)C[pc].x - mx), py = (int32_t)((double)C[pc].y - my);
    const long long dx = (long long)(x - px), dy = (long long)(y - py);
    l2o_pt p;
    p.x = x;
    p.y = y;
    p.d = l2o::metric_dist(dx * dx + dy * dy, metric);
    p.w = l2o::ceil_sqrtg(p.d, l2o::tier1_g(metric));
    P[k] = p;
}

// the work list in row-block order (the engine's order before its sorting):
// one thread per row block writes its column tiles after those of the blocks
// above it
__global__ __launch_bounds__(64) void k_na_items(unsigned char *__restrict__ base, uint32_t n)
{
    const uint32_t rb = blockIdx.x * 64 + threadIdx.x;
    const uint32_t a0 = rb * NA_T;
    if (a0 + 4 > n) return;
    uint32_t off = 0;
    for (uint32_t a = 0; a < a0; a += NA_T) off += (n - 4 - a) / NA_COLS + 1;
    l2o_sweep_item *it = (l2o_sweep_item *)(base + NA_OFF_ITEMS);
    for (uint32_t b0 = a0 + 2; b0 + 2 <= n; b0 += NA_COLS) {
        l2o_sweep_item q;
        q.inst = 0;
        q.a0 = a0;
        q.b0 = b0;
        q.rows = NA_T;
        it[off++] = q;
    }
}

// device address of `scratch` per device, taken once (hipGetSymbolAddress)
constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mu;
unsigned char *g_na_ptr[NA_MAX_DEVICES] = {};

int na_scratch(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LOGO2OPT_EDEVICE;
        g_na_ptr[device] = (unsigned char *)p;
    }
    *out = g_na_ptr[device];
    return LOGO2OPT_OK;
}

}  // namespace

extern "C" int logo2opt_step_noalloc(int device, uint32_t n, const logo2opt_city_coords *coords, int metric,
                                     const uint32_t *route, logo2opt_best2_out *out)
{
    if (!coords || !route || !out || n < 3) return LOGO2OPT_EINVAL;
    if (metric < LOGO2OPT_METRIC_EUC_2D || metric > LOGO2OPT_METRIC_ATT) return LOGO2OPT_EINVAL;
    if (n > NA_MAX_N) return LOGO2OPT_ESIZE;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || device >= NA_MAX_DEVICES)
        return LOGO2OPT_EDEVICE;
    // coordinates: integral, spans below 2^22 (as logo2opt_create checks)
    double mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
    for (uint32_t k = 0; k < n; ++k) {
        const double x = coords[k].x, y = coords[k].y;
        if (!std::isfinite(x) || !std::isfinite(y) || x != std::floor(x) || y != std::floor(y)) return LOGO2OPT_ECOORD;
        mnx = std::fmin(mnx, x);
        mny = std::fmin(mny, y);
        mxx = std::fmax(mxx, x);
        mxy = std::fmax(mxy, y);
    }
    if (mxx - mnx >= (double)(1u << 22) || mxy - mny >= (double)(1u << 22)) return LOGO2OPT_ECOORD;
    // the tour: a permutation of 0..n-1
    uint64_t seen[NA_MAX_N / 64];
    std::memset(seen, 0, sizeof(uint64_t) * ((n + 63) / 64));
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t c = route[k];
        if (c >= n || (seen[c >> 6] >> (c & 63) & 1)) return LOGO2OPT_ETOUR;
        seen[c >> 6] |= 1ull << (c & 63);
    }
    out->i = out->j = 0;
    out->minchange = 0;
    if (n < 4) return LOGO2OPT_OK;  // no candidate move
    if (hipSetDevice(device) != hipSuccess) return LOGO2OPT_EDEVICE;
    unsigned char *base = nullptr;
    int rc = na_scratch(device, &base);
    if (rc) return rc;

    // the instance record (same fields as the engine's upload for a local
    // search), with its pointers into the scratch array
    l2o_inst in;
    std::memset(&in, 0, sizeof(in));
    in.pts[0] = in.pts[1] = (l2o_pt *)(base + NA_OFF_PTS);
    in.tour[0] = in.tour[1] = (uint32_t *)(base + NA_OFF_TOUR);
    in.n = n;
    // bound on |exact gain - FP32 estimate| for the coordinate span (the
    // engine's formula, l2o_engine.hip check_coords)
    const double dmax = std::sqrt((mxx - mnx) * (mxx - mnx) + (mxy - mny) * (mxy - mny));
    in.slack = std::nextafter((float)(1.0 + dmax * std::ldexp(1.0, -18) + std::ldexp(1.0, -8)), INFINITY);
    in.metric = (uint32_t)metric;
    in.group_lead = L2O_NO_GROUP;
    in.est_scale = metric == LOGO2OPT_METRIC_ATT ? (float)(1.0 / std::sqrt(10.0)) : 1.0f;
    in.est_add = metric == LOGO2OPT_METRIC_EUC_2D ? 0.0f : 0.5f;
    in.qshift = l2o::tier0_shift((uint32_t)std::fmax(mxx - mnx, mxy - mny));
    in.phase = L2O_LO;
    in.key = L2O_SENTINEL;

    hipStream_t st = hipStreamPerThread;
#define NA_TRY(call)                                      \
    do {                                                  \
        if ((call) != hipSuccess) return LOGO2OPT_EDEVICE; \
    } while (0)
    NA_TRY(hipMemcpyAsync(base + NA_OFF_COORDS, coords, sizeof(logo2opt_city_coords) * n, hipMemcpyHostToDevice, st));
    NA_TRY(hipMemcpyAsync(base + NA_OFF_TOUR, route, sizeof(uint32_t) * n, hipMemcpyHostToDevice, st));
    NA_TRY(hipMemcpyAsync(base + NA_OFF_INST, &in, sizeof(in), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_na_points, dim3((n + 255) / 256), dim3(256), 0, st, base, n, (uint32_t)metric, mnx, mny);
    NA_TRY(hipGetLastError());
    const uint32_t rblocks = (n - 4) / NA_T + 1;
    hipLaunchKernelGGL(k_na_items, dim3((rblocks + 63) / 64), dim3(64), 0, st, base, n);
    NA_TRY(hipGetLastError());
    NA_TRY(l2o::launch_sweep(l2o::sweep_cfg{NA_K, NA_T, NA_W}, (const l2o_sweep_item *)(base + NA_OFF_ITEMS),
                             na_items(n), (l2o_inst *)(base + NA_OFF_INST), st, false));
    unsigned long long key = L2O_SENTINEL;
    NA_TRY(hipMemcpyAsync(&key, base + NA_OFF_INST + offsetof(l2o_inst, key), sizeof(key), hipMemcpyDeviceToHost,
                          st));
    NA_TRY(hipStreamSynchronize(st));
#undef NA_TRY
    if (key != L2O_SENTINEL) {
        out->minchange = (int32_t)((long long)(key >> 40) - L2O_KEY_GAIN_BIAS);
        out->i = (uint32_t)(key >> 20) & 0xFFFFFu;
        out->j = (uint32_t)key & 0xFFFFFu;
    }
    return LOGO2OPT_OK;
}
