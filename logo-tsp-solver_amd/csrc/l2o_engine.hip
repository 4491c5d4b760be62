// l2o_engine.hip -- host side of the C ABI declared in include/logo2opt.h.
//
// An engine owns one HIP stream, one HBM arena (the admission budget) and an
// array of instance slots (l2o_inst, device resident).  A batched step is
//   sweep  (one launch per tile shape, covering every live instance)
//   apply  (one launch, gathers the winning move of every instance)
//   ils    (one launch, accept/revert + kick for instances at a local optimum)
// Steps are issued back to back; the host looks at the device only every
// `sync_every` steps, and then through one 4-byte counter of live instances.
// With sync_every > 1 the steps are captured once into a HIP graph and
// replayed.  The single-instance context (the CUDASolver replacement,
// CUDASolver.cpp:48-434) is an engine with one slot.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/logo2opt.h"
#include "l2o_launch.h"
#include "l2o_state.h"

using l2o::sweep_cfg;

// a failed HIP call returns LOGO2OPT_EDEVICE; with LOGO2OPT_DEBUG set in the
// environment it is also reported on stderr (call, file:line, HIP's message)
#define L2O_TRY(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            if (std::getenv("LOGO2OPT_DEBUG"))                                                     \
                std::fprintf(stderr, "logo2opt: %s (%s:%d): %s\n", #call, __FILE__, __LINE__,      \
                             hipGetErrorString(e_));                                               \
            return LOGO2OPT_EDEVICE;                                                               \
        }                                                                                          \
    } while (0)

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// tile-shape table per n bucket (written by the autotuner, tune.py)
// ---------------------------------------------------------------------------
struct tune_row {
    uint32_t max_n;
    sweep_cfg cfg;
};

// Two tables: 0 = one instance in flight (latency: small tiles fill the
// device), 1 = a batch (throughput: the batch fills the device, larger tiles
// waste less).  An engine uses table 1 when it holds L2O_BATCH_MIN or more
// live instances.
#define L2O_BATCH_MIN 8
std::mutex g_tune_mu;
std::vector<tune_row> g_tune[2] = {
    {{640, {1, 64, 1}}, {8192, {4, 128, 1}}, {LOGO2OPT_MAX_N, {8, 256, 1}}},
    {{640, {1, 64, 1}}, {8192, {4, 128, 1}}, {LOGO2OPT_MAX_N, {8, 256, 1}}},
};
bool g_tune_loaded[2] = {false, false};
const char *g_tune_file[2] = {"tuning_gfx950.txt", "tuning_batch_gfx950.txt"};
const char *g_tune_env[2] = {"LOGO2OPT_TUNING", "LOGO2OPT_TUNING_BATCH"};

// text table, one "max_n k t waves" row per line; returns rows accepted
int read_tuning_file(const char *path, std::vector<tune_row> &rows)
{
    FILE *f = std::fopen(path, "r");
    if (!f) return -1;
    unsigned mx;
    int k, t, w;
    while (std::fscanf(f, "%u %d %d %d", &mx, &k, &t, &w) == 4) {
        sweep_cfg c{k, t, w};
        if (!l2o::sweep_cfg_supported(c)) {
            std::fclose(f);
            return -1;
        }
        rows.push_back({mx, c});
    }
    std::fclose(f);
    return (int)rows.size();
}

void install_rows_locked(int table, std::vector<tune_row> rows)
{
    std::sort(rows.begin(), rows.end(), [](const tune_row &a, const tune_row &b) { return a.max_n < b.max_n; });
    rows.back().max_n = LOGO2OPT_MAX_N;
    g_tune[table] = rows;
}

// $LOGO2OPT_TUNING[_BATCH], else the table file next to this library
// (written by tune.py), else the built-in defaults above
void load_tuning_locked(int table)
{
    if (g_tune_loaded[table]) return;
    g_tune_loaded[table] = true;
    std::string path;
    const char *p = std::getenv(g_tune_env[table]);
    if (p && *p) {
        path = p;
    } else {
        Dl_info info;
        if (dladdr((void *)&read_tuning_file, &info) && info.dli_fname) {
            path = info.dli_fname;
            const size_t slash = path.rfind('/');
            path = (slash == std::string::npos ? std::string() : path.substr(0, slash + 1)) + g_tune_file[table];
        }
    }
    if (path.empty()) return;
    std::vector<tune_row> rows;
    if (read_tuning_file(path.c_str(), rows) > 0) install_rows_locked(table, rows);
}

sweep_cfg cfg_for(uint32_t n, int table)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    load_tuning_locked(table);
    for (const auto &r : g_tune[table])
        if (n <= r.max_n) return r.cfg;
    return g_tune[table].back().cfg;
}

// ---------------------------------------------------------------------------
// first-fit arena over one hipMalloc (the HBM admission budget)
// ---------------------------------------------------------------------------
struct arena {
    char *base = nullptr;
    uint64_t size = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> length

    void init(char *b, uint64_t s)
    {
        base = b;
        size = s;
        free_.clear();
        free_[0] = s;
    }
    bool alloc(uint64_t bytes, uint64_t *off)
    {
        bytes = (bytes + 255) & ~255ull;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const uint64_t rest = it->second - bytes;
                const uint64_t o = it->first + bytes;
                free_.erase(it);
                if (rest) free_[o] = rest;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes)
    {
        bytes = (bytes + 255) & ~255ull;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t free_bytes() const
    {
        uint64_t s = 0;
        for (auto &kv : free_) s += kv.second;
        return s;
    }
};

// exact metric length of the edge a -> b on the host (same function as the kernels)
inline uint32_t dist_i(const int32_t *xy, uint32_t a, uint32_t b, uint32_t metric)
{
    const int64_t dx = (int64_t)xy[2 * a] - xy[2 * b], dy = (int64_t)xy[2 * a + 1] - xy[2 * b + 1];
    return (uint32_t)l2o::metric_dist((long long)(dx * dx + dy * dy), metric);
}

uint64_t pairs_of(uint32_t n) { return n >= 4 ? (uint64_t)(n - 2) * (uint64_t)(n - 3) / 2 : 0; }

struct slot_t {
    bool used = false;
    uint32_t n = 0;
    uint64_t off = 0, bytes = 0;
    logo2opt_job job{};
    std::vector<int32_t> xy;  // translated integer coordinates
    bool done = false;
    bool done_reported = false;
    float slack = 1.0f;
    uint32_t group_lead = L2O_NO_GROUP;  // multi-start ILS group (first slot)
    uint32_t group_size = 0;
};

}  // namespace

struct logo2opt_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    char *d_arena = nullptr;
    arena ar;
    uint32_t max_slots = 0;
    std::vector<slot_t> slots;
    l2o_inst *d_insts = nullptr;
    unsigned *d_live = nullptr;
    unsigned *h_live = nullptr;  // pinned
    unsigned live = 0;           // host mirror, valid between runs
    unsigned long long *d_scratch = nullptr;
    unsigned long long *h_scratch = nullptr;  // pinned

    // work lists
    struct group {
        sweep_cfg cfg;
        uint32_t off, count, id;
        int mode;  // sweep variant (L2O_SWEEP_*)
    };
    std::vector<group> groups;
    l2o_sweep_item *d_sitems = nullptr;
    size_t sitems_cap = 0;

    l2o_apply_item *d_aitems = nullptr;
    size_t aitems_cap = 0;
    uint32_t n_aitems = 0;
    uint32_t ils_grid = 0;    // k_ils covers slots [0, ils_grid)
    bool any_ils = false;
    uint32_t group_grid = 0;  // k_group covers slots [0, group_grid)
    bool dirty = true;
    unsigned live_at_build = 0;

    hipGraphExec_t gexec = nullptr;
    uint32_t gsteps = 0;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double wall_ms = 0.0, sweep_ms = 0.0;
    uint64_t sweep_launches = 0;
    std::vector<int32_t> done_queue;

    // timing mode: every step's sweep launches are bracketed by HIP events on
    // the engine stream (no graphs); elapsed times are summed at each sync
    bool timing = false;
    // batches of steps are captured into a HIP graph unless disabled
    // (logo2opt_engine_set_graphs, or LOGO2OPT_GRAPHS=0 in the environment)
    bool graphs = true;
    // sweeps skip the blocks of rows that the bounding-box bound rules out
    // (logo2opt_engine_set_pruning; same moves, fewer candidates evaluated):
    // LOGO2OPT_PRUNE_OFF, _ALL, or _AUTO (instances of L2O_PRUNE_AUTO_MIN_N
    // cities or more)
    int prune = LOGO2OPT_PRUNE_OFF;
    std::vector<hipEvent_t> tev;
    uint32_t tev_used = 0;
};

struct logo2opt_ctx {
    logo2opt_engine *e = nullptr;
    logo2opt_engine *ge = nullptr;  // multi-start ILS groups (created on first use)
    uint32_t ge_chains = 0;
    uint32_t n = 0;
    uint32_t metric = LOGO2OPT_METRIC_EUC_2D;
    int prune = LOGO2OPT_PRUNE_OFF;  // applied to ge when it is created
    std::vector<logo2opt_city_coords> coords;
    int32_t slot = -1;
    uint64_t sweeps = 0, moves = 0, evals = 0, fallbacks = 0;
    double wall_ms = 0.0, sweep_ms = 0.0;
};

namespace {

int check_coords(uint32_t n, const logo2opt_city_coords *c, std::vector<int32_t> &xy, float *slack)
{
    if (!c) return LOGO2OPT_EINVAL;
    xy.resize(2 * (size_t)n);
    double mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
    for (uint32_t k = 0; k < n; ++k) {
        const double x = c[k].x, y = c[k].y;
        if (!std::isfinite(x) || !std::isfinite(y) || x != std::floor(x) || y != std::floor(y))
            return LOGO2OPT_ECOORD;
        mnx = std::min(mnx, x);
        mny = std::min(mny, y);
        mxx = std::max(mxx, x);
        mxy = std::max(mxy, y);
    }
    const double lim = (double)(1u << 22);
    if (mxx - mnx >= lim || mxy - mny >= lim) return LOGO2OPT_ECOORD;
    for (uint32_t k = 0; k < n; ++k) {
        xy[2 * k] = (int32_t)(c[k].x - mnx);
        xy[2 * k + 1] = (int32_t)(c[k].y - mny);
    }
    // |exact gain - FP32 estimate| <= 1 (|D - (sqrt(s) scale + add)| <= 1/2 per
    // new edge for every metric) + 2 e_d + rounding of the three FP32 adds and
    // of the threshold, with e_d <= dmax * 2^-21 (covers FP32 rounding of s,
    // up to 3.5 ulp of v_sqrt_f32 -- tests check v_sqrt_f32 against the
    // correctly rounded sqrt -- and for ATT the rounding of 1/sqrt(10)), each
    // add within ulp(4 dmax) <= dmax * 2^-21 and each estimate's fma within
    // dmax * 2^-23.  Total <= 1 + dmax * 2^-18; 2^-8 margin.
    const double dmax = std::sqrt((mxx - mnx) * (mxx - mnx) + (mxy - mny) * (mxy - mny));
    const double s = 1.0 + dmax * std::ldexp(1.0, -18) + std::ldexp(1.0, -8);
    *slack = std::nextafter((float)s, INFINITY);
    return LOGO2OPT_OK;
}

int check_tour(uint32_t n, const uint32_t *r)
{
    if (!r) return LOGO2OPT_EINVAL;
    std::vector<unsigned char> seen(n, 0);
    for (uint32_t k = 0; k < n; ++k) {
        if (r[k] >= n || seen[r[k]]) return LOGO2OPT_ETOUR;
        seen[r[k]] = 1;
    }
    return LOGO2OPT_OK;
}

uint64_t bytes_for(uint32_t n, const logo2opt_job *job)
{
    auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
    uint64_t b = 2 * al(16ull * n) + 2 * al(4ull * n);
    if (job && job->ils_rounds) b += al(16ull * n) + al(4ull * n) + al(8ull * ((uint64_t)job->ils_rounds + 1));
    if (job && job->log_cap) b += al(12ull * job->log_cap);
    return b;
}

// the context's engine budget: one instance with `job`'s move log and round
// log, plus 16 MiB
uint64_t ctx_budget(uint32_t n, const logo2opt_job &job) { return bytes_for(n, &job) + (1ull << 24); }

// writes the ordered points / tour of `route` into slot buffers and the
// instance record; the stream must be idle for this slot.
int upload_instance(logo2opt_engine *e, int32_t id, const uint32_t *route, uint32_t phase)
{
    slot_t &s = e->slots[id];
    const uint32_t n = s.n;
    auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
    char *base = e->d_arena + s.off;
    l2o_inst in;
    std::memset(&in, 0, sizeof(in));
    in.pts[0] = (l2o_pt *)base;
    in.pts[1] = (l2o_pt *)(base + al(16ull * n));
    in.tour[0] = (uint32_t *)(base + 2 * al(16ull * n));
    in.tour[1] = (uint32_t *)(base + 2 * al(16ull * n) + al(4ull * n));
    char *p = base + 2 * al(16ull * n) + 2 * al(4ull * n);
    if (s.job.ils_rounds) {
        in.bpts = (l2o_pt *)p;
        p += al(16ull * n);
        in.btour = (uint32_t *)p;
        p += al(4ull * n);
        in.rlog = (unsigned long long *)p;
        p += al(8ull * ((uint64_t)s.job.ils_rounds + 1));
    }
    if (s.job.log_cap) in.log = (l2o_move *)p;
    in.n = n;
    in.log_cap = s.job.log_cap;
    in.slack = s.slack;
    in.ils_rounds = s.job.ils_rounds;
    in.kick = s.job.kick;
    in.metric = s.job.metric;
    in.group_lead = s.group_lead;
    in.group_size = s.group_size;
    in.kick_src = (uint32_t)id;
    in.est_scale = s.job.metric == l2o::L2O_ATT ? (float)(1.0 / std::sqrt(10.0)) : 1.0f;
    in.est_add = s.job.metric == l2o::L2O_EUC_2D ? 0.0f : 0.5f;
    {
        int32_t m = 0;
        for (const int32_t v : s.xy) m = std::max(m, v);
        in.qshift = l2o::tier0_shift((uint32_t)m);
    }
    in.stop_below = s.job.ils_rounds ? s.job.stop_below : 0;
    in.phase = phase;
    in.key = L2O_SENTINEL;
    // SplitMix64 -> xoshiro256** state (same seeding as the workload generators)
    uint64_t x = s.job.seed;
    for (int k = 0; k < 4; ++k) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        in.rng[k] = z ^ (z >> 31);
    }
    std::vector<l2o_pt> pts(n);
    uint64_t L = 0;
    const int32_t *xy = s.xy.data();
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t c = route[k], pc = route[k == 0 ? n - 1 : k - 1];
        const uint32_t d = dist_i(xy, pc, c, s.job.metric);
        L += d;
        pts[k].x = xy[2 * c];
        pts[k].y = xy[2 * c + 1];
        pts[k].d = (int32_t)d;
        pts[k].w = l2o::ceil_sqrtg((int32_t)d, l2o::tier1_g(s.job.metric));
    }
    in.length = L;
    L2O_TRY(hipMemcpyAsync(in.pts[0], pts.data(), 16ull * n, hipMemcpyHostToDevice, e->stream));
    L2O_TRY(hipMemcpyAsync(in.tour[0], route, 4ull * n, hipMemcpyHostToDevice, e->stream));
    L2O_TRY(hipMemcpyAsync(e->d_insts + id, &in, sizeof(in), hipMemcpyHostToDevice, e->stream));
    L2O_TRY(hipStreamSynchronize(e->stream));
    return LOGO2OPT_OK;
}

void drop_graph(logo2opt_engine *e)
{
    if (e->gexec) {
        (void)hipGraphExecDestroy(e->gexec);
        e->gexec = nullptr;
    }
    e->gsteps = 0;
}

// LOGO2OPT_PRUNE_AUTO prunes instances of at least this many cities: below
// it the pruned sweep measured 4-9% slower (per-wave fixed costs dominate),
// from 16,000 on it wins (1.04x there, 1.2x at 24,000, 1.6x at 33,810)
#define L2O_PRUNE_AUTO_MIN_N 16000u
// Pruned instances below this many cities send their live blocks through
// tier 0; larger ones, whose live blocks lie close to the columns and mostly
// pass it, go straight to tier 1 (measured: 33,810 and 100,000 cities 3-5%
// faster without tier 0, the 2,048-instance batch of 100-5,000 cities 7%
// faster with it)
#ifndef L2O_PRUNE_T0_MAX_N
#define L2O_PRUNE_T0_MAX_N 16000u
#endif

#ifndef L2O_DIAG_WEIGHT
#define L2O_DIAG_WEIGHT 4
#endif

// 1/L2O_TAIL_SPLIT of a launch's work, at its end, goes in items of
// 1/L2O_TAIL_DIV of the shape's rows (T/L2O_TAIL_DIV must stay a multiple of 4)
#ifndef L2O_TAIL_SPLIT
#define L2O_TAIL_SPLIT 8
#endif
#ifndef L2O_TAIL_DIV
#define L2O_TAIL_DIV 4
#endif

// candidates (a, b), a+2 <= b <= n-2, of a sweep item (closed form)
uint64_t item_work(const l2o_sweep_item &it, const sweep_cfg &c, uint32_t n)
{
    const int64_t b0 = it.b0, a0 = it.a0;
    const int64_t a_end = std::min<int64_t>(a0 + it.rows, (int64_t)n - 3);  // rows a0 .. a_end-1
    const int64_t bhi = std::min<int64_t>(b0 + c.cols() - 1, (int64_t)n - 2);
    if (a_end <= a0 || bhi < b0) return 0;
    // rows a <= b0-2 see all of [b0, bhi]
    const int64_t r1 = std::max<int64_t>(0, std::min(a_end, b0 - 1) - a0);
    uint64_t w = (uint64_t)r1 * (uint64_t)(bhi - b0 + 1);
    // rows a >= b0-1 see [a+2, bhi]: bhi-1-a candidates while a <= bhi-2
    const int64_t lo = std::max(a0, b0 - 1), hi = std::min(a_end, bhi - 1);  // a in [lo, hi)
    if (hi > lo) w += (uint64_t)((2 * (bhi - 1) - lo - (hi - 1)) * (hi - lo) / 2);
    return w;
}

int rebuild_lists(logo2opt_engine *e)
{
    std::map<std::tuple<int, int, int, int>, std::vector<l2o_sweep_item>> by_cfg;
    std::vector<l2o_apply_item> ai;
    e->ils_grid = 0;
    e->any_ils = false;
    e->group_grid = 0;
    // largest instances first: their long sweeps start early and the small
    // ones fill the tail (longest-processing-time order)
    std::vector<uint32_t> order;
    for (uint32_t id = 0; id < e->max_slots; ++id)
        if (e->slots[id].used && !e->slots[id].done) order.push_back(id);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return e->slots[a].n > e->slots[b].n; });
    const int table = order.size() >= L2O_BATCH_MIN ? 1 : 0;
    for (uint32_t id : order) {
        const slot_t &s = e->slots[id];
        const uint32_t n = s.n;
        const sweep_cfg c = cfg_for(n, table);
        const bool pr = e->prune == LOGO2OPT_PRUNE_ALL || (e->prune == LOGO2OPT_PRUNE_AUTO && n >= L2O_PRUNE_AUTO_MIN_N);
        const int mode = !pr ? l2o::L2O_SWEEP_FULL : n < L2O_PRUNE_T0_MAX_N ? l2o::L2O_SWEEP_PRUNE : l2o::L2O_SWEEP_PRUNE_T1;
        auto &v = by_cfg[std::make_tuple(c.k, c.t, c.waves, mode)];
        // rows a <= n-4; the first tile of a row block starts at the
        // diagonal, b0 = a0 + 2
        for (uint32_t a0 = 0; a0 + 4 <= n; a0 += c.t)
            for (uint32_t b0 = a0 + 2; b0 + 2 <= n; b0 += c.cols()) v.push_back({id, a0, b0, (uint32_t)c.t});
        const uint32_t nch = (n + L2O_APPLY_CHUNK - 1) / L2O_APPLY_CHUNK;
        for (uint32_t ch = 0; ch < nch; ++ch) ai.push_back({id, ch});
        if (s.job.ils_rounds) {
            e->any_ils = true;
            e->ils_grid = std::max(e->ils_grid, id + 1);  // `order` is not by id
        }
        if (s.group_lead != L2O_NO_GROUP) e->group_grid = std::max(e->group_grid, id + 1);
    }
    std::vector<l2o_sweep_item> all;
    e->groups.clear();
    for (auto &kv : by_cfg) {
        const sweep_cfg c{std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first)};
        const int mode = std::get<3>(kv.first);
        // Items are dispatched in list order (largest instances first), so at
        // any moment the device works on a window of consecutive items of one
        // or a few instances, whose points stay hot in every XCD's L2.
        // longest items first (candidates per item: full tiles, then the
        // partial tiles at the diagonal and the right and bottom edges), so
        // the dispatch tail is made of the short ones; stable, so equal items
        // keep their instance order
        // (With row pruning the tiles near the diagonal keep most of their
        // rows and the far ones are mostly skipped; the same order, which
        // spreads the near-diagonal tiles over the launch, measured faster
        // than issuing them first.)
        auto &v = kv.second;
        std::vector<uint64_t> work(v.size());
        // In a batch the first tile of each row block (at the diagonal) is
        // ordered as if it had L2O_DIAG_WEIGHT times its candidates: its rows
        // and columns are tour neighbours, close in the plane, so more of its
        // blocks pass tier 0 and it runs longer than its count says.  Issued
        // earlier it no longer ends the launch: the 2,048-instance batch
        // measured 2.15 -> 1.84 s (the same for weights 3 to 8).  One
        // instance keeps the plain order: there every slow path publishes to
        // the same key, and the near-diagonal tiles run best spread over the
        // launch (weighted, the config-2 local search took 0.39 s, not 0.285).
        for (size_t q = 0; q < v.size(); ++q) {
            work[q] = item_work(v[q], c, e->slots[v[q].inst].n);
            if (table == 1 && v[q].b0 == v[q].a0 + 2) work[q] *= L2O_DIAG_WEIGHT;
        }
        std::vector<uint32_t> idx(v.size());
        for (uint32_t q = 0; q < idx.size(); ++q) idx[q] = q;
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) { return work[x] > work[y]; });
        std::vector<l2o_sweep_item> sorted;
        sorted.reserve(v.size() + v.size() / 2);
        // The last L2O_TAIL_SPLIT of the work (the end of the dispatch) is
        // issued as items of t/L2O_TAIL_DIV rows, so the launch drains in
        // short pieces.
        uint64_t total = 0, done = 0;
        for (uint64_t w : work) total += w;
        for (size_t q = 0; q < idx.size(); ++q) {
            const l2o_sweep_item it = v[idx[q]];
            if (c.t >= 4 * L2O_TAIL_DIV && done >= total - total / L2O_TAIL_SPLIT) {
                const uint32_t r4 = (uint32_t)c.t / L2O_TAIL_DIV;
                for (uint32_t r = 0; r < it.rows; r += r4) sorted.push_back({it.inst, it.a0 + r, it.b0, r4});
            } else {
                sorted.push_back(it);
            }
            done += work[idx[q]];
        }
        v.swap(sorted);
        const uint32_t cnt = (uint32_t)v.size();
        if (e->groups.size() >= L2O_MAX_GROUPS) return LOGO2OPT_EINVAL;
        e->groups.push_back({c, (uint32_t)all.size(), cnt, (uint32_t)e->groups.size(), mode});
        all.insert(all.end(), v.begin(), v.end());
    }
    if (all.size() > e->sitems_cap) {
        if (e->d_sitems) (void)hipFree(e->d_sitems);
        e->sitems_cap = std::max<size_t>(all.size(), 2 * e->sitems_cap);
        L2O_TRY(hipMalloc(&e->d_sitems, e->sitems_cap * sizeof(l2o_sweep_item)));
    }
    if (ai.size() > e->aitems_cap) {
        if (e->d_aitems) (void)hipFree(e->d_aitems);
        e->aitems_cap = std::max<size_t>(ai.size(), 2 * e->aitems_cap);
        L2O_TRY(hipMalloc(&e->d_aitems, e->aitems_cap * sizeof(l2o_apply_item)));
    }
    if (!all.empty())
        L2O_TRY(hipMemcpyAsync(e->d_sitems, all.data(), all.size() * sizeof(l2o_sweep_item), hipMemcpyHostToDevice,
                               e->stream));
    if (!ai.empty())
        L2O_TRY(hipMemcpyAsync(e->d_aitems, ai.data(), ai.size() * sizeof(l2o_apply_item), hipMemcpyHostToDevice,
                               e->stream));
    L2O_TRY(hipStreamSynchronize(e->stream));
    e->n_aitems = (uint32_t)ai.size();
    e->dirty = false;
    e->live_at_build = e->live;
    drop_graph(e);
    return LOGO2OPT_OK;
}

// the sweep launch of one group
hipError_t launch_group_sweep(const logo2opt_engine *e, const logo2opt_engine::group &g)
{
    return l2o::launch_sweep(g.cfg, e->d_sitems + g.off, g.count, e->d_insts, e->stream, g.mode);
}

int issue_step(logo2opt_engine *e)
{
    const bool tm = e->timing && !e->groups.empty();
    if (tm) {
        if (e->tev_used + 2 > e->tev.size()) {
            for (int q = 0; q < 64; ++q) {
                hipEvent_t ev;
                L2O_TRY(hipEventCreate(&ev));
                e->tev.push_back(ev);
            }
        }
        L2O_TRY(hipEventRecord(e->tev[e->tev_used], e->stream));
    }
    for (const auto &g : e->groups) {
        L2O_TRY(launch_group_sweep(e, g));
        e->sweep_launches += 1;
    }
    if (tm) {
        L2O_TRY(hipEventRecord(e->tev[e->tev_used + 1], e->stream));
        e->tev_used += 2;
    }
    L2O_TRY(l2o::launch_apply(e->d_aitems, e->n_aitems, e->d_insts, e->d_live, e->stream));
    if (e->any_ils) L2O_TRY(l2o::launch_ils(e->d_insts, e->ils_grid, e->d_live, e->stream));
    if (e->group_grid) L2O_TRY(l2o::launch_group(e->d_insts, e->group_grid, e->d_live, e->stream));
    return LOGO2OPT_OK;
}

// after a stream sync: add the timed sweep intervals
int collect_timing(logo2opt_engine *e)
{
    for (uint32_t q = 0; q + 1 < e->tev_used; q += 2) {
        float ms = 0.0f;
        L2O_TRY(hipEventElapsedTime(&ms, e->tev[q], e->tev[q + 1]));
        e->sweep_ms += ms;
    }
    e->tev_used = 0;
    return LOGO2OPT_OK;
}

int issue_steps(logo2opt_engine *e, uint32_t k, uint32_t sync_every)
{
    if (sync_every > 1 && k == sync_every && !e->timing && e->graphs) {
        if (!e->gexec || e->gsteps != k) {
            drop_graph(e);
            hipGraph_t g;
            L2O_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            int rc = LOGO2OPT_OK;
            for (uint32_t s = 0; s < k && rc == LOGO2OPT_OK; ++s) rc = issue_step(e);
            hipError_t ce = hipStreamEndCapture(e->stream, &g);
            if (rc != LOGO2OPT_OK) return rc;
            L2O_TRY(ce);
            hipError_t ie = hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            L2O_TRY(ie);
            e->gsteps = k;
        }
        L2O_TRY(hipGraphLaunch(e->gexec, e->stream));
        return LOGO2OPT_OK;
    }
    for (uint32_t s = 0; s < k; ++s) {
        int rc = issue_step(e);
        if (rc) return rc;
    }
    return LOGO2OPT_OK;
}

// refresh host knowledge of finished slots
int scan_done(logo2opt_engine *e)
{
    std::vector<l2o_inst> h(e->max_slots);
    L2O_TRY(hipMemcpyAsync(h.data(), e->d_insts, sizeof(l2o_inst) * e->max_slots, hipMemcpyDeviceToHost, e->stream));
    L2O_TRY(hipStreamSynchronize(e->stream));
    for (uint32_t id = 0; id < e->max_slots; ++id) {
        slot_t &s = e->slots[id];
        if (s.used && !s.done && h[id].phase == L2O_DONE) {
            s.done = true;
            e->done_queue.push_back((int32_t)id);
        }
    }
    return LOGO2OPT_OK;
}

int engine_run(logo2opt_engine *e, uint32_t max_steps, uint32_t sync_every, uint32_t *live_out)
{
    if (sync_every == 0) sync_every = 1;
    const double t0 = now_ms();
    if (e->dirty) {
        int rc = rebuild_lists(e);
        if (rc) return rc;
    }
    uint32_t steps = 0;
    while (steps < max_steps && e->live > 0) {
        const uint32_t k = std::min(sync_every, max_steps - steps);
        int rc = issue_steps(e, k, sync_every);
        if (rc) return rc;
        steps += k;
        L2O_TRY(hipMemcpyAsync(e->h_live, e->d_live, sizeof(unsigned), hipMemcpyDeviceToHost, e->stream));
        L2O_TRY(hipStreamSynchronize(e->stream));
        if (e->timing) {
            rc = collect_timing(e);
            if (rc) return rc;
        }
        const unsigned lv = *e->h_live;
        if (lv != e->live) {
            e->live = lv;
            rc = scan_done(e);
            if (rc) return rc;
            // drop finished instances from the work lists once they are a
            // noticeable share of the launched workgroups
            if (e->live == 0 || e->live * 4 < e->live_at_build * 3) {
                rc = rebuild_lists(e);
                if (rc) return rc;
            }
        }
    }
    e->wall_ms += now_ms() - t0;
    if (live_out) *live_out = e->live;
    return LOGO2OPT_OK;
}

// admits an instance into slot `id` (free); group_lead / group_size as in l2o_inst
int engine_add_at(logo2opt_engine *e, int32_t id, uint32_t n, const logo2opt_city_coords *coords,
                  const uint32_t *route, const logo2opt_job *job, uint32_t group_lead, uint32_t group_size)
{
    if (n < 3 || n > LOGO2OPT_MAX_N) return LOGO2OPT_EINVAL;
    slot_t tmp;
    int rc = check_coords(n, coords, tmp.xy, &tmp.slack);
    if (rc) return rc;
    rc = check_tour(n, route);
    if (rc) return rc;
    tmp.group_lead = group_lead;
    tmp.group_size = group_size;
    logo2opt_job j{};
    if (job) j = *job;
    if (j.kick > 1 || j.metric > LOGO2OPT_METRIC_ATT) return LOGO2OPT_EINVAL;
    if (j.ils_rounds && n < 8) return LOGO2OPT_EINVAL;  // double-bridge needs room for 3 cuts
    const uint64_t bytes = bytes_for(n, &j);
    // larger than the whole budget: it can never be admitted
    if (bytes > e->ar.size) return LOGO2OPT_ENOMEM;
    uint64_t off;
    if (!e->ar.alloc(bytes, &off)) return LOGO2OPT_EAGAIN;
    slot_t &s = e->slots[id];
    s = std::move(tmp);
    s.used = true;
    s.n = n;
    s.off = off;
    s.bytes = bytes;
    s.job = j;
    s.done = false;
    s.done_reported = false;
    rc = upload_instance(e, id, route, L2O_LO);
    if (rc) {
        e->ar.release(off, bytes);
        s = slot_t();
        return rc;
    }
    e->live += 1;
    *e->h_live = e->live;
    L2O_TRY(hipMemcpyAsync(e->d_live, e->h_live, sizeof(unsigned), hipMemcpyHostToDevice, e->stream));
    L2O_TRY(hipStreamSynchronize(e->stream));
    e->dirty = true;
    return LOGO2OPT_OK;
}

int engine_add(logo2opt_engine *e, uint32_t n, const logo2opt_city_coords *coords, const uint32_t *route,
               const logo2opt_job *job, int32_t *slot_out)
{
    int32_t id = -1;
    for (uint32_t k = 0; k < e->max_slots; ++k)
        if (!e->slots[k].used) {
            id = (int32_t)k;
            break;
        }
    if (id < 0) return LOGO2OPT_EAGAIN;
    const int rc = engine_add_at(e, id, n, coords, route, job, L2O_NO_GROUP, 0);
    if (rc == LOGO2OPT_OK) *slot_out = id;
    return rc;
}

int engine_remove(logo2opt_engine *e, int32_t id);

// `chains` ILS chains of one instance in consecutive slots; chain c uses the
// seed job->seed + c
int engine_add_group(logo2opt_engine *e, uint32_t n, const logo2opt_city_coords *coords, const uint32_t *route,
                     const logo2opt_job *job, uint32_t chains, int32_t *first_out)
{
    if (!job || job->ils_rounds == 0 || chains == 0 || chains > e->max_slots) return LOGO2OPT_EINVAL;
    int32_t first = -1;
    for (uint32_t k = 0; k + chains <= e->max_slots && first < 0; ++k) {
        uint32_t c = 0;
        while (c < chains && !e->slots[k + c].used) ++c;
        if (c == chains) first = (int32_t)k;
        else k += c;  // slot k + c is used: resume after it
    }
    if (first < 0) return LOGO2OPT_EAGAIN;
    for (uint32_t c = 0; c < chains; ++c) {
        logo2opt_job j = *job;
        j.seed = job->seed + c;
        const int rc = engine_add_at(e, first + (int32_t)c, n, coords, route, &j, (uint32_t)first, chains);
        if (rc) {
            for (uint32_t q = 0; q < c; ++q) e->slots[first + q].group_lead = L2O_NO_GROUP;
            for (uint32_t q = 0; q < c; ++q) (void)engine_remove(e, first + (int32_t)q);
            return rc;
        }
    }
    *first_out = first;
    return LOGO2OPT_OK;
}

int engine_result(logo2opt_engine *e, int32_t id, uint32_t *route_out, logo2opt_result *res, logo2opt_best2_out *log)
{
    if (id < 0 || (uint32_t)id >= e->max_slots || !e->slots[id].used) return LOGO2OPT_ENOTFOUND;
    l2o_inst in;
    L2O_TRY(hipMemcpyAsync(&in, e->d_insts + id, sizeof(in), hipMemcpyDeviceToHost, e->stream));
    L2O_TRY(hipStreamSynchronize(e->stream));
    const bool ils = in.ils_rounds > 0;
    if (res) {
        res->length = ils && in.have_best ? in.best_length : in.length;
        res->moves = in.moves;
        res->sweeps = in.sweeps;
        res->fallbacks = in.fallbacks;
        res->rounds_done = in.rounds_done;
        res->done = in.phase == L2O_DONE;
    }
    if (route_out) {
        const uint32_t *src = ils && in.have_best ? in.btour : in.tour[in.parity];
        L2O_TRY(hipMemcpyAsync(route_out, src, 4ull * in.n, hipMemcpyDeviceToHost, e->stream));
    }
    if (log && in.log_cap) {
        const uint64_t m = std::min<uint64_t>(in.moves, in.log_cap);
        if (m) L2O_TRY(hipMemcpyAsync(log, in.log, 12ull * m, hipMemcpyDeviceToHost, e->stream));
    }
    L2O_TRY(hipStreamSynchronize(e->stream));
    return LOGO2OPT_OK;
}

int engine_remove(logo2opt_engine *e, int32_t id)
{
    if (id < 0 || (uint32_t)id >= e->max_slots || !e->slots[id].used) return LOGO2OPT_ENOTFOUND;
    {
        // a chain of a multi-start group cannot run on without the others:
        // the whole group is retired
        const uint32_t lead = e->slots[id].group_lead;
        if (lead != L2O_NO_GROUP) {
            const uint32_t m = e->slots[lead].group_size;
            for (uint32_t c = 0; c < m; ++c) e->slots[lead + c].group_lead = L2O_NO_GROUP;
            int rc = LOGO2OPT_OK;
            for (uint32_t c = 0; c < m; ++c)
                if (e->slots[lead + c].used && rc == LOGO2OPT_OK) rc = engine_remove(e, (int32_t)(lead + c));
            return rc;
        }
    }
    slot_t &s = e->slots[id];
    if (!s.done) {
        // retire a running instance: mark it done on the device first
        l2o_inst in;
        L2O_TRY(hipMemcpyAsync(&in, e->d_insts + id, sizeof(in), hipMemcpyDeviceToHost, e->stream));
        L2O_TRY(hipStreamSynchronize(e->stream));
        if (in.phase != L2O_DONE) {
            e->live -= 1;
            *e->h_live = e->live;
            L2O_TRY(hipMemcpyAsync(e->d_live, e->h_live, sizeof(unsigned), hipMemcpyHostToDevice, e->stream));
        }
    }
    l2o_inst blank;
    std::memset(&blank, 0, sizeof(blank));
    blank.phase = L2O_FREE;
    L2O_TRY(hipMemcpyAsync(e->d_insts + id, &blank, sizeof(blank), hipMemcpyHostToDevice, e->stream));
    L2O_TRY(hipStreamSynchronize(e->stream));
    e->ar.release(s.off, s.bytes);
    s = slot_t();
    e->dirty = true;
    return LOGO2OPT_OK;
}

int engine_create(int device, uint64_t budget, uint32_t max_slots, logo2opt_engine **out)
{
    if (!out || max_slots == 0) return LOGO2OPT_EINVAL;
    if (budget == 0) budget = 1ull << 30;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return LOGO2OPT_EDEVICE;
    L2O_TRY(hipSetDevice(device));
    logo2opt_engine *e = new (std::nothrow) logo2opt_engine();
    if (!e) return LOGO2OPT_ENOMEM;
    e->device = device;
    e->max_slots = max_slots;
    e->slots.resize(max_slots);
    if (const char *g = std::getenv("LOGO2OPT_GRAPHS")) e->graphs = std::strcmp(g, "0") != 0;
    bool ok = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc(&e->d_arena, budget) == hipSuccess &&
              hipMalloc(&e->d_insts, sizeof(l2o_inst) * max_slots) == hipSuccess &&
              hipMalloc(&e->d_live, sizeof(unsigned)) == hipSuccess &&
              hipMalloc(&e->d_scratch, sizeof(unsigned long long) * 4) == hipSuccess &&
              hipHostMalloc(&e->h_live, sizeof(unsigned), hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc(&e->h_scratch, sizeof(unsigned long long) * 4, hipHostMallocDefault) == hipSuccess &&
              hipEventCreate(&e->ev0) == hipSuccess && hipEventCreate(&e->ev1) == hipSuccess;
    if (ok) {
        e->ar.init(e->d_arena, budget);
        ok = hipMemsetAsync(e->d_insts, 0, sizeof(l2o_inst) * max_slots, e->stream) == hipSuccess &&
             hipMemsetAsync(e->d_live, 0, sizeof(unsigned), e->stream) == hipSuccess &&
             hipStreamSynchronize(e->stream) == hipSuccess;
    }
    if (!ok) {
        logo2opt_engine_destroy(e);
        return LOGO2OPT_EDEVICE;
    }
    *out = e;
    return LOGO2OPT_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *logo2opt_strerror(int code)
{
    switch (code) {
    case LOGO2OPT_OK: return "ok";
    case LOGO2OPT_EINVAL: return "invalid argument";
    case LOGO2OPT_ECOORD: return "coordinates must be integral with x/y spans below 2^22";
    case LOGO2OPT_ETOUR: return "route is not a permutation of 0..n-1";
    case LOGO2OPT_ENOMEM: return "out of memory";
    case LOGO2OPT_EDEVICE: return "HIP runtime error";
    case LOGO2OPT_EAGAIN: return "engine full (HBM budget or slots)";
    case LOGO2OPT_ESTATE: return "invalid state for this call";
    case LOGO2OPT_ENOTFOUND: return "unknown slot";
    case LOGO2OPT_EFORMAT: return "not a supported TSPLIB file";
    case LOGO2OPT_EIO: return "file could not be opened, read or written";
    case LOGO2OPT_ESIZE: return "more cities than the no-allocation entry point's fixed array holds";
    default: return "unknown error";
    }
}

int logo2opt_engine_create(int device, uint64_t hbm_budget_bytes, uint32_t max_slots, logo2opt_engine **out)
{
    return engine_create(device, hbm_budget_bytes, max_slots, out);
}

int logo2opt_tuning_shapes(int32_t *out, uint32_t cap, uint32_t *count)
{
    if (!count) return LOGO2OPT_EINVAL;
    int tmp[64][3];
    const int m = l2o::sweep_shapes(tmp, 64);
    for (int q = 0; q < m && (uint32_t)q < cap && out; ++q)
        for (int c = 0; c < 3; ++c) out[3 * q + c] = tmp[q][c];
    *count = (uint32_t)m;
    return LOGO2OPT_OK;
}

int logo2opt_tuning_set(int table, const uint32_t *max_n, const int32_t *shapes, uint32_t rows)
{
    if (table < 0 || table > 1 || !max_n || !shapes || rows == 0) return LOGO2OPT_EINVAL;
    std::vector<tune_row> v;
    for (uint32_t r = 0; r < rows; ++r) {
        sweep_cfg c{shapes[3 * r], shapes[3 * r + 1], shapes[3 * r + 2]};
        if (!l2o::sweep_cfg_supported(c)) return LOGO2OPT_EINVAL;
        v.push_back({max_n[r], c});
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_loaded[table] = true;
    install_rows_locked(table, v);
    return LOGO2OPT_OK;
}

int logo2opt_tuning_get(int table, uint32_t *max_n, int32_t *shapes, uint32_t cap, uint32_t *rows)
{
    if (table < 0 || table > 1 || !rows) return LOGO2OPT_EINVAL;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    load_tuning_locked(table);
    const auto &g = g_tune[table];
    for (uint32_t r = 0; r < g.size() && r < cap; ++r) {
        if (max_n) max_n[r] = g[r].max_n;
        if (shapes) {
            shapes[3 * r] = g[r].cfg.k;
            shapes[3 * r + 1] = g[r].cfg.t;
            shapes[3 * r + 2] = g[r].cfg.waves;
        }
    }
    *rows = (uint32_t)g.size();
    return LOGO2OPT_OK;
}

int logo2opt_tuning_load(int table, const char *path)
{
    if (table < 0 || table > 1 || !path) return LOGO2OPT_EINVAL;
    std::vector<tune_row> v;
    if (read_tuning_file(path, v) <= 0) return LOGO2OPT_EINVAL;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_loaded[table] = true;
    install_rows_locked(table, v);
    return LOGO2OPT_OK;
}

int logo2opt_engine_destroy(logo2opt_engine *e)
{
    if (!e) return LOGO2OPT_EINVAL;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    if (e->d_sitems) (void)hipFree(e->d_sitems);
    if (e->d_aitems) (void)hipFree(e->d_aitems);
    if (e->d_arena) (void)hipFree(e->d_arena);
    if (e->d_insts) (void)hipFree(e->d_insts);
    if (e->d_live) (void)hipFree(e->d_live);
    if (e->d_scratch) (void)hipFree(e->d_scratch);
    if (e->h_live) (void)hipHostFree(e->h_live);
    if (e->h_scratch) (void)hipHostFree(e->h_scratch);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    for (hipEvent_t ev : e->tev) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return LOGO2OPT_OK;
}

uint64_t logo2opt_engine_bytes_for(uint32_t n, const logo2opt_job *job) { return bytes_for(n, job); }

int logo2opt_engine_add(logo2opt_engine *e, uint32_t n, const logo2opt_city_coords *coords, const uint32_t *route,
                        const logo2opt_job *job, int32_t *slot)
{
    if (!e || !slot) return LOGO2OPT_EINVAL;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    return engine_add(e, n, coords, route, job, slot);
}

int logo2opt_engine_add_group(logo2opt_engine *e, uint32_t n, const logo2opt_city_coords *coords,
                              const uint32_t *route, const logo2opt_job *job, uint32_t chains, int32_t *first_slot)
{
    if (!e || !first_slot) return LOGO2OPT_EINVAL;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    return engine_add_group(e, n, coords, route, job, chains, first_slot);
}

int logo2opt_engine_group_result(logo2opt_engine *e, int32_t first_slot, uint32_t *route_out, logo2opt_result *res,
                                 int32_t *best_chain)
{
    if (!e) return LOGO2OPT_EINVAL;
    if (first_slot < 0 || (uint32_t)first_slot >= e->max_slots || !e->slots[first_slot].used ||
        e->slots[first_slot].group_lead != (uint32_t)first_slot)
        return LOGO2OPT_ENOTFOUND;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    const uint32_t m = e->slots[first_slot].group_size;
    std::vector<l2o_inst> h(m);
    L2O_TRY(hipMemcpyAsync(h.data(), e->d_insts + first_slot, sizeof(l2o_inst) * m, hipMemcpyDeviceToHost,
                           e->stream));
    L2O_TRY(hipStreamSynchronize(e->stream));
    uint32_t b = 0;
    for (uint32_t c = 1; c < m; ++c)
        if (h[c].have_best && (!h[b].have_best || h[c].best_length < h[b].best_length)) b = c;
    logo2opt_result r;
    int rc = engine_result(e, first_slot + (int32_t)b, route_out, &r, nullptr);
    if (rc) return rc;
    if (res) {
        *res = r;
        res->moves = res->sweeps = res->fallbacks = 0;
        res->done = 1;
        for (uint32_t c = 0; c < m; ++c) {
            res->moves += h[c].moves;
            res->sweeps += h[c].sweeps;
            res->fallbacks += h[c].fallbacks;
            res->done &= h[c].phase == L2O_DONE ? 1u : 0u;
        }
    }
    if (best_chain) *best_chain = (int32_t)b;
    return LOGO2OPT_OK;
}

int logo2opt_engine_run(logo2opt_engine *e, uint32_t max_steps, uint32_t sync_every, uint32_t *live)
{
    if (!e) return LOGO2OPT_EINVAL;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    return engine_run(e, max_steps, sync_every, live);
}

int logo2opt_engine_result(logo2opt_engine *e, int32_t slot, uint32_t *route_out, logo2opt_result *res,
                           logo2opt_best2_out *log)
{
    if (!e) return LOGO2OPT_EINVAL;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    return engine_result(e, slot, route_out, res, log);
}

int logo2opt_engine_remove(logo2opt_engine *e, int32_t slot)
{
    if (!e) return LOGO2OPT_EINVAL;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    return engine_remove(e, slot);
}

int logo2opt_engine_poll_done(logo2opt_engine *e, int32_t *out, uint32_t cap, uint32_t *count)
{
    if (!e || !count) return LOGO2OPT_EINVAL;
    uint32_t c = 0;
    auto &q = e->done_queue;
    while (c < cap && c < q.size()) {
        out[c] = q[c];
        ++c;
    }
    q.erase(q.begin(), q.begin() + c);
    *count = c;
    return LOGO2OPT_OK;
}

int logo2opt_engine_stats(logo2opt_engine *e, logo2opt_stats *st)
{
    if (!e || !st) return LOGO2OPT_EINVAL;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    std::vector<l2o_inst> h(e->max_slots);
    L2O_TRY(hipMemcpyAsync(h.data(), e->d_insts, sizeof(l2o_inst) * e->max_slots, hipMemcpyDeviceToHost, e->stream));
    L2O_TRY(hipStreamSynchronize(e->stream));
    std::memset(st, 0, sizeof(*st));
    for (uint32_t id = 0; id < e->max_slots; ++id) {
        if (!e->slots[id].used) continue;
        st->sweeps += h[id].sweeps;
        st->moves += h[id].moves;
        st->evals += h[id].sweeps * pairs_of(h[id].n);
        st->fallbacks += h[id].fallbacks;
    }
    st->wall_ms = e->wall_ms;
    st->sweep_ms = e->sweep_ms;
    st->sweep_launches = e->sweep_launches;
    return LOGO2OPT_OK;
}

int logo2opt_engine_set_timing(logo2opt_engine *e, int on)
{
    if (!e) return LOGO2OPT_EINVAL;
    e->timing = on != 0;
    drop_graph(e);
    return LOGO2OPT_OK;
}

int logo2opt_set_timing(logo2opt_ctx *ctx, int on)
{
    if (!ctx) return LOGO2OPT_EINVAL;
    return logo2opt_engine_set_timing(ctx->e, on);
}

int logo2opt_engine_set_graphs(logo2opt_engine *e, int on)
{
    if (!e) return LOGO2OPT_EINVAL;
    e->graphs = on != 0;
    drop_graph(e);
    return LOGO2OPT_OK;
}

int logo2opt_engine_set_pruning(logo2opt_engine *e, int mode)
{
    if (!e || mode < LOGO2OPT_PRUNE_OFF || mode > LOGO2OPT_PRUNE_AUTO) return LOGO2OPT_EINVAL;
    e->prune = mode;
    e->dirty = true;  // the work lists group launches by it
    drop_graph(e);
    return LOGO2OPT_OK;
}

int logo2opt_set_pruning(logo2opt_ctx *ctx, int mode)
{
    if (!ctx) return LOGO2OPT_EINVAL;
    int rc = logo2opt_engine_set_pruning(ctx->e, mode);
    if (!rc && ctx->ge) rc = logo2opt_engine_set_pruning(ctx->ge, mode);
    if (!rc) ctx->prune = mode;
    return rc;
}

// ---------------------------------------------------------------------------
// single-instance context
// ---------------------------------------------------------------------------
int logo2opt_create(int device, uint32_t n, const logo2opt_city_coords *coords, logo2opt_ctx **out)
{
    return logo2opt_create_metric(device, n, coords, LOGO2OPT_METRIC_EUC_2D, out);
}

int logo2opt_create_metric(int device, uint32_t n, const logo2opt_city_coords *coords, int metric,
                           logo2opt_ctx **out)
{
    if (!out || !coords || n < 3 || n > LOGO2OPT_MAX_N) return LOGO2OPT_EINVAL;
    if (metric < LOGO2OPT_METRIC_EUC_2D || metric > LOGO2OPT_METRIC_ATT) return LOGO2OPT_EINVAL;
    std::vector<int32_t> xy;
    float slack;
    int rc = check_coords(n, coords, xy, &slack);
    if (rc) return rc;
    logo2opt_ctx *c = new (std::nothrow) logo2opt_ctx();
    if (!c) return LOGO2OPT_ENOMEM;
    c->n = n;
    c->metric = (uint32_t)metric;
    c->coords.assign(coords, coords + n);
    // room for one ILS instance with a modest move log; ctx_load replaces the
    // engine by a larger one when a call needs more
    logo2opt_job j{};
    j.ils_rounds = 1;
    j.log_cap = 1u << 16;
    rc = engine_create(device, ctx_budget(n, j), 1, &c->e);
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return LOGO2OPT_OK;
}

int logo2opt_destroy(logo2opt_ctx *ctx)
{
    if (!ctx) return LOGO2OPT_EINVAL;
    if (ctx->e) logo2opt_engine_destroy(ctx->e);
    if (ctx->ge) logo2opt_engine_destroy(ctx->ge);
    delete ctx;
    return LOGO2OPT_OK;
}

namespace {

void ctx_accumulate(logo2opt_ctx *c)
{
    logo2opt_stats st;
    if (logo2opt_engine_stats(c->e, &st) == LOGO2OPT_OK) {
        c->sweeps += st.sweeps;
        c->moves += st.moves;
        c->evals += st.evals;
        c->fallbacks += st.fallbacks;
    }
}

int ctx_load(logo2opt_ctx *c, const uint32_t *route, logo2opt_job job)
{
    job.metric = c->metric;
    if (c->slot >= 0) {
        ctx_accumulate(c);
        int rc = engine_remove(c->e, c->slot);
        if (rc) return rc;
        c->slot = -1;
    }
    if (bytes_for(c->n, &job) > c->e->ar.size) {
        // a longer move log or more ILS rounds than the engine was sized
        // for: replace it by a larger one with the same settings
        logo2opt_engine *old = c->e, *e = nullptr;
        int rc = engine_create(old->device, ctx_budget(c->n, job), 1, &e);
        if (rc) return rc;
        e->timing = old->timing;
        e->graphs = old->graphs;
        e->prune = old->prune;
        c->sweep_ms += old->sweep_ms;
        logo2opt_engine_destroy(old);
        c->e = e;
    }
    return engine_add(c->e, c->n, c->coords.data(), route, &job, &c->slot);
}

}  // namespace

int logo2opt_step(logo2opt_ctx *ctx, const uint32_t *route, logo2opt_best2_out *out, logo2opt_process_time *t)
{
    if (!ctx || !route || !out) return LOGO2OPT_EINVAL;
    logo2opt_engine *e = ctx->e;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    const double h0 = now_ms();
    logo2opt_job job{};
    int rc = ctx_load(ctx, route, job);
    if (rc) return rc;
    e = ctx->e;  // ctx_load replaces the engine when it needs a larger one
    rc = rebuild_lists(e);
    if (rc) return rc;
    const double h1 = now_ms();
    L2O_TRY(hipEventRecord(e->ev0, e->stream));
    for (const auto &g : e->groups)
        L2O_TRY(launch_group_sweep(e, g));
    L2O_TRY(hipEventRecord(e->ev1, e->stream));
    L2O_TRY(hipEventSynchronize(e->ev1));
    float ms = 0.0f;
    L2O_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    const double d0 = now_ms();
    l2o_inst in;
    L2O_TRY(hipMemcpyAsync(&in, e->d_insts + ctx->slot, sizeof(in), hipMemcpyDeviceToHost, e->stream));
    L2O_TRY(hipStreamSynchronize(e->stream));
    const double d1 = now_ms();
    if (in.key == L2O_SENTINEL) {
        out->i = out->j = 0;
        out->minchange = 0;
    } else {
        out->minchange = (int32_t)((long long)(in.key >> 40) - L2O_KEY_GAIN_BIAS);
        out->i = (uint32_t)(in.key >> 20) & 0xFFFFFu;
        out->j = (uint32_t)in.key & 0xFFFFFu;
    }
    ctx->sweep_ms += ms;
    if (t) {
        t->HtD_time = (h1 - h0) * 1e-3;
        t->kernel_time = ms * 1e-3;
        t->DtH_time = (d1 - d0) * 1e-3;
        t->out = *out;
    }
    return LOGO2OPT_OK;
}

int logo2opt_optimize(logo2opt_ctx *ctx, uint32_t *route_inout, uint64_t *length, uint64_t *moves,
                      logo2opt_best2_out *log, uint64_t log_cap, uint32_t sync_every)
{
    if (!ctx || !route_inout) return LOGO2OPT_EINVAL;
    logo2opt_engine *e = ctx->e;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    if (log && log_cap > 0xFFFFFFFFull) return LOGO2OPT_EINVAL;
    logo2opt_job job{};
    job.log_cap = log ? (uint32_t)log_cap : 0u;
    int rc = ctx_load(ctx, route_inout, job);
    if (rc) return rc;
    e = ctx->e;  // ctx_load replaces the engine when it needs a larger one
    const double t0 = now_ms();
    uint32_t live = 1;
    rc = engine_run(e, 0xFFFFFFFFu, sync_every, &live);
    if (rc) return rc;
    ctx->wall_ms += now_ms() - t0;
    logo2opt_result res;
    rc = engine_result(e, ctx->slot, route_inout, &res, log);
    if (rc) return rc;
    if (length) *length = res.length;
    if (moves) *moves = res.moves;
    return LOGO2OPT_OK;
}

int logo2opt_ils(logo2opt_ctx *ctx, uint32_t *route_inout, uint64_t seed, uint32_t rounds, int kick,
                 uint64_t *best_length, uint64_t *round_lengths)
{
    if (!ctx || !route_inout || (kick != 0 && kick != 1)) return LOGO2OPT_EINVAL;
    logo2opt_engine *e = ctx->e;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    logo2opt_job job{};
    job.ils_rounds = rounds;
    job.kick = (uint32_t)kick;
    job.seed = seed;
    if (rounds == 0) {
        // a plain local search
        uint64_t L = 0;
        int rc = logo2opt_optimize(ctx, route_inout, &L, nullptr, nullptr, 0, 64);
        if (best_length) *best_length = L;
        if (round_lengths) round_lengths[0] = L;
        return rc;
    }
    int rc = ctx_load(ctx, route_inout, job);
    if (rc) return rc;
    e = ctx->e;  // ctx_load replaces the engine when it needs a larger one
    const double t0 = now_ms();
    uint32_t live = 1;
    rc = engine_run(e, 0xFFFFFFFFu, 64, &live);
    if (rc) return rc;
    ctx->wall_ms += now_ms() - t0;
    logo2opt_result res;
    rc = engine_result(e, ctx->slot, route_inout, &res, nullptr);
    if (rc) return rc;
    if (best_length) *best_length = res.length;
    if (round_lengths) {
        rc = logo2opt_engine_round_lengths(e, ctx->slot, round_lengths, rounds + 1);
        if (rc) return rc;
    }
    return LOGO2OPT_OK;
}

uint64_t logo2opt_stop_below(uint64_t solution, float err)
{
    // the reference's loop goes on while solution * (1.0f + error) < best,
    // in float (ILSGlobalSolverSequential.cpp:98); the largest best that ends
    // it, found by bisection (the float conversion is monotone)
    const float thr = (float)solution * (1.0f + err);
    auto ends = [&](uint64_t L) { return !(thr < (float)L); };
    if (!ends(0)) return 0;  // unreachable target (thr is negative or NaN)
    if (ends(UINT64_MAX)) return UINT64_MAX;  // every length ends it
    uint64_t lo = 0, hi = UINT64_MAX;  // ends(lo), !ends(hi)
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (ends(mid)) lo = mid;
        else hi = mid;
    }
    return lo + 1;
}

int logo2opt_ils_until(logo2opt_ctx *ctx, uint32_t *route_inout, uint64_t seed, uint32_t max_rounds, int kick,
                       uint64_t solution, float err, double time_limit_s, uint64_t *best_length,
                       uint64_t *round_lengths, uint32_t *rounds_done)
{
    if (!ctx || !route_inout || (kick != 0 && kick != 1) || max_rounds == 0 || !(time_limit_s >= 0.0))
        return LOGO2OPT_EINVAL;
    logo2opt_engine *e = ctx->e;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    const double t_start = now_ms();
    logo2opt_job job{};
    job.ils_rounds = max_rounds;
    job.kick = (uint32_t)kick;
    job.seed = seed;
    job.stop_below = logo2opt_stop_below(solution, err);
    int rc = ctx_load(ctx, route_inout, job);
    if (rc) return rc;
    e = ctx->e;  // ctx_load replaces the engine when it needs a larger one
    const double t0 = now_ms();
    uint32_t live = 1;
    bool capped = false;
    while (live) {
        rc = engine_run(e, 64, 64, &live);
        if (rc) return rc;
        if (live && !capped && time_limit_s > 0.0 && (now_ms() - t_start) * 1e-3 >= time_limit_s) {
            // the clock has run out: the round in progress is the last one
            l2o_inst in;
            L2O_TRY(hipMemcpyAsync(&in, e->d_insts + ctx->slot, sizeof(in), hipMemcpyDeviceToHost, e->stream));
            L2O_TRY(hipStreamSynchronize(e->stream));
            const uint32_t cap = std::min(in.ils_rounds, in.rounds_done);
            L2O_TRY(hipMemcpyAsync((char *)(e->d_insts + ctx->slot) + offsetof(l2o_inst, ils_rounds), &cap,
                                   sizeof(cap), hipMemcpyHostToDevice, e->stream));
            L2O_TRY(hipStreamSynchronize(e->stream));
            capped = true;
        }
    }
    ctx->wall_ms += now_ms() - t0;
    logo2opt_result res;
    rc = engine_result(e, ctx->slot, route_inout, &res, nullptr);
    if (rc) return rc;
    if (best_length) *best_length = res.length;
    if (rounds_done) *rounds_done = res.rounds_done;
    if (round_lengths) {
        rc = logo2opt_engine_round_lengths(e, ctx->slot, round_lengths, max_rounds + 1);
        if (rc) return rc;
    }
    return LOGO2OPT_OK;
}

int logo2opt_ils_multistart(logo2opt_ctx *ctx, uint32_t *route_inout, uint64_t seed, uint32_t rounds,
                            uint32_t chains, int kick, uint64_t *best_length, uint64_t *round_lengths)
{
    if (!ctx || !route_inout || (kick != 0 && kick != 1) || rounds == 0 || chains == 0 || chains > 4096)
        return LOGO2OPT_EINVAL;
    if (hipSetDevice(ctx->e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    logo2opt_job job{};
    job.ils_rounds = rounds;
    job.kick = (uint32_t)kick;
    job.seed = seed;
    job.metric = ctx->metric;
    int rc;
    if (ctx->ge && (ctx->ge_chains < chains || ctx->ge->ar.size < chains * bytes_for(ctx->n, &job))) {
        logo2opt_engine_destroy(ctx->ge);
        ctx->ge = nullptr;
    }
    if (!ctx->ge) {
        rc = engine_create(ctx->e->device, chains * bytes_for(ctx->n, &job) + (1ull << 24), chains, &ctx->ge);
        if (rc) return rc;
        ctx->ge_chains = chains;
        ctx->ge->prune = ctx->prune;
    }
    logo2opt_engine *e = ctx->ge;
    for (uint32_t k = 0; k < e->max_slots; ++k)
        if (e->slots[k].used) {
            rc = engine_remove(e, (int32_t)k);
            if (rc) return rc;
        }
    int32_t first = -1;
    rc = engine_add_group(e, ctx->n, ctx->coords.data(), route_inout, &job, chains, &first);
    if (rc) return rc;
    const double t0 = now_ms();
    uint32_t live = 1;
    rc = engine_run(e, 0xFFFFFFFFu, 64, &live);
    if (rc) return rc;
    ctx->wall_ms += now_ms() - t0;
    logo2opt_result res;
    rc = logo2opt_engine_group_result(e, first, route_inout, &res, nullptr);
    if (rc) return rc;
    ctx->sweeps += res.sweeps;
    ctx->moves += res.moves;
    ctx->fallbacks += res.fallbacks;
    if (best_length) *best_length = res.length;
    if (round_lengths) rc = logo2opt_engine_round_lengths(e, first, round_lengths, rounds + 1);
    return rc;
}

int logo2opt_engine_round_lengths(logo2opt_engine *e, int32_t slot, uint64_t *out, uint32_t cap)
{
    if (!e || !out) return LOGO2OPT_EINVAL;
    if (slot < 0 || (uint32_t)slot >= e->max_slots || !e->slots[slot].used) return LOGO2OPT_ENOTFOUND;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    l2o_inst in;
    L2O_TRY(hipMemcpyAsync(&in, e->d_insts + slot, sizeof(in), hipMemcpyDeviceToHost, e->stream));
    L2O_TRY(hipStreamSynchronize(e->stream));
    if (!in.rlog) return LOGO2OPT_ESTATE;
    const uint32_t m = std::min<uint32_t>(cap, in.rounds_done + (in.phase == L2O_DONE ? 1u : 0u));
    if (m) {
        L2O_TRY(hipMemcpyAsync(out, in.rlog, 8ull * m, hipMemcpyDeviceToHost, e->stream));
        L2O_TRY(hipStreamSynchronize(e->stream));
    }
    return LOGO2OPT_OK;
}

int logo2opt_route_length(logo2opt_ctx *ctx, const uint32_t *route, uint64_t *length)
{
    if (!ctx || !route || !length) return LOGO2OPT_EINVAL;
    logo2opt_engine *e = ctx->e;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    logo2opt_job job{};
    int rc = ctx_load(ctx, route, job);
    if (rc) return rc;
    e = ctx->e;  // ctx_load replaces the engine when it needs a larger one
    L2O_TRY(hipMemsetAsync(e->d_scratch, 0, sizeof(unsigned long long), e->stream));
    L2O_TRY(l2o::launch_length(e->d_insts, (uint32_t)ctx->slot, ctx->n, e->d_scratch, e->stream));
    L2O_TRY(hipMemcpyAsync(e->h_scratch, e->d_scratch, sizeof(unsigned long long), hipMemcpyDeviceToHost, e->stream));
    L2O_TRY(hipStreamSynchronize(e->stream));
    *length = e->h_scratch[0];
    return LOGO2OPT_OK;
}

int logo2opt_get_stats(logo2opt_ctx *ctx, logo2opt_stats *st)
{
    if (!ctx || !st) return LOGO2OPT_EINVAL;
    logo2opt_stats cur;
    std::memset(&cur, 0, sizeof(cur));
    if (ctx->slot >= 0) {
        int rc = logo2opt_engine_stats(ctx->e, &cur);
        if (rc) return rc;
    }
    st->sweeps = ctx->sweeps + cur.sweeps;
    st->moves = ctx->moves + cur.moves;
    st->evals = ctx->evals + cur.evals;
    st->fallbacks = ctx->fallbacks + cur.fallbacks;
    st->sweep_ms = ctx->sweep_ms + ctx->e->sweep_ms;
    st->wall_ms = ctx->wall_ms;
    st->sweep_launches = ctx->e->sweep_launches;
    return LOGO2OPT_OK;
}

int logo2opt_engine_time_sweep(logo2opt_engine *e, uint32_t sweeps, double *avg_ms)
{
    if (!e || !avg_ms || sweeps == 0) return LOGO2OPT_EINVAL;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    int rc = rebuild_lists(e);
    if (rc) return rc;
    for (const auto &g : e->groups)  // untimed first launch
        L2O_TRY(launch_group_sweep(e, g));
    L2O_TRY(hipEventRecord(e->ev0, e->stream));
    for (uint32_t s = 0; s < sweeps; ++s)
        for (const auto &g : e->groups)
            L2O_TRY(launch_group_sweep(e, g));
    L2O_TRY(hipEventRecord(e->ev1, e->stream));
    L2O_TRY(hipEventSynchronize(e->ev1));
    float ms = 0.0f;
    L2O_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    *avg_ms = (double)ms / sweeps;
    return LOGO2OPT_OK;
}

int logo2opt_sqrt_ulp_check(int device, uint32_t lo_bits, uint32_t hi_bits, uint32_t *max_ulp)
{
    if (!max_ulp || lo_bits >= hi_bits) return LOGO2OPT_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return LOGO2OPT_EDEVICE;
    unsigned *d = nullptr;
    L2O_TRY(hipMalloc(&d, sizeof(unsigned)));
    hipError_t err = hipMemset(d, 0, sizeof(unsigned));
    if (err == hipSuccess) err = l2o::launch_sqrt_ulp(lo_bits, hi_bits, d, nullptr);
    unsigned h = 0;
    if (err == hipSuccess) err = hipMemcpy(&h, d, sizeof(unsigned), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    L2O_TRY(err);
    *max_ulp = h;
    return LOGO2OPT_OK;
}

int logo2opt_time_sweep(logo2opt_ctx *ctx, const uint32_t *route, uint32_t sweeps, double *avg_ms)
{
    if (!ctx || !route || !avg_ms || sweeps == 0) return LOGO2OPT_EINVAL;
    logo2opt_engine *e = ctx->e;
    if (hipSetDevice(e->device) != hipSuccess) return LOGO2OPT_EDEVICE;
    logo2opt_job job{};
    int rc = ctx_load(ctx, route, job);
    if (rc) return rc;
    e = ctx->e;  // ctx_load replaces the engine when it needs a larger one
    rc = rebuild_lists(e);
    if (rc) return rc;
    // one untimed launch (code object load, caches)
    for (const auto &g : e->groups)
        L2O_TRY(launch_group_sweep(e, g));
    L2O_TRY(hipEventRecord(e->ev0, e->stream));
    for (uint32_t s = 0; s < sweeps; ++s)
        for (const auto &g : e->groups)
            L2O_TRY(launch_group_sweep(e, g));
    L2O_TRY(hipEventRecord(e->ev1, e->stream));
    L2O_TRY(hipEventSynchronize(e->ev1));
    float ms = 0.0f;
    L2O_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    *avg_ms = (double)ms / sweeps;
    ctx->sweep_ms += ms;
    return LOGO2OPT_OK;
}

}  // extern "C"
