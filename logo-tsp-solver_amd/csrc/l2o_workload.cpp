// l2o_workload.cpp -- deterministic synthetic workloads and the greedy
// nearest-neighbour and multiple-fragment constructions (host C++).
//
// The reference builds its inputs from TSPLIB files and time-seeded libc
// rand() (construction.c:118-150, common.c:253-259); the benchmark configs
// here are synthetic, so every stream is a SplitMix64-seeded xoshiro256**
// and results are reproducible across runs and machines with the same libm.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <numeric>
#include <queue>
#include <thread>
#include <vector>

#include "../../include/logo2opt.h"

namespace {

struct xoshiro {
    uint64_t s[4];
    explicit xoshiro(uint64_t seed)
    {
        uint64_t x = seed;
        for (auto &w : s) {
            uint64_t z = (x += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            w = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
};

int32_t round_clamp(double v, uint32_t side)
{
    const double r = std::nearbyint(v);
    if (r < 0.0) return 0;
    if (r > (double)(side - 1)) return (int32_t)(side - 1);
    return (int32_t)r;
}

// Points bucketed on a uniform G x G grid (about `per_cell` points per cell)
// with O(1) removal, for exact nearest-point queries.
struct point_grid {
    const int32_t *xy = nullptr;
    int64_t mnx = 0, mny = 0, cw = 1, G = 1;
    std::vector<std::vector<uint32_t>> cells;
    std::vector<uint32_t> pos;

    void build(uint32_t n, const int32_t *p, double per_cell)
    {
        xy = p;
        int64_t mxx = p[0], mxy = p[1];
        mnx = p[0];
        mny = p[1];
        for (uint32_t k = 1; k < n; ++k) {
            mnx = std::min<int64_t>(mnx, p[2 * k]);
            mxx = std::max<int64_t>(mxx, p[2 * k]);
            mny = std::min<int64_t>(mny, p[2 * k + 1]);
            mxy = std::max<int64_t>(mxy, p[2 * k + 1]);
        }
        const int64_t span = std::max<int64_t>(1, std::max(mxx - mnx, mxy - mny) + 1);
        G = (int64_t)std::ceil(std::sqrt((double)n / per_cell));
        G = std::max<int64_t>(1, std::min<int64_t>(G, 4096));
        cw = (span + G - 1) / G;  // G * cw >= span
        cells.assign((size_t)(G * G), {});
        pos.assign(n, 0);
        for (uint32_t c = 0; c < n; ++c) {
            auto &v = cells[cell(c)];
            pos[c] = (uint32_t)v.size();
            v.push_back(c);
        }
    }
    size_t cell(uint32_t c) const
    {
        return (size_t)(((xy[2 * c + 1] - mny) / cw) * G + (xy[2 * c] - mnx) / cw);
    }
    void remove(uint32_t c)
    {
        auto &v = cells[cell(c)];
        const uint32_t p = pos[c], last = v.back();
        v[p] = last;
        pos[last] = p;
        v.pop_back();
    }
    // The point u != c in the grid with ok(u) that minimises (squared
    // distance, u); UINT32_MAX if there is none.  Rings of cells are searched
    // outward: ring R holds no point closer than (R-1) cell widths, so the
    // search ends once that bound exceeds the best distance (equal distances
    // in farther rings are still examined for the id tie-break).
    uint32_t nearest(uint32_t c, const std::function<bool(uint32_t)> &ok, uint64_t *s_out) const
    {
        const int64_t px = xy[2 * c], py = xy[2 * c + 1];
        const int64_t gx = (px - mnx) / cw, gy = (py - mny) / cw;
        uint64_t best = UINT64_MAX;
        uint32_t arg = UINT32_MAX;
        for (int64_t R = 0; R < G; ++R) {
            if (R >= 2 && best != UINT64_MAX) {
                const uint64_t lb = (uint64_t)((R - 1) * cw) * (uint64_t)((R - 1) * cw);
                if (lb > best) break;
            }
            const int64_t y0 = std::max<int64_t>(0, gy - R), y1 = std::min<int64_t>(G - 1, gy + R);
            for (int64_t y = y0; y <= y1; ++y) {
                const bool full = (y == gy - R || y == gy + R);
                const int64_t step = full ? 1 : 2 * R;
                for (int64_t x = gx - R; x <= gx + R; x += (step > 0 ? step : 1)) {
                    if (x < 0 || x >= G) continue;
                    for (uint32_t u : cells[(size_t)(y * G + x)]) {
                        if (u == c) continue;
                        const int64_t dx = xy[2 * u] - px, dy = xy[2 * u + 1] - py;
                        const uint64_t s = (uint64_t)(dx * dx + dy * dy);
                        if ((s < best || (s == best && u < arg)) && ok(u)) {
                            best = s;
                            arg = u;
                        }
                    }
                }
            }
        }
        *s_out = best;
        return arg;
    }

    // The k points u != c that come first in (squared distance, u) order,
    // ascending, into ids / ss; returns how many (fewer than k only when the
    // grid holds fewer points).  Same ring bound as nearest(): once the k-th
    // best distance is below the ring's lower bound, no later point can enter.
    int knn(uint32_t c, int k, uint32_t *ids, uint64_t *ss) const
    {
        const int64_t px = xy[2 * c], py = xy[2 * c + 1];
        const int64_t gx = (px - mnx) / cw, gy = (py - mny) / cw;
        int cnt = 0;
        for (int64_t R = 0; R < G; ++R) {
            if (R >= 2 && cnt == k) {
                const uint64_t lb = (uint64_t)((R - 1) * cw) * (uint64_t)((R - 1) * cw);
                if (lb > ss[k - 1]) break;
            }
            const int64_t y0 = std::max<int64_t>(0, gy - R), y1 = std::min<int64_t>(G - 1, gy + R);
            for (int64_t y = y0; y <= y1; ++y) {
                const bool full = (y == gy - R || y == gy + R);
                const int64_t step = full ? 1 : 2 * R;
                for (int64_t x = gx - R; x <= gx + R; x += (step > 0 ? step : 1)) {
                    if (x < 0 || x >= G) continue;
                    for (uint32_t u : cells[(size_t)(y * G + x)]) {
                        if (u == c) continue;
                        const int64_t dx = xy[2 * u] - px, dy = xy[2 * u + 1] - py;
                        const uint64_t s = (uint64_t)(dx * dx + dy * dy);
                        if (cnt == k && (s > ss[k - 1] || (s == ss[k - 1] && u > ids[k - 1]))) continue;
                        int q = cnt < k ? cnt++ : k - 1;  // insertion into the sorted list
                        while (q > 0 && (ss[q - 1] > s || (ss[q - 1] == s && ids[q - 1] > u))) {
                            ss[q] = ss[q - 1];
                            ids[q] = ids[q - 1];
                            --q;
                        }
                        ss[q] = s;
                        ids[q] = u;
                    }
                }
            }
        }
        return cnt;
    }
};

// Each city's K nearest other cities in (squared distance, id) order,
// computed on several host threads (the queries only read the grid).  The
// greedy constructions below take their next city from these lists: the
// first entry that is still allowed is the allowed city of smallest
// (distance, id), because every city outside the list comes later in that
// order than the whole list; only when no entry is allowed any more do they
// search the grid.
constexpr int L2O_KNN = 10;

struct knn_lists {
    std::vector<uint32_t> ids;  // n x L2O_KNN
    std::vector<uint8_t> cnt;

    void build(const point_grid &g, uint32_t n)
    {
        ids.assign((size_t)n * L2O_KNN, UINT32_MAX);
        cnt.assign(n, 0);
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const uint32_t threads = std::min<uint32_t>(std::min(16u, hw), n / 2048 + 1);
        auto work = [&](uint32_t lo, uint32_t hi) {
            uint64_t ss[L2O_KNN];
            for (uint32_t c = lo; c < hi; ++c) cnt[c] = (uint8_t)g.knn(c, L2O_KNN, &ids[(size_t)c * L2O_KNN], ss);
        };
        if (threads <= 1) {
            work(0, n);
            return;
        }
        std::vector<std::thread> pool;
        for (uint32_t t = 0; t < threads; ++t)
            pool.emplace_back(work, (uint32_t)((uint64_t)n * t / threads), (uint32_t)((uint64_t)n * (t + 1) / threads));
        for (auto &th : pool) th.join();
    }
    // the first city u of c's list with ok(u), or UINT32_MAX
    template <class F>
    uint32_t first(uint32_t c, F ok) const
    {
        const uint32_t *l = &ids[(size_t)c * L2O_KNN];
        for (int q = 0; q < cnt[c]; ++q)
            if (ok(l[q])) return l[q];
        return UINT32_MAX;
    }
};

}  // namespace

extern "C" {

void logo2opt_gen_uniform(uint32_t n, uint32_t side, uint64_t seed, int32_t *xy)
{
    xoshiro g(seed);
    for (uint32_t k = 0; k < 2 * n; ++k) xy[k] = (int32_t)(g.next() % side);
}

// Clustered instances: centres uniform in [0,side)^2; the first
// n - round(noise_frac*n) points are dealt round-robin to the centres with a
// Box-Muller N(0, (sigma_frac*side)^2) offset per axis, rounded and clamped;
// the remaining points are uniform noise.
void logo2opt_gen_clustered(uint32_t n, uint32_t side, uint32_t nclusters, double sigma_frac, double noise_frac,
                            uint64_t seed, int32_t *xy)
{
    xoshiro g(seed);
    std::vector<double> cx(nclusters), cy(nclusters);
    for (uint32_t c = 0; c < nclusters; ++c) {
        cx[c] = g.u01() * (double)side;
        cy[c] = g.u01() * (double)side;
    }
    uint32_t noise = (uint32_t)std::nearbyint(noise_frac * (double)n);
    noise = std::min(noise, n);
    const uint32_t clustered = n - noise;
    const double sigma = sigma_frac * (double)side;
    const double tau = 6.283185307179586476925286766559;
    for (uint32_t k = 0; k < clustered; ++k) {
        const uint32_t c = k % nclusters;
        const double u1 = ((double)(g.next() >> 11) + 1.0) * 0x1.0p-53;
        const double u2 = g.u01();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double a = tau * u2;
        xy[2 * k] = round_clamp(cx[c] + sigma * (r * std::cos(a)), side);
        xy[2 * k + 1] = round_clamp(cy[c] + sigma * (r * std::sin(a)), side);
    }
    for (uint32_t k = clustered; k < n; ++k) {
        xy[2 * k] = (int32_t)(g.next() % side);
        xy[2 * k + 1] = (int32_t)(g.next() % side);
    }
}

void logo2opt_random_perm(uint32_t n, uint64_t seed, uint32_t *route)
{
    xoshiro g(seed);
    for (uint32_t k = 0; k < n; ++k) route[k] = k;
    for (uint32_t k = n - 1; k >= 1; --k) {
        const uint32_t j = (uint32_t)(g.next() % ((uint64_t)k + 1));
        std::swap(route[k], route[j]);
    }
}

void logo2opt_xy_to_coords(uint32_t n, const int32_t *xy, logo2opt_city_coords *coords)
{
    for (uint32_t k = 0; k < n; ++k) {
        coords[k].x = (float)xy[2 * k];
        coords[k].y = (float)xy[2 * k + 1];
    }
}

// Greedy nearest neighbour: from the current city the next one is the
// unvisited city of smallest (squared distance, id) -- lowest city id on
// ties.  It comes from the current city's nearest-city list when one of its
// entries is unvisited; otherwise from the grid of unvisited cities, whose
// rings of cells are searched outward until the ring's lower bound exceeds the
// best distance found (equal distances in farther rings are still examined).
int logo2opt_nn_tour(uint32_t n, const int32_t *xy, uint32_t start, uint32_t *route)
{
    if (n == 0 || !xy || !route || start >= n) return LOGO2OPT_EINVAL;
    point_grid g;
    g.build(n, xy, 2.0);
    knn_lists nl;
    nl.build(g, n);
    std::vector<uint8_t> visited(n, 0);
    uint32_t cur = start;
    visited[cur] = 1;
    g.remove(cur);
    route[0] = cur;
    for (uint32_t k = 1; k < n; ++k) {
        uint32_t u = nl.first(cur, [&](uint32_t v) { return !visited[v]; });
        if (u == UINT32_MAX) {
            uint64_t s;
            u = g.nearest(cur, [](uint32_t) { return true; }, &s);  // the grid holds the unvisited cities
        }
        visited[u] = 1;
        g.remove(u);
        route[k] = u;
        cur = u;
    }
    return LOGO2OPT_OK;
}

// Multiple-fragment (greedy edge) construction, the reference's default
// initial tour (defs.h:70 "mf"; multipleFragment_KD, kdtree_utils.c:56-248):
// edges are added in increasing order of (squared length, lower id, higher id)
// whenever both ends have degree < 2 and they lie in different fragments,
// until one Hamiltonian path remains, which is closed.  As in the reference,
// each city with degree < 2 keeps its nearest valid neighbour in a priority
// queue and is re-queried when that candidate has become invalid (a valid
// edge's key can only grow, so the queue's top, once valid, is the smallest
// valid edge).  The reference keeps the candidates in a kd-tree over float
// coordinates that drops duplicate points (kdtree_utils.c:42-48); here a grid
// over the integer coordinates answers exact queries, so the result is the
// greedy-edge tour for the (squared Euclidean, id) order, duplicates
// included.  The squared length orders EUC_2D, CEIL_2D and ATT edges alike
// (all are non-decreasing in it).  route[0] = 0, route[1] = the lower-id
// neighbour of city 0.
int logo2opt_mf_tour(uint32_t n, const int32_t *xy, uint32_t *route)
{
    if (n < 3 || n > LOGO2OPT_MAX_N || !xy || !route) return LOGO2OPT_EINVAL;
    point_grid g;
    g.build(n, xy, 2.0);
    std::vector<uint8_t> deg(n, 0);
    std::vector<uint32_t> uf(n), adj(2 * (size_t)n, UINT32_MAX);
    std::iota(uf.begin(), uf.end(), 0u);
    auto find = [&](uint32_t v) {
        while (uf[v] != v) {
            uf[v] = uf[uf[v]];
            v = uf[v];
        }
        return v;
    };
    struct cand {
        uint64_t s;
        uint32_t lo, hi, x, y;  // edge (lo, hi) = {x, y}, found for x
        bool operator>(const cand &o) const
        {
            if (s != o.s) return s > o.s;
            if (lo != o.lo) return lo > o.lo;
            if (hi != o.hi) return hi > o.hi;
            return x > o.x;
        }
    };
    knn_lists nl;
    nl.build(g, n);
    auto sq = [&](uint32_t a, uint32_t b) {
        const int64_t dx = (int64_t)xy[2 * a] - xy[2 * b], dy = (int64_t)xy[2 * a + 1] - xy[2 * b + 1];
        return (uint64_t)(dx * dx + dy * dy);
    };
    std::priority_queue<cand, std::vector<cand>, std::greater<cand>> pq;
    // x's nearest valid partner: degree < 2 (the grid holds only those) and
    // in another fragment
    auto query = [&](uint32_t x) {
        const uint32_t rx = find(x);
        uint32_t y = nl.first(x, [&](uint32_t u) { return deg[u] < 2 && find(u) != rx; });
        uint64_t s;
        if (y != UINT32_MAX) s = sq(x, y);
        else y = g.nearest(x, [&](uint32_t u) { return find(u) != rx; }, &s);
        if (y != UINT32_MAX) pq.push(cand{s, std::min(x, y), std::max(x, y), x, y});
    };
    {
        // every city starts as its own fragment of degree 0: its first
        // candidate is its nearest city
        std::vector<cand> init;
        init.reserve(n);
        for (uint32_t x = 0; x < n; ++x) {
            const uint32_t y = nl.cnt[x] ? nl.ids[(size_t)x * L2O_KNN] : UINT32_MAX;
            if (y != UINT32_MAX) init.push_back(cand{sq(x, y), std::min(x, y), std::max(x, y), x, y});
        }
        pq = std::priority_queue<cand, std::vector<cand>, std::greater<cand>>(std::greater<cand>(), std::move(init));
    }
    uint32_t added = 0;
    while (added + 1 < n) {
        if (pq.empty()) return LOGO2OPT_ESTATE;  // cannot happen while two fragments remain
        const cand c = pq.top();
        pq.pop();
        if (deg[c.x] == 2) continue;
        if (deg[c.y] < 2 && find(c.x) != find(c.y)) {
            for (uint32_t v : {c.x, c.y}) {
                adj[2 * (size_t)v + deg[v]] = v == c.x ? c.y : c.x;
                if (++deg[v] == 2) g.remove(v);
            }
            uf[find(c.x)] = find(c.y);
            ++added;
            if (deg[c.x] < 2) query(c.x);
        } else {
            query(c.x);  // stale candidate: find x's nearest valid neighbour again
        }
    }
    // close the path: its two ends have degree 1
    uint32_t e[2], ne = 0;
    for (uint32_t v = 0; v < n && ne < 2; ++v)
        if (deg[v] < 2) e[ne++] = v;
    if (ne != 2) return LOGO2OPT_ESTATE;
    for (uint32_t v : {e[0], e[1]}) adj[2 * (size_t)v + deg[v]++] = v == e[0] ? e[1] : e[0];
    uint32_t prev = 0, cur = std::min(adj[0], adj[1]);
    route[0] = 0;
    for (uint32_t k = 1; k < n; ++k) {
        route[k] = cur;
        const uint32_t nx = adj[2 * (size_t)cur] == prev ? adj[2 * (size_t)cur + 1] : adj[2 * (size_t)cur];
        prev = cur;
        cur = nx;
    }
    return cur == 0 ? LOGO2OPT_OK : LOGO2OPT_ESTATE;
}

}  // extern "C"
