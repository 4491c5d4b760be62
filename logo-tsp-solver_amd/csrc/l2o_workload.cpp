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

// Greedy nearest neighbour on a uniform bucket grid.  From the current city
// the rings of cells around it are searched outward; ring R cannot hold a
// point closer than (R-1) cell widths, so the search stops once that bound
// exceeds the best squared distance found (ties: lowest city id, so equal
// distances in farther rings are still examined).
int logo2opt_nn_tour(uint32_t n, const int32_t *xy, uint32_t start, uint32_t *route)
{
    if (n == 0 || !xy || !route || start >= n) return LOGO2OPT_EINVAL;
    int64_t mnx = xy[0], mny = xy[1], mxx = xy[0], mxy = xy[1];
    for (uint32_t k = 1; k < n; ++k) {
        mnx = std::min<int64_t>(mnx, xy[2 * k]);
        mxx = std::max<int64_t>(mxx, xy[2 * k]);
        mny = std::min<int64_t>(mny, xy[2 * k + 1]);
        mxy = std::max<int64_t>(mxy, xy[2 * k + 1]);
    }
    const int64_t span = std::max<int64_t>(1, std::max(mxx - mnx, mxy - mny) + 1);
    int64_t G = (int64_t)std::ceil(std::sqrt((double)n / 2.0));
    G = std::max<int64_t>(1, std::min<int64_t>(G, 4096));
    const int64_t cw = (span + G - 1) / G;  // cell width, G * cw >= span
    auto cell_of = [&](uint32_t c, int64_t *gx, int64_t *gy) {
        *gx = (xy[2 * c] - mnx) / cw;
        *gy = (xy[2 * c + 1] - mny) / cw;
    };
    std::vector<std::vector<uint32_t>> cells((size_t)(G * G));
    std::vector<uint32_t> pos(n);
    for (uint32_t c = 0; c < n; ++c) {
        int64_t gx, gy;
        cell_of(c, &gx, &gy);
        auto &v = cells[(size_t)(gy * G + gx)];
        pos[c] = (uint32_t)v.size();
        v.push_back(c);
    }
    auto remove = [&](uint32_t c) {
        int64_t gx, gy;
        cell_of(c, &gx, &gy);
        auto &v = cells[(size_t)(gy * G + gx)];
        const uint32_t p = pos[c], last = v.back();
        v[p] = last;
        pos[last] = p;
        v.pop_back();
    };
    uint32_t cur = start;
    remove(cur);
    route[0] = cur;
    for (uint32_t k = 1; k < n; ++k) {
        int64_t gx, gy;
        cell_of(cur, &gx, &gy);
        const int64_t px = xy[2 * cur], py = xy[2 * cur + 1];
        uint64_t best = UINT64_MAX;
        uint32_t arg = UINT32_MAX;
        for (int64_t R = 0; R < G; ++R) {
            if (R >= 2 && best != UINT64_MAX) {
                const uint64_t lb = (uint64_t)((R - 1) * cw) * (uint64_t)((R - 1) * cw);
                if (lb > best) break;
            }
            const int64_t x0 = gx - R, x1 = gx + R, y0 = gy - R, y1 = gy + R;
            for (int64_t y = std::max<int64_t>(0, y0); y <= std::min<int64_t>(G - 1, y1); ++y) {
                const bool edge_row = (y == y0 || y == y1);
                for (int64_t x = std::max<int64_t>(0, x0); x <= std::min<int64_t>(G - 1, x1); ++x) {
                    if (!edge_row && x != x0 && x != x1) continue;
                    for (uint32_t u : cells[(size_t)(y * G + x)]) {
                        const int64_t dx = xy[2 * u] - px, dy = xy[2 * u + 1] - py;
                        const uint64_t s = (uint64_t)(dx * dx + dy * dy);
                        if (s < best || (s == best && u < arg)) {
                            best = s;
                            arg = u;
                        }
                    }
                }
            }
        }
        remove(arg);
        route[k] = arg;
        cur = arg;
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
    std::priority_queue<cand, std::vector<cand>, std::greater<cand>> pq;
    auto query = [&](uint32_t x) {
        const uint32_t rx = find(x);
        uint64_t s;
        const uint32_t y = g.nearest(x, [&](uint32_t u) { return find(u) != rx; }, &s);
        if (y != UINT32_MAX) pq.push(cand{s, std::min(x, y), std::max(x, y), x, y});
    };
    for (uint32_t x = 0; x < n; ++x) query(x);
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
