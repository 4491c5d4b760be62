/*
 * logo_oracle.c -- CPU restatement of LOGO's 2-opt local-search path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker; never linked by the product).
 * Each function cites the reference file:line whose behaviour it restates.
 * Arithmetic is exact integer EUC_2D; see logo_oracle.h for what differs
 * from the reference on purpose.
 */
#include "logo_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* PRNG (build-defined; the reference seeds libc rand() from the clock,      */
/* common.c:253-259)                                                          */
/* ------------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void orc_rng_seed(orc_rng_t *r, uint64_t seed)
{
    uint64_t x = seed;
    for (int k = 0; k < 4; k++) r->s[k] = splitmix64(&x);
}

static inline uint64_t rotl64(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }

uint64_t orc_rng_next(orc_rng_t *r)
{
    uint64_t *s = r->s;
    uint64_t result = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

double orc_rng_u01(orc_rng_t *r) { return (double)(orc_rng_next(r) >> 11) * 0x1.0p-53; }

/* ------------------------------------------------------------------------ */
/* Workloads                                                                  */
/* ------------------------------------------------------------------------ */
void orc_gen_uniform(uint32_t n, uint32_t side, uint64_t seed, int32_t *xy)
{
    orc_rng_t r;
    orc_rng_seed(&r, seed);
    for (uint32_t k = 0; k < n; k++) {
        xy[2 * k] = (int32_t)(orc_rng_next(&r) % side);
        xy[2 * k + 1] = (int32_t)(orc_rng_next(&r) % side);
    }
}

static int32_t clamp_coord(double v, uint32_t side)
{
    double r = nearbyint(v); /* round half to even, default FP environment */
    if (r < 0.0) return 0;
    if (r > (double)(side - 1)) return (int32_t)(side - 1);
    return (int32_t)r;
}

/* Clustered instance: nclusters centres uniform in [0,side)^2; the first
 * n - round(noise_frac*n) points go round-robin to the centres with a
 * Box-Muller Gaussian offset of sigma_frac*side per axis, rounded and clamped;
 * the rest are uniform noise points. */
void orc_gen_clustered(uint32_t n, uint32_t side, uint32_t nclusters,
                       double sigma_frac, double noise_frac, uint64_t seed,
                       int32_t *xy)
{
    orc_rng_t r;
    orc_rng_seed(&r, seed);
    double *cx = (double *)malloc(sizeof(double) * nclusters);
    double *cy = (double *)malloc(sizeof(double) * nclusters);
    for (uint32_t c = 0; c < nclusters; c++) {
        cx[c] = orc_rng_u01(&r) * (double)side;
        cy[c] = orc_rng_u01(&r) * (double)side;
    }
    uint32_t nnoise = (uint32_t)nearbyint(noise_frac * (double)n);
    if (nnoise > n) nnoise = n;
    uint32_t nclust = n - nnoise;
    double sigma = sigma_frac * (double)side;
    const double two_pi = 6.283185307179586476925286766559;
    for (uint32_t k = 0; k < nclust; k++) {
        uint32_t c = k % nclusters;
        double u1 = ((double)(orc_rng_next(&r) >> 11) + 1.0) * 0x1.0p-53; /* (0,1] */
        double u2 = orc_rng_u01(&r);
        double rad = sqrt(-2.0 * log(u1));
        double ang = two_pi * u2;
        double gx = rad * cos(ang);
        double gy = rad * sin(ang);
        xy[2 * k] = clamp_coord(cx[c] + sigma * gx, side);
        xy[2 * k + 1] = clamp_coord(cy[c] + sigma * gy, side);
    }
    for (uint32_t k = nclust; k < n; k++) {
        xy[2 * k] = (int32_t)(orc_rng_next(&r) % side);
        xy[2 * k + 1] = (int32_t)(orc_rng_next(&r) % side);
    }
    free(cx);
    free(cy);
}

/* Seeded Fisher-Yates; the reference uses std::random_shuffle
 * (construction.c:129-137). */
void orc_random_perm(uint32_t n, uint64_t seed, uint32_t *route)
{
    orc_rng_t r;
    orc_rng_seed(&r, seed);
    for (uint32_t k = 0; k < n; k++) route[k] = k;
    for (uint32_t k = n - 1; k >= 1; k--) {
        uint32_t j = (uint32_t)(orc_rng_next(&r) % ((uint64_t)k + 1));
        uint32_t t = route[k];
        route[k] = route[j];
        route[j] = t;
    }
}

/* Greedy nearest neighbour (the reference builds NN with a kd-tree that drops
 * duplicate points, kdtree_utils.c:279-301; this one is exact and keeps all
 * points). */
void orc_nn_tour(uint32_t n, const int32_t *xy, uint32_t start, uint32_t *route)
{
    unsigned char *used = (unsigned char *)calloc(n, 1);
    uint32_t cur = start;
    used[cur] = 1;
    route[0] = cur;
    for (uint32_t k = 1; k < n; k++) {
        int64_t bx = xy[2 * cur], by = xy[2 * cur + 1];
        uint64_t best = UINT64_MAX;
        uint32_t arg = 0;
        for (uint32_t u = 0; u < n; u++) {
            if (used[u]) continue;
            int64_t dx = xy[2 * u] - bx, dy = xy[2 * u + 1] - by;
            uint64_t s = (uint64_t)(dx * dx + dy * dy);
            if (s < best) { best = s; arg = u; }
        }
        used[arg] = 1;
        route[k] = arg;
        cur = arg;
    }
    free(used);
}

/* Greedy edge (multiple fragment) by brute force: all n(n-1)/2 edges sorted
 * by (squared length, lower id, higher id), each accepted when both ends have
 * degree < 2 and lie in different fragments (union-find); the remaining path
 * is closed.  Same selection rule as kdtree_utils.c:56-248 without its
 * kd-tree (which drops duplicate points).  O(n^2 log n): small n only.
 * route[0] = 0, route[1] = the lower-id neighbour of city 0.  Returns 0, or
 * -1 on n < 3 / allocation failure. */
typedef struct { uint64_t s; uint32_t a, b; } mf_edge;

static int mf_cmp(const void *p, const void *q)
{
    const mf_edge *x = (const mf_edge *)p, *y = (const mf_edge *)q;
    if (x->s != y->s) return x->s < y->s ? -1 : 1;
    if (x->a != y->a) return x->a < y->a ? -1 : 1;
    if (x->b != y->b) return x->b < y->b ? -1 : 1;
    return 0;
}

static uint32_t uf_find(uint32_t *uf, uint32_t v)
{
    while (uf[v] != v) v = uf[v] = uf[uf[v]];
    return v;
}

int orc_mf_tour(uint32_t n, const int32_t *xy, uint32_t *route)
{
    if (n < 3) return -1;
    uint64_t m = (uint64_t)n * (n - 1) / 2, k = 0;
    mf_edge *e = (mf_edge *)malloc(sizeof(mf_edge) * m);
    uint32_t *uf = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *adj = (uint32_t *)malloc(sizeof(uint32_t) * 2 * n);
    unsigned char *deg = (unsigned char *)calloc(n, 1);
    if (!e || !uf || !adj || !deg) { free(e); free(uf); free(adj); free(deg); return -1; }
    for (uint32_t a = 0; a < n; a++)
        for (uint32_t b = a + 1; b < n; b++) {
            int64_t dx = (int64_t)xy[2 * a] - xy[2 * b], dy = (int64_t)xy[2 * a + 1] - xy[2 * b + 1];
            e[k].s = (uint64_t)(dx * dx + dy * dy);
            e[k].a = a;
            e[k].b = b;
            k++;
        }
    qsort(e, m, sizeof(mf_edge), mf_cmp);
    for (uint32_t v = 0; v < n; v++) uf[v] = v;
    uint32_t added = 0;
    for (uint64_t q = 0; q < m && added + 1 < n; q++) {
        uint32_t a = e[q].a, b = e[q].b;
        if (deg[a] == 2 || deg[b] == 2) continue;
        uint32_t ra = uf_find(uf, a), rb = uf_find(uf, b);
        if (ra == rb) continue;
        adj[2 * a + deg[a]++] = b;
        adj[2 * b + deg[b]++] = a;
        uf[ra] = rb;
        added++;
    }
    uint32_t ends[2], ne = 0;
    for (uint32_t v = 0; v < n && ne < 2; v++)
        if (deg[v] < 2) ends[ne++] = v;
    adj[2 * ends[0] + deg[ends[0]]++] = ends[1];
    adj[2 * ends[1] + deg[ends[1]]++] = ends[0];
    uint32_t prev = 0, cur = adj[0] < adj[1] ? adj[0] : adj[1];
    route[0] = 0;
    for (uint32_t i = 1; i < n; i++) {
        route[i] = cur;
        uint32_t nx = adj[2 * cur] == prev ? adj[2 * cur + 1] : adj[2 * cur];
        prev = cur;
        cur = nx;
    }
    free(e);
    free(uf);
    free(adj);
    free(deg);
    return 0;
}

/* Greedy edge in two phases, the same tour as orc_mf_tour for large n:
 *  1. every pair with s <= R^2 (found through a uniform grid of cells at
 *     least R wide), sorted by (s, lower id, higher id), under the greedy
 *     rule;
 *  2. afterwards a valid edge can only join two cities of degree < 2 in
 *     different fragments, and a pair of them not yet accepted is either
 *     longer than R or in one fragment (it was rejected while both had
 *     degree < 2, and fragments only merge): so all pairs among the cities
 *     of degree < 2, sorted the same way, finish the greedy order exactly.
 * R = 0 leaves everything to phase 2 (orc_mf_tour itself).  Returns 0, or -1
 * when phase 2 would hold more than max_pairs pairs (call again with a larger
 * R) or on allocation failure. */
int orc_mf_tour_radius(uint32_t n, const int32_t *xy, uint64_t R, uint64_t max_pairs, uint32_t *route)
{
    if (n < 3) return -1;
    uint32_t *uf = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *adj = (uint32_t *)malloc(sizeof(uint32_t) * 2 * n);
    unsigned char *deg = (unsigned char *)calloc(n, 1);
    mf_edge *e = NULL;
    uint32_t *cell_start = NULL, *order = NULL, *rest = NULL;
    int rc = -1;
    if (!uf || !adj || !deg) goto out;
    for (uint32_t v = 0; v < n; v++) uf[v] = v;
    uint32_t added = 0;
    if (R > 0) {
        int64_t mnx = xy[0], mny = xy[1], mxx = xy[0], mxy = xy[1];
        for (uint32_t v = 1; v < n; v++) {
            if (xy[2 * v] < mnx) mnx = xy[2 * v];
            if (xy[2 * v] > mxx) mxx = xy[2 * v];
            if (xy[2 * v + 1] < mny) mny = xy[2 * v + 1];
            if (xy[2 * v + 1] > mxy) mxy = xy[2 * v + 1];
        }
        /* cells at least R wide and at most ~4n of them */
        int64_t cw = (int64_t)R;
        int64_t span = (mxx - mnx > mxy - mny ? mxx - mnx : mxy - mny) + 1;
        int64_t minw = span / (int64_t)(2.0 * sqrt((double)n)) + 1;
        if (cw < minw) cw = minw;
        int64_t gx = (mxx - mnx) / cw + 1, gy = (mxy - mny) / cw + 1;
        cell_start = (uint32_t *)calloc((size_t)(gx * gy + 1), sizeof(uint32_t));
        order = (uint32_t *)malloc(sizeof(uint32_t) * n);
        if (!cell_start || !order) goto out;
        for (uint32_t v = 0; v < n; v++)
            cell_start[((xy[2 * v + 1] - mny) / cw) * gx + (xy[2 * v] - mnx) / cw + 1]++;
        for (int64_t c = 0; c < gx * gy; c++) cell_start[c + 1] += cell_start[c];
        {
            uint32_t *fill = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(gx * gy));
            if (!fill) goto out;
            memcpy(fill, cell_start, sizeof(uint32_t) * (size_t)(gx * gy));
            for (uint32_t v = 0; v < n; v++) order[fill[((xy[2 * v + 1] - mny) / cw) * gx + (xy[2 * v] - mnx) / cw]++] = v;
            free(fill);
        }
        const uint64_t R2 = R * R;
        uint64_t cap = 1024, m = 0;
        e = (mf_edge *)malloc(sizeof(mf_edge) * cap);
        if (!e) goto out;
        for (uint32_t a = 0; a < n; a++) {
            const int64_t cx = (xy[2 * a] - mnx) / cw, cy = (xy[2 * a + 1] - mny) / cw;
            for (int64_t y = cy - 1; y <= cy + 1; y++) {
                if (y < 0 || y >= gy) continue;
                for (int64_t x = cx - 1; x <= cx + 1; x++) {
                    if (x < 0 || x >= gx) continue;
                    for (uint32_t q = cell_start[y * gx + x]; q < cell_start[y * gx + x + 1]; q++) {
                        const uint32_t b = order[q];
                        if (b <= a) continue;
                        int64_t dx = (int64_t)xy[2 * a] - xy[2 * b], dy = (int64_t)xy[2 * a + 1] - xy[2 * b + 1];
                        const uint64_t sq = (uint64_t)(dx * dx + dy * dy);
                        if (sq > R2) continue;
                        if (m == cap) {
                            cap *= 2;
                            mf_edge *ne = (mf_edge *)realloc(e, sizeof(mf_edge) * cap);
                            if (!ne) goto out;
                            e = ne;
                        }
                        e[m].s = sq;
                        e[m].a = a;
                        e[m].b = b;
                        m++;
                    }
                }
            }
        }
        qsort(e, m, sizeof(mf_edge), mf_cmp);
        for (uint64_t q = 0; q < m && added + 1 < n; q++) {
            uint32_t a = e[q].a, b = e[q].b;
            if (deg[a] == 2 || deg[b] == 2) continue;
            uint32_t ra = uf_find(uf, a), rb = uf_find(uf, b);
            if (ra == rb) continue;
            adj[2 * a + deg[a]++] = b;
            adj[2 * b + deg[b]++] = a;
            uf[ra] = rb;
            added++;
        }
        free(e);
        e = NULL;
    }
    if (added + 1 < n) {
        uint32_t nr = 0;
        rest = (uint32_t *)malloc(sizeof(uint32_t) * n);
        if (!rest) goto out;
        for (uint32_t v = 0; v < n; v++)
            if (deg[v] < 2) rest[nr++] = v;
        const uint64_t m = (uint64_t)nr * (nr - 1) / 2;
        if (m > max_pairs) goto out;
        e = (mf_edge *)malloc(sizeof(mf_edge) * (m ? m : 1));
        if (!e) goto out;
        uint64_t k = 0;
        for (uint32_t i = 0; i < nr; i++)
            for (uint32_t j = i + 1; j < nr; j++) {
                const uint32_t a = rest[i], b = rest[j];
                int64_t dx = (int64_t)xy[2 * a] - xy[2 * b], dy = (int64_t)xy[2 * a + 1] - xy[2 * b + 1];
                e[k].s = (uint64_t)(dx * dx + dy * dy);
                e[k].a = a;
                e[k].b = b;
                k++;
            }
        qsort(e, m, sizeof(mf_edge), mf_cmp);
        for (uint64_t q = 0; q < m && added + 1 < n; q++) {
            uint32_t a = e[q].a, b = e[q].b;
            if (deg[a] == 2 || deg[b] == 2) continue;
            uint32_t ra = uf_find(uf, a), rb = uf_find(uf, b);
            if (ra == rb) continue;
            adj[2 * a + deg[a]++] = b;
            adj[2 * b + deg[b]++] = a;
            uf[ra] = rb;
            added++;
        }
    }
    {
        uint32_t ends[2], ne = 0;
        for (uint32_t v = 0; v < n && ne < 2; v++)
            if (deg[v] < 2) ends[ne++] = v;
        if (ne != 2) goto out;
        adj[2 * ends[0] + deg[ends[0]]++] = ends[1];
        adj[2 * ends[1] + deg[ends[1]]++] = ends[0];
        uint32_t prev = 0, cur = adj[0] < adj[1] ? adj[0] : adj[1];
        route[0] = 0;
        for (uint32_t i = 1; i < n; i++) {
            route[i] = cur;
            uint32_t nx = adj[2 * cur] == prev ? adj[2 * cur + 1] : adj[2 * cur];
            prev = cur;
            cur = nx;
        }
    }
    rc = 0;
out:
    free(e);
    free(cell_start);
    free(order);
    free(rest);
    free(uf);
    free(adj);
    free(deg);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Distances                                                                  */
/* ------------------------------------------------------------------------ */
static inline uint64_t isqrt_u64(uint64_t s)
{
    uint64_t k = (uint64_t)sqrt((double)s);
    while (k * k > s) k--;
    while ((k + 1) * (k + 1) <= s) k++;
    return k;
}

/* TSPLIB EUC_2D nint(sqrt(s)) for integer s.  sqrt(s) is never k + 1/2 for
 * integer s, so nint(sqrt(s)) = k + [s > k^2 + k] exactly, k = isqrt(s).
 * The reference's float version is headers.h:126-142. */
uint32_t orc_nint_sqrt(uint64_t s)
{
    uint64_t k = isqrt_u64(s);
    return (uint32_t)(k + (s > k * k + k ? 1u : 0u));
}

/* TSPLIB95 CEIL_2D and ATT (Reinelt, "TSPLIB 95", section 2.1 and 2.5) on
 * integer s = dx^2 + dy^2, written as ceilings of square roots:
 *   ceil(sqrt(s))      = isqrt(s - 1) + 1 for s >= 1 (the smallest m with
 *                        m^2 >= s), 0 for s = 0;
 *   ceil(sqrt(s / 10)) = isqrt(c - 1) + 1 with c = ceil(s / 10) (m^2 >= s/10
 *                        iff m^2 >= c for integer m), 0 for s = 0.
 * TSPLIB's ATT recipe t = nint(r), r = sqrt(s/10), d = t + [t < r] equals
 * ceil(r).  The reference has neither (tsp_loader.c:49-162 reads every file
 * as EUC_2D). */
uint32_t orc_metric_len(uint64_t s, int metric)
{
    if (metric == ORC_METRIC_CEIL_2D) return s == 0 ? 0u : (uint32_t)(isqrt_u64(s - 1) + 1);
    if (metric == ORC_METRIC_ATT) {
        uint64_t c = (s + 9) / 10;
        return c == 0 ? 0u : (uint32_t)(isqrt_u64(c - 1) + 1);
    }
    return orc_nint_sqrt(s);
}

void orc_metric_len_batch(const uint64_t *s, uint32_t *out, uint64_t count, int metric)
{
    for (uint64_t k = 0; k < count; k++) out[k] = orc_metric_len(s[k], metric);
}

static inline uint32_t dist_xy(int64_t x1, int64_t y1, int64_t x2, int64_t y2, int metric)
{
    int64_t dx = x1 - x2, dy = y1 - y2;
    return orc_metric_len((uint64_t)(dx * dx + dy * dy), metric);
}

void orc_nint_sqrt_batch(const uint64_t *s, uint32_t *out, uint64_t count)
{
    for (uint64_t k = 0; k < count; k++) out[k] = orc_nint_sqrt(s[k]);
}

uint32_t orc_dist_m(const int32_t *xy, uint32_t a, uint32_t b, int metric)
{
    return dist_xy(xy[2 * a], xy[2 * a + 1], xy[2 * b], xy[2 * b + 1], metric);
}

uint32_t orc_dist(const int32_t *xy, uint32_t a, uint32_t b)
{
    return orc_dist_m(xy, a, b, ORC_METRIC_EUC_2D);
}

/* headers.h:170-189: sum of consecutive edges plus the closing edge. */
uint64_t orc_route_length_m(uint32_t n, const int32_t *xy, const uint32_t *route, int metric)
{
    uint64_t d = 0;
    for (uint32_t k = 1; k < n; k++) d += orc_dist_m(xy, route[k - 1], route[k], metric);
    if (n > 0) d += orc_dist_m(xy, route[n - 1], route[0], metric);
    return d;
}

uint64_t orc_route_length(uint32_t n, const int32_t *xy, const uint32_t *route)
{
    return orc_route_length_m(n, xy, route, ORC_METRIC_EUC_2D);
}

/* construction.c:152-183 (returns an error code instead of exit(-1)). */
int orc_check_tour(uint32_t n, const uint32_t *route)
{
    if (n < 3) return -1;
    unsigned char *seen = (unsigned char *)calloc(n, 1);
    int ok = 0;
    for (uint32_t k = 0; k < n; k++) {
        if (route[k] >= n || seen[route[k]]) { ok = -1; break; }
        seen[route[k]] = 1;
    }
    free(seen);
    return ok;
}

/* ------------------------------------------------------------------------ */
/* Sweep                                                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t n;
    const int64_t *px, *py; /* tour-ordered coordinates, cpuSolver.cpp:67-79 */
    const uint32_t *dk;     /* dk[k] = d(P[k-1],P[k]) for k >= 1            */
    uint32_t i_lo, i_hi;    /* i in [i_lo, i_hi)                            */
    int metric;
    int32_t best;
    uint32_t bi, bj;
} sweep_job;

/* cpuSolver.cpp:104-146: i ascending, J ascending, update on strict '<',
 * only negative gains are recorded. */
static void *sweep_range(void *arg)
{
    sweep_job *jb = (sweep_job *)arg;
    const int64_t *px = jb->px, *py = jb->py;
    const uint32_t *dk = jb->dk;
    uint32_t n = jb->n;
    const int metric = jb->metric;
    int32_t best = 0;
    uint32_t bi = 0, bj = 0;
    for (uint32_t i = jb->i_lo; i < jb->i_hi; i++) {
        int64_t xi = px[i], yi = py[i], xh = px[i - 1], yh = py[i - 1];
        int32_t c1 = (int32_t)dk[i];
        for (uint32_t J = i + 2; J <= n - 1; J++) {
            int32_t g = (int32_t)dist_xy(xi, yi, px[J], py[J], metric) +
                        (int32_t)dist_xy(xh, yh, px[J - 1], py[J - 1], metric) - c1 -
                        (int32_t)dk[J];
            if (g < best) {
                best = g;
                bi = i;
                bj = J;
            }
        }
    }
    jb->best = best;
    jb->bi = bi;
    jb->bj = bj;
    return NULL;
}

int orc_step_m(uint32_t n, const int32_t *xy, const uint32_t *route, int threads, int metric,
               orc_best2_out *out)
{
    out->i = 0;
    out->j = 0;
    out->minchange = 0;
    if (n < 4) return 0;
    int64_t *px = (int64_t *)malloc(sizeof(int64_t) * n);
    int64_t *py = (int64_t *)malloc(sizeof(int64_t) * n);
    uint32_t *dk = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (uint32_t k = 0; k < n; k++) {
        px[k] = xy[2 * route[k]];
        py[k] = xy[2 * route[k] + 1];
    }
    dk[0] = dist_xy(px[n - 1], py[n - 1], px[0], py[0], metric);
    for (uint32_t k = 1; k < n; k++) dk[k] = dist_xy(px[k - 1], py[k - 1], px[k], py[k], metric);

    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    uint32_t rows = n - 3; /* i in [1, n-3] */
    if ((uint32_t)threads > rows) threads = (int)rows;
    sweep_job jobs[256];
    pthread_t tid[256];
    /* split the triangle into ranges of roughly equal pair counts
     * (cpuMTSolver.cpp:109-138 splits the pair range per thread) */
    double total = (double)(n - 2) * (double)(n - 3) / 2.0;
    uint32_t lo = 1;
    for (int t = 0; t < threads; t++) {
        uint32_t hi;
        if (t == threads - 1) {
            hi = n - 2;
        } else {
            double target = total * (double)(t + 1) / (double)threads, acc = 0.0;
            /* pairs with i < h: sum_{i=1}^{h-1} (n-2-i) */
            hi = lo;
            while (hi < n - 2) {
                double h = (double)hi;
                acc = (h - 1.0) * (double)(n - 2) - (h - 1.0) * h / 2.0;
                if (acc >= target) break;
                hi++;
            }
        }
        jobs[t].n = n;
        jobs[t].px = px;
        jobs[t].py = py;
        jobs[t].dk = dk;
        jobs[t].i_lo = lo;
        jobs[t].i_hi = hi;
        jobs[t].metric = metric;
        lo = hi;
    }
    if (threads == 1) {
        sweep_range(&jobs[0]);
    } else {
        for (int t = 0; t < threads; t++) pthread_create(&tid[t], NULL, sweep_range, &jobs[t]);
        for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
    }
    /* merge in ascending i order with strict '<': identical to one scan */
    int32_t best = 0;
    uint32_t bi = 0, bj = 0;
    for (int t = 0; t < threads; t++) {
        if (jobs[t].best < best) {
            best = jobs[t].best;
            bi = jobs[t].bi;
            bj = jobs[t].bj;
        }
    }
    free(px);
    free(py);
    free(dk);
    if (best < 0) {
        out->i = bi;
        out->j = bj;
        out->minchange = best;
        return 1;
    }
    return 0;
}

int orc_step(uint32_t n, const int32_t *xy, const uint32_t *route, int threads,
             orc_best2_out *out)
{
    return orc_step_m(n, xy, route, threads, ORC_METRIC_EUC_2D, out);
}

/* twoOptLocalSolver.cpp:41-48 in closed form:
 *   rotate(begin, begin+i, end); reverse(begin+(J-i), end)
 * gives r' = r[i..J-1] ++ r[i-1], ..., r[0] ++ r[n-1], ..., r[J]. */
void orc_apply(uint32_t n, uint32_t *route, uint32_t i, uint32_t j)
{
    if (i > j) { uint32_t t = i; i = j; j = t; }
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t w = 0;
    for (uint32_t k = i; k < j; k++) tmp[w++] = route[k];
    for (uint32_t k = i; k-- > 0;) tmp[w++] = route[k];
    for (uint32_t k = n; k-- > j;) tmp[w++] = route[k];
    memcpy(route, tmp, sizeof(uint32_t) * n);
    free(tmp);
}

/* The sweep of the loops below: the plain scan (orc_step_m), or the exact
 * SIMD sweep (orc_step_fast), which returns the same move and makes full-size
 * runs (tens of thousands of cities, thousands of sweeps) take minutes. */
static int loop_step(int fast, uint32_t n, const int32_t *xy, const uint32_t *route, int threads, int metric,
                     orc_best2_out *o)
{
    return fast ? orc_step_fast(n, xy, route, threads, metric, o, NULL) : orc_step_m(n, xy, route, threads, metric, o);
}

/* twoOptLocalSolver.cpp:22-64: step, apply while the gain is negative. */
static uint64_t ls_core(uint32_t n, const int32_t *xy, uint32_t *route, int threads, int metric, int fast,
                        uint64_t max_moves, orc_best2_out *log, uint64_t log_cap, uint64_t *sweeps,
                        uint64_t *length)
{
    uint64_t moves = 0, sw = 0;
    uint64_t L = orc_route_length_m(n, xy, route, metric);
    for (;;) {
        if (max_moves && moves >= max_moves) break;
        orc_best2_out o;
        sw++;
        if (!loop_step(fast, n, xy, route, threads, metric, &o)) break;
        orc_apply(n, route, o.i, o.j);
        L = (uint64_t)((int64_t)L + o.minchange);
        if (log && moves < log_cap) log[moves] = o;
        moves++;
    }
    if (sweeps) *sweeps = sw;
    if (length) *length = L;
    return moves;
}

uint64_t orc_local_search_m(uint32_t n, const int32_t *xy, uint32_t *route,
                            int threads, int metric, uint64_t max_moves,
                            orc_best2_out *log, uint64_t log_cap,
                            uint64_t *sweeps, uint64_t *length)
{
    return ls_core(n, xy, route, threads, metric, 0, max_moves, log, log_cap, sweeps, length);
}

uint64_t orc_local_search_fast(uint32_t n, const int32_t *xy, uint32_t *route,
                               int threads, int metric, uint64_t max_moves,
                               orc_best2_out *log, uint64_t log_cap,
                               uint64_t *sweeps, uint64_t *length)
{
    return ls_core(n, xy, route, threads, metric, 1, max_moves, log, log_cap, sweeps, length);
}

uint64_t orc_local_search(uint32_t n, const int32_t *xy, uint32_t *route,
                          int threads, uint64_t max_moves,
                          orc_best2_out *log, uint64_t log_cap,
                          uint64_t *sweeps, uint64_t *length)
{
    return orc_local_search_m(n, xy, route, threads, ORC_METRIC_EUC_2D, max_moves, log, log_cap,
                              sweeps, length);
}

/* Build-defined kick: three distinct cut points from 1..n-1, sorted. */
void orc_double_bridge(uint32_t n, uint32_t *route, orc_rng_t *rng, uint32_t *cuts)
{
    uint32_t p[3];
    int have = 0;
    while (have < 3) {
        uint32_t v = 1u + (uint32_t)(orc_rng_next(rng) % (uint64_t)(n - 1));
        int dup = 0;
        for (int k = 0; k < have; k++) dup |= (p[k] == v);
        if (!dup) p[have++] = v;
    }
    /* sort three values */
    if (p[0] > p[1]) { uint32_t t = p[0]; p[0] = p[1]; p[1] = t; }
    if (p[1] > p[2]) { uint32_t t = p[1]; p[1] = p[2]; p[2] = t; }
    if (p[0] > p[1]) { uint32_t t = p[0]; p[0] = p[1]; p[1] = t; }
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t w = 0;
    for (uint32_t k = 0; k < p[0]; k++) tmp[w++] = route[k];    /* A */
    for (uint32_t k = p[1]; k < p[2]; k++) tmp[w++] = route[k]; /* C */
    for (uint32_t k = p[0]; k < p[1]; k++) tmp[w++] = route[k]; /* B */
    for (uint32_t k = p[2]; k < n; k++) tmp[w++] = route[k];    /* D */
    memcpy(route, tmp, sizeof(uint32_t) * n);
    free(tmp);
    if (cuts) { cuts[0] = p[0]; cuts[1] = p[1]; cuts[2] = p[2]; }
}

/* ILSGlobalSolver.cpp:46-47: rotate(begin, begin+a, end);
 * reverse(begin+(b-a), end), for a <= b. */
void orc_rotate_reverse(uint32_t n, uint32_t *route, uint32_t a, uint32_t b)
{
    if (b < a) { uint32_t t = a; a = b; b = t; }
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (uint32_t k = 0; k < n; k++) tmp[k] = route[(a + k) % n];
    uint32_t lo = b - a, hi = n - 1;
    while (lo < hi) {
        uint32_t t = tmp[lo];
        tmp[lo] = tmp[hi];
        tmp[hi] = t;
        lo++;
        hi--;
    }
    memcpy(route, tmp, sizeof(uint32_t) * n);
    free(tmp);
}

/* ILSGlobalSolver.cpp:37-45: a, b = rand() % n, sorted (here from the seeded
 * generator), then the rotate+reverse above. */
void orc_ref_kick(uint32_t n, uint32_t *route, orc_rng_t *rng)
{
    uint32_t a = (uint32_t)(orc_rng_next(rng) % n);
    uint32_t b = (uint32_t)(orc_rng_next(rng) % n);
    orc_rotate_reverse(n, route, a, b);
}

/* ILSGlobalSolverSequential.cpp:60-100 with a fixed round count in place of
 * the target/time stop rule: LO, accept iff strictly shorter else revert to
 * the best, kick the best. */
static uint64_t ils_core(uint32_t n, const int32_t *xy, uint32_t *route, uint64_t seed, uint32_t rounds,
                         int kick_kind, int threads, int metric, int fast, uint64_t *round_lengths,
                         uint64_t *total_moves)
{
    orc_rng_t rng;
    orc_rng_seed(&rng, seed);
    uint64_t L, tm = 0;
    tm += ls_core(n, xy, route, threads, metric, fast, 0, NULL, 0, NULL, &L);
    uint64_t Lbest = L;
    uint32_t *best = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *cur = (uint32_t *)malloc(sizeof(uint32_t) * n);
    memcpy(best, route, sizeof(uint32_t) * n);
    if (round_lengths) round_lengths[0] = Lbest;
    for (uint32_t r = 0; r < rounds; r++) {
        memcpy(cur, best, sizeof(uint32_t) * n);
        if (kick_kind == ORC_KICK_REFERENCE)
            orc_ref_kick(n, cur, &rng);
        else
            orc_double_bridge(n, cur, &rng, NULL);
        tm += ls_core(n, xy, cur, threads, metric, fast, 0, NULL, 0, NULL, &L);
        if (L < Lbest) {
            Lbest = L;
            memcpy(best, cur, sizeof(uint32_t) * n);
        }
        if (round_lengths) round_lengths[r + 1] = Lbest;
    }
    memcpy(route, best, sizeof(uint32_t) * n);
    free(best);
    free(cur);
    if (total_moves) *total_moves = tm;
    return Lbest;
}

uint64_t orc_ils_m(uint32_t n, const int32_t *xy, uint32_t *route, uint64_t seed,
                   uint32_t rounds, int kick_kind, int threads, int metric,
                   uint64_t *round_lengths, uint64_t *total_moves)
{
    return ils_core(n, xy, route, seed, rounds, kick_kind, threads, metric, 0, round_lengths, total_moves);
}

uint64_t orc_ils_fast(uint32_t n, const int32_t *xy, uint32_t *route, uint64_t seed,
                      uint32_t rounds, int kick_kind, int threads, int metric,
                      uint64_t *round_lengths, uint64_t *total_moves)
{
    return ils_core(n, xy, route, seed, rounds, kick_kind, threads, metric, 1, round_lengths, total_moves);
}

uint64_t orc_ils(uint32_t n, const int32_t *xy, uint32_t *route, uint64_t seed,
                 uint32_t rounds, int kick_kind, int threads,
                 uint64_t *round_lengths, uint64_t *total_moves)
{
    return orc_ils_m(n, xy, route, seed, rounds, kick_kind, threads, ORC_METRIC_EUC_2D,
                     round_lengths, total_moves);
}

/* ILSGlobalSolverMT::optimize with comm = 1 (ILSGlobalSolverMT.cpp:86-195)
 * restated deterministically: `chains` chains start from route; round 0 is a
 * local search of every chain; in every later round each chain kicks (its own
 * generator, seed + c) its own best if its last local search improved that
 * best, else the group best, and runs a local search; a chain keeps a result
 * iff it is strictly shorter than its best.  After each round the group best
 * is the chain with the smallest best (lowest chain on ties).  The reference
 * shares the global best through a mutex between free-running threads, so
 * its sequence depends on timing; here rounds are synchronised.  route
 * receives the group best; round_lengths (rounds+1) the group best per round.
 * Returns the group best length. */
uint64_t orc_ils_multistart(uint32_t n, const int32_t *xy, uint32_t *route, uint64_t seed,
                            uint32_t rounds, uint32_t chains, int kick_kind, int threads,
                            int metric, uint64_t *round_lengths, uint64_t *total_moves)
{
    uint32_t **best = (uint32_t **)malloc(sizeof(uint32_t *) * chains);
    uint32_t *cur = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *gbest = (uint32_t *)malloc(sizeof(uint32_t) * n); /* group best at the last barrier */
    uint64_t *bl = (uint64_t *)malloc(sizeof(uint64_t) * chains);
    int *improved = (int *)malloc(sizeof(int) * chains);
    orc_rng_t *rng = (orc_rng_t *)malloc(sizeof(orc_rng_t) * chains);
    uint64_t tm = 0, L;
    uint32_t g = 0;
    for (uint32_t c = 0; c < chains; c++) {
        best[c] = (uint32_t *)malloc(sizeof(uint32_t) * n);
        orc_rng_seed(&rng[c], seed + c);
        memcpy(cur, route, sizeof(uint32_t) * n);
        tm += orc_local_search_m(n, xy, cur, threads, metric, 0, NULL, 0, NULL, &L);
        memcpy(best[c], cur, sizeof(uint32_t) * n);
        bl[c] = L;
        improved[c] = 1;
        if (bl[c] < bl[g]) g = c;
    }
    if (round_lengths) round_lengths[0] = bl[g];
    for (uint32_t r = 0; r < rounds; r++) {
        /* every chain of this round sees the group best of the previous one,
         * even after the chain holding it has moved on within the round */
        memcpy(gbest, best[g], sizeof(uint32_t) * n);
        for (uint32_t c = 0; c < chains; c++) {
            memcpy(cur, improved[c] ? best[c] : gbest, sizeof(uint32_t) * n);
            if (kick_kind == ORC_KICK_REFERENCE)
                orc_ref_kick(n, cur, &rng[c]);
            else
                orc_double_bridge(n, cur, &rng[c], NULL);
            tm += orc_local_search_m(n, xy, cur, threads, metric, 0, NULL, 0, NULL, &L);
            improved[c] = L < bl[c];
            if (improved[c]) {
                bl[c] = L;
                memcpy(best[c], cur, sizeof(uint32_t) * n);
            }
        }
        g = 0;
        for (uint32_t c = 1; c < chains; c++)
            if (bl[c] < bl[g]) g = c;
        if (round_lengths) round_lengths[r + 1] = bl[g];
    }
    memcpy(route, best[g], sizeof(uint32_t) * n);
    L = bl[g];
    for (uint32_t c = 0; c < chains; c++) free(best[c]);
    free(best);
    free(cur);
    free(gbest);
    free(bl);
    free(improved);
    free(rng);
    if (total_moves) *total_moves = tm;
    return L;
}

/* ------------------------------------------------------------------------ */
/* Exact SIMD sweep: the CPU baseline (same result as orc_step_m)             */
/* ------------------------------------------------------------------------ */
/* The cpuSolver.cpp:148-197 / cpuMTSolver.cpp:79-340 analogue done exactly:
 * a candidate (i, J) can only beat the running best `best` if
 *   L1(i-1, J-1) + L1(i, J) - W[J] <= ceil(sqrt(g) best) + W[i],
 * with W[k] = ceil(sqrt(g) D[k]) and g = 2 (EUC_2D, CEIL_2D) or 20 (ATT):
 * |dx| + |dy| <= sqrt(2) d and d < D + 1/2, d <= D, d <= sqrt(10) D for the
 * three metrics.  The integer test runs 16 (AVX-512, four rows per block of
 * column loads) or 8 (AVX2) columns at a time; the rare survivors get the
 * exact lengths and the minimum of the key (gain, i, J), so the result is
 * the lexicographic minimum like orc_step_m.  Threads take i-ranges of equal
 * pair counts and merge in i order. */
#include <immintrin.h>

static int32_t ceil_sqrtg_i(int32_t d, int g)
{
    if (d <= 0) return 0;
    uint64_t s = (uint64_t)g * (uint64_t)d * (uint64_t)d;
    return (int32_t)(isqrt_u64(s) + 1); /* g d^2 is never a square for d >= 1 */
}

/* ceil(sqrt(g) * (best_gain_allowed + 1)) for the largest gain still wanted:
 * improving moves only (best = 0) -> 0; else best - 1 + 1 = best <= -1 */
static int32_t tier1_threshold(int32_t best, int g)
{
    return best == 0 ? 0 : 1 - ceil_sqrtg_i(-best, g);
}

typedef struct {
    uint32_t n, i_lo, i_hi;
    int metric, g;
    const int32_t *x, *y, *D, *W;
    int32_t best;
    uint32_t bi, bj;
    uint64_t survivors;
    int isa;
} fast_job;

/* exact check of candidate column b of row a (i = a+1, J = b+1).  Rows are
 * scanned in blocks, so candidates do not arrive in (i, J) order: the best is
 * the minimum of the full key (gain, i, J), and the threshold admits gains
 * equal to the best (tb = ceil(sqrt(g) (best + 1))). */
static inline void fast_exact(fast_job *jb, uint32_t a, uint32_t b, int32_t *tb)
{
    const int32_t *x = jb->x, *y = jb->y;
    int32_t g = (int32_t)dist_xy(x[a], y[a], x[b], y[b], jb->metric) +
                (int32_t)dist_xy(x[a + 1], y[a + 1], x[b + 1], y[b + 1], jb->metric) - jb->D[a + 1] -
                jb->D[b + 1];
    jb->survivors++;
    if (g < 0 && (g < jb->best || (g == jb->best && (a + 1 < jb->bi || (a + 1 == jb->bi && b + 1 < jb->bj))))) {
        jb->best = g;
        jb->bi = a + 1;
        jb->bj = b + 1;
        *tb = tier1_threshold(g + 1, jb->g);
    }
}

static inline int32_t fast_l1(const fast_job *jb, uint32_t a, uint32_t b)
{
    int32_t dx = jb->x[a] - jb->x[b], dy = jb->y[a] - jb->y[b];
    return (dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy);
}

static int32_t fast_tb(const fast_job *jb)
{
    return jb->best == 0 ? 0 : tier1_threshold(jb->best + 1, jb->g);
}

static void fast_rows_scalar(fast_job *jb)
{
    int32_t tb = fast_tb(jb);
    for (uint32_t a = jb->i_lo - 1; a + 1 < jb->i_hi; a++)
        for (uint32_t b = a + 2; b + 2 <= jb->n; b++)
            if (fast_l1(jb, a, b) + fast_l1(jb, a + 1, b + 1) - jb->W[b + 1] <= tb + jb->W[a + 1])
                fast_exact(jb, a, b, &tb);
}

#define FAST_L1_SUM8(r)                                                                                    \
    _mm256_sub_epi32(                                                                                      \
        _mm256_add_epi32(_mm256_add_epi32(_mm256_abs_epi32(_mm256_sub_epi32(cx0, px##r)),                   \
                                          _mm256_abs_epi32(_mm256_sub_epi32(cy0, py##r))),                  \
                         _mm256_add_epi32(_mm256_abs_epi32(_mm256_sub_epi32(cx1, px##r##n)),                \
                                          _mm256_abs_epi32(_mm256_sub_epi32(cy1, py##r##n)))),              \
        w1)
/* s <= t as an 8-bit mask */
#define FAST_LE8(s_, t_) ((unsigned)_mm256_movemask_ps(_mm256_castsi256_ps(_mm256_cmpgt_epi32((t_), (s_)))))

__attribute__((target("avx2"))) static void fast_rows_avx2(fast_job *jb)
{
    const int32_t *x = jb->x, *y = jb->y, *W = jb->W;
    int32_t tb = fast_tb(jb);
    const uint32_t end = jb->n - 1; /* columns b <= n-2 */
    uint32_t a0 = jb->i_lo - 1;
    const uint32_t a_end = jb->i_hi - 1;
    for (; a0 + 4 <= a_end; a0 += 4) {
        const uint32_t bv = a0 + 5;
        for (uint32_t r = 0; r < 4; r++)
            for (uint32_t b = a0 + r + 2; b < bv && b < end; b++)
                if (fast_l1(jb, a0 + r, b) + fast_l1(jb, a0 + r + 1, b + 1) - W[b + 1] <= tb + W[a0 + r + 1])
                    fast_exact(jb, a0 + r, b, &tb);
        const __m256i px0 = _mm256_set1_epi32(x[a0]), py0 = _mm256_set1_epi32(y[a0]);
        const __m256i px1 = _mm256_set1_epi32(x[a0 + 1]), py1 = _mm256_set1_epi32(y[a0 + 1]);
        const __m256i px2 = _mm256_set1_epi32(x[a0 + 2]), py2 = _mm256_set1_epi32(y[a0 + 2]);
        const __m256i px3 = _mm256_set1_epi32(x[a0 + 3]), py3 = _mm256_set1_epi32(y[a0 + 3]);
        const __m256i px4 = _mm256_set1_epi32(x[a0 + 4]), py4 = _mm256_set1_epi32(y[a0 + 4]);
        const __m256i px0n = px1, py0n = py1, px1n = px2, py1n = py2, px2n = px3, py2n = py3, px3n = px4,
                      py3n = py4;
        /* s <= tb + W  <=>  tb + W + 1 > s */
        __m256i t0 = _mm256_set1_epi32(tb + W[a0 + 1] + 1), t1 = _mm256_set1_epi32(tb + W[a0 + 2] + 1);
        __m256i t2 = _mm256_set1_epi32(tb + W[a0 + 3] + 1), t3 = _mm256_set1_epi32(tb + W[a0 + 4] + 1);
        uint32_t b = bv;
        for (; b + 8 <= end; b += 8) {
            const __m256i cx0 = _mm256_loadu_si256((const __m256i *)(x + b));
            const __m256i cy0 = _mm256_loadu_si256((const __m256i *)(y + b));
            const __m256i cx1 = _mm256_loadu_si256((const __m256i *)(x + b + 1));
            const __m256i cy1 = _mm256_loadu_si256((const __m256i *)(y + b + 1));
            const __m256i w1 = _mm256_loadu_si256((const __m256i *)(W + b + 1));
            const unsigned m0 = FAST_LE8(FAST_L1_SUM8(0), t0), m1 = FAST_LE8(FAST_L1_SUM8(1), t1);
            const unsigned m2 = FAST_LE8(FAST_L1_SUM8(2), t2), m3 = FAST_LE8(FAST_L1_SUM8(3), t3);
            if (__builtin_expect((m0 | m1 | m2 | m3) != 0, 0)) {
                const unsigned mm[4] = {m0, m1, m2, m3};
                for (uint32_t r = 0; r < 4; r++)
                    for (unsigned m = mm[r]; m; m &= m - 1) fast_exact(jb, a0 + r, b + (unsigned)__builtin_ctz(m), &tb);
                t0 = _mm256_set1_epi32(tb + W[a0 + 1] + 1);
                t1 = _mm256_set1_epi32(tb + W[a0 + 2] + 1);
                t2 = _mm256_set1_epi32(tb + W[a0 + 3] + 1);
                t3 = _mm256_set1_epi32(tb + W[a0 + 4] + 1);
            }
        }
        for (uint32_t r = 0; r < 4; r++)
            for (uint32_t c = b; c < end; c++)
                if (fast_l1(jb, a0 + r, c) + fast_l1(jb, a0 + r + 1, c + 1) - W[c + 1] <= tb + W[a0 + r + 1])
                    fast_exact(jb, a0 + r, c, &tb);
    }
    for (; a0 < a_end; a0++)
        for (uint32_t b = a0 + 2; b < end; b++)
            if (fast_l1(jb, a0, b) + fast_l1(jb, a0 + 1, b + 1) - W[b + 1] <= tb + W[a0 + 1]) fast_exact(jb, a0, b, &tb);
}

/* four rows per block of column loads (AVX-512); leftover rows go one by one */
#define FAST_L1_SUM(r)                                                                                     \
    _mm512_sub_epi32(                                                                                      \
        _mm512_add_epi32(_mm512_add_epi32(_mm512_abs_epi32(_mm512_sub_epi32(cx0, px##r)),                   \
                                          _mm512_abs_epi32(_mm512_sub_epi32(cy0, py##r))),                  \
                         _mm512_add_epi32(_mm512_abs_epi32(_mm512_sub_epi32(cx1, px##r##n)),                \
                                          _mm512_abs_epi32(_mm512_sub_epi32(cy1, py##r##n)))),              \
        w1)

__attribute__((target("avx512f"))) static void fast_row1_avx512(fast_job *jb, uint32_t a, int32_t *tb)
{
    const int32_t *x = jb->x, *y = jb->y, *W = jb->W;
    const uint32_t end = jb->n - 1; /* columns b <= n-2 */
    const __m512i px0 = _mm512_set1_epi32(x[a]), py0 = _mm512_set1_epi32(y[a]);
    const __m512i px0n = _mm512_set1_epi32(x[a + 1]), py0n = _mm512_set1_epi32(y[a + 1]);
    uint32_t b = a + 2;
    for (; b + 16 <= end; b += 16) {
        const __m512i cx0 = _mm512_loadu_si512(x + b), cy0 = _mm512_loadu_si512(y + b);
        const __m512i cx1 = _mm512_loadu_si512(x + b + 1), cy1 = _mm512_loadu_si512(y + b + 1);
        const __m512i w1 = _mm512_loadu_si512(W + b + 1);
        unsigned m = (unsigned)_mm512_cmple_epi32_mask(FAST_L1_SUM(0), _mm512_set1_epi32(*tb + W[a + 1]));
        while (m) {
            unsigned k = (unsigned)__builtin_ctz(m);
            m &= m - 1;
            fast_exact(jb, a, b + k, tb);
        }
    }
    for (; b < end; b++)
        if (fast_l1(jb, a, b) + fast_l1(jb, a + 1, b + 1) - W[b + 1] <= *tb + W[a + 1]) fast_exact(jb, a, b, tb);
}

__attribute__((target("avx512f"))) static void fast_rows_avx512(fast_job *jb)
{
    const int32_t *x = jb->x, *y = jb->y, *W = jb->W;
    int32_t tb = fast_tb(jb);
    const uint32_t end = jb->n - 1;
    uint32_t a0 = jb->i_lo - 1;
    const uint32_t a_end = jb->i_hi - 1; /* rows a < a_end */
    for (; a0 + 4 <= a_end; a0 += 4) {
        /* columns below a0 + 5 belong to only some of the four rows: scalar */
        const uint32_t bv = a0 + 5;
        for (uint32_t r = 0; r < 4; r++)
            for (uint32_t b = a0 + r + 2; b < bv && b < end; b++)
                if (fast_l1(jb, a0 + r, b) + fast_l1(jb, a0 + r + 1, b + 1) - W[b + 1] <= tb + W[a0 + r + 1])
                    fast_exact(jb, a0 + r, b, &tb);
        const __m512i px0 = _mm512_set1_epi32(x[a0]), py0 = _mm512_set1_epi32(y[a0]);
        const __m512i px1 = _mm512_set1_epi32(x[a0 + 1]), py1 = _mm512_set1_epi32(y[a0 + 1]);
        const __m512i px2 = _mm512_set1_epi32(x[a0 + 2]), py2 = _mm512_set1_epi32(y[a0 + 2]);
        const __m512i px3 = _mm512_set1_epi32(x[a0 + 3]), py3 = _mm512_set1_epi32(y[a0 + 3]);
        const __m512i px4 = _mm512_set1_epi32(x[a0 + 4]), py4 = _mm512_set1_epi32(y[a0 + 4]);
        const __m512i px0n = px1, py0n = py1, px1n = px2, py1n = py2, px2n = px3, py2n = py3, px3n = px4,
                      py3n = py4;
        __m512i t0 = _mm512_set1_epi32(tb + W[a0 + 1]), t1 = _mm512_set1_epi32(tb + W[a0 + 2]);
        __m512i t2 = _mm512_set1_epi32(tb + W[a0 + 3]), t3 = _mm512_set1_epi32(tb + W[a0 + 4]);
        uint32_t b = bv;
        for (; b + 16 <= end; b += 16) {
            const __m512i cx0 = _mm512_loadu_si512(x + b), cy0 = _mm512_loadu_si512(y + b);
            const __m512i cx1 = _mm512_loadu_si512(x + b + 1), cy1 = _mm512_loadu_si512(y + b + 1);
            const __m512i w1 = _mm512_loadu_si512(W + b + 1);
            const __mmask16 m0 = _mm512_cmple_epi32_mask(FAST_L1_SUM(0), t0);
            const __mmask16 m1 = _mm512_cmple_epi32_mask(FAST_L1_SUM(1), t1);
            const __mmask16 m2 = _mm512_cmple_epi32_mask(FAST_L1_SUM(2), t2);
            const __mmask16 m3 = _mm512_cmple_epi32_mask(FAST_L1_SUM(3), t3);
            if (__builtin_expect((m0 | m1 | m2 | m3) != 0, 0)) {
                const unsigned mm[4] = {m0, m1, m2, m3};
                for (uint32_t r = 0; r < 4; r++)
                    for (unsigned m = mm[r]; m; m &= m - 1) fast_exact(jb, a0 + r, b + (unsigned)__builtin_ctz(m), &tb);
                t0 = _mm512_set1_epi32(tb + W[a0 + 1]);
                t1 = _mm512_set1_epi32(tb + W[a0 + 2]);
                t2 = _mm512_set1_epi32(tb + W[a0 + 3]);
                t3 = _mm512_set1_epi32(tb + W[a0 + 4]);
            }
        }
        for (uint32_t r = 0; r < 4; r++)
            for (uint32_t c = b; c < end; c++)
                if (fast_l1(jb, a0 + r, c) + fast_l1(jb, a0 + r + 1, c + 1) - W[c + 1] <= tb + W[a0 + r + 1])
                    fast_exact(jb, a0 + r, c, &tb);
    }
    for (; a0 < a_end; a0++) fast_row1_avx512(jb, a0, &tb);
}

/* the widest variant this CPU runs; ORC_FAST_ISA=avx2|scalar narrows it
 * (used by the tests to cover every variant on one machine) */
static int fast_isa_id(void)
{
    const char *want = getenv("ORC_FAST_ISA");
    __builtin_cpu_init();
    if (want && !strcmp(want, "scalar")) return 0;
    if (want && !strcmp(want, "avx2")) return __builtin_cpu_supports("avx2") ? 1 : 0;
    if (__builtin_cpu_supports("avx512f")) return 2;
    if (__builtin_cpu_supports("avx2")) return 1;
    return 0;
}

const char *orc_fast_isa(void)
{
    static const char *names[] = {"scalar", "avx2", "avx512"};
    return names[fast_isa_id()];
}

static void *fast_thread(void *arg)
{
    fast_job *jb = (fast_job *)arg;
    if (jb->isa == 2) fast_rows_avx512(jb);
    else if (jb->isa == 1) fast_rows_avx2(jb);
    else fast_rows_scalar(jb);
    return NULL;
}

int orc_step_fast(uint32_t n, const int32_t *xy, const uint32_t *route, int threads, int metric,
                  orc_best2_out *out, uint64_t *survivors)
{
    out->i = out->j = 0;
    out->minchange = 0;
    if (survivors) *survivors = 0;
    if (n < 4) return 0;
    int32_t *x = (int32_t *)malloc(sizeof(int32_t) * (n + 16));
    int32_t *y = (int32_t *)malloc(sizeof(int32_t) * (n + 16));
    int32_t *D = (int32_t *)malloc(sizeof(int32_t) * (n + 16));
    int32_t *W = (int32_t *)malloc(sizeof(int32_t) * (n + 16));
    const int g = metric == ORC_METRIC_ATT ? 20 : 2;
    for (uint32_t k = 0; k < n; k++) {
        x[k] = xy[2 * route[k]];
        y[k] = xy[2 * route[k] + 1];
    }
    for (uint32_t k = 0; k < n; k++) {
        uint32_t p = k == 0 ? n - 1 : k - 1;
        D[k] = (int32_t)dist_xy(x[p], y[p], x[k], y[k], metric);
        W[k] = ceil_sqrtg_i(D[k], g);
    }
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if ((uint32_t)threads > n - 3) threads = (int)(n - 3);
    fast_job jobs[256];
    pthread_t tid[256];
    double total = (double)(n - 2) * (double)(n - 3) / 2.0;
    uint32_t lo = 1;
    const int isa = fast_isa_id();
    for (int t = 0; t < threads; t++) {
        uint32_t hi;
        if (t == threads - 1) {
            hi = n - 2;
        } else {
            double target = total * (double)(t + 1) / (double)threads;
            hi = lo;
            while (hi < n - 2) {
                double h = (double)hi;
                if ((h - 1.0) * (double)(n - 2) - (h - 1.0) * h / 2.0 >= target) break;
                hi++;
            }
        }
        fast_job *jb = &jobs[t];
        jb->isa = isa;
        jb->n = n;
        jb->i_lo = lo;
        jb->i_hi = hi;
        jb->metric = metric;
        jb->g = g;
        jb->x = x;
        jb->y = y;
        jb->D = D;
        jb->W = W;
        jb->best = 0;
        jb->bi = jb->bj = 0;
        jb->survivors = 0;
        lo = hi;
    }
    if (threads == 1) {
        fast_thread(&jobs[0]);
    } else {
        for (int t = 0; t < threads; t++) pthread_create(&tid[t], NULL, fast_thread, &jobs[t]);
        for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
    }
    int32_t best = 0;
    uint32_t bi = 0, bj = 0;
    uint64_t sv = 0;
    for (int t = 0; t < threads; t++) {
        sv += jobs[t].survivors;
        if (jobs[t].i_lo < jobs[t].i_hi && jobs[t].best < best) {
            best = jobs[t].best;
            bi = jobs[t].bi;
            bj = jobs[t].bj;
        }
    }
    free(x);
    free(y);
    free(D);
    free(W);
    if (survivors) *survivors = sv;
    if (best < 0) {
        out->i = bi;
        out->j = bj;
        out->minchange = best;
        return 1;
    }
    return 0;
}
