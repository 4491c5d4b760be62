// l2o_server.cpp -- continuous-batching request loop over the engine.
//
// Requests (a local search or an ILS solve) wait in a FIFO queue on the host.
// Each pump() admits queued requests into free engine slots while the HBM
// budget and the slot count allow (engine_add copies coordinates and the start
// tour to HBM once; they stay resident until the result is harvested), runs a
// bounded number of batched steps in which every admitted request advances
// together, and harvests the requests that finished.  A request that arrives
// while others are on the GPU therefore joins the very next batch of steps.
// Admission is FIFO by default: when the head of the queue does not fit, later
// requests wait too (no overtaking), so latency order follows arrival order.
// The shortest-first order (logo2opt_server_set_order) admits the request with
// the fewest cities whose start tour is ready among the first 1,024 queued; a
// request that later arrivals have overtaken max_bypass times is admitted
// next, so no request waits behind an unbounded stream of smaller ones.
//
// The reference has no serving layer; its closest relative is the multi-start
// ILS driver with one solver per pthread (src/solvers/ILSGlobalSolverMT.cpp:86-195).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/logo2opt.h"

namespace {

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum req_state { Q_QUEUED = 0, Q_RUNNING = 1, Q_DONE = 2, Q_REJECTED = 3 };

struct build_job;

struct request {
    int64_t id = 0;
    uint32_t n = 0;
    std::vector<logo2opt_city_coords> coords;
    std::vector<uint32_t> route;
    logo2opt_job job{};
    double t_submit = 0, t_admit = 0, t_done = 0;
    int32_t slot = -1;
    int state = Q_QUEUED;
    int err = LOGO2OPT_OK;  // Q_REJECTED: why
    double t_construct = 0;  // seconds the server spent building its start tour
    uint32_t bypassed = 0;   // later arrivals admitted before it (shortest-first order)
    // no start tour given: the server's builder thread makes the greedy NN
    // tour while the GPU runs other requests; admission waits for it
    std::shared_ptr<build_job> build;
    logo2opt_result res{};
};

// A start tour for the builder thread: inputs are fixed at submission, the
// outputs are published by `ready` (release / acquire).
struct build_job {
    std::vector<logo2opt_city_coords> coords;
    std::vector<uint32_t> route;
    int rc = LOGO2OPT_OK;
    double seconds = 0;
    std::atomic<bool> ready{false};
};

// Greedy nearest-neighbour tour from city 0 (the serving workload's start)
// over the coordinates translated so the minimum is 0.  Invalid coordinates
// are left to engine_add to reject.
void build_nn(build_job &b)
{
    const double t0 = now_s();
    const uint32_t n = (uint32_t)b.coords.size();
    double mnx = b.coords[0].x, mny = b.coords[0].y;
    for (uint32_t k = 1; k < n; ++k) {
        mnx = std::min(mnx, (double)b.coords[k].x);
        mny = std::min(mny, (double)b.coords[k].y);
    }
    std::vector<int32_t> xy(2 * (size_t)n);
    b.rc = LOGO2OPT_OK;
    for (uint32_t k = 0; k < n && b.rc == LOGO2OPT_OK; ++k) {
        const double x = (double)b.coords[k].x - mnx, y = (double)b.coords[k].y - mny;
        if (!(x >= 0 && x < 2147483648.0 && y >= 0 && y < 2147483648.0)) b.rc = LOGO2OPT_ECOORD;
        xy[2 * k] = (int32_t)x;
        xy[2 * k + 1] = (int32_t)y;
    }
    if (b.rc == LOGO2OPT_OK) {
        b.route.resize(n);
        b.rc = logo2opt_nn_tour(n, xy.data(), 0, b.route.data());
    }
    b.seconds = now_s() - t0;
    b.ready.store(true, std::memory_order_release);
}

}  // namespace

struct logo2opt_server {
    logo2opt_engine *e = nullptr;
    uint32_t queue_cap = 0;
    // admission limits beyond HBM and slots: running requests, and the sum
    // of candidate moves per sweep over running requests (0 = no limit).  A
    // saturated GPU gains nothing from more concurrent requests; they only
    // stretch every request's latency (processor sharing).
    uint32_t max_running = 0;
    uint64_t max_inflight_pairs = 0;
    uint64_t inflight_pairs = 0;
    // admission order: LOGO2OPT_ORDER_FIFO or LOGO2OPT_ORDER_SHORTEST (with
    // its overtaking bound, 0 = none)
    int order = LOGO2OPT_ORDER_FIFO;
    uint32_t max_bypass = 0;
    int64_t next_id = 1;
    std::deque<int64_t> queue;
    std::map<int64_t, request> reqs;
    std::map<int32_t, int64_t> slot_req;
    std::vector<int64_t> completed;  // not yet polled
    uint64_t n_submitted = 0, n_rejected = 0, n_admitted = 0, n_done = 0;
    // builder thread for start tours (started with the first request that
    // has none), FIFO like admission
    std::thread builder;
    std::mutex bmu;
    std::condition_variable bcv;
    std::deque<std::shared_ptr<build_job>> bq;
    bool bstop = false;

    // Queue position to admit next, or -1 when it has to wait (FIFO: the
    // head's start tour is not built yet; shortest-first: no start tour is).
    // Shortest-first looks at the first ORDER_WINDOW queued requests, so an
    // admission costs at most that many lookups however long the queue is.
    // Ties in size go to the earlier arrival.
    static constexpr size_t ORDER_WINDOW = 1024;
    ptrdiff_t pick()
    {
        const request &head = reqs[queue.front()];
        const bool head_ready = !head.build || head.build->ready.load(std::memory_order_acquire);
        if (order == LOGO2OPT_ORDER_FIFO || (max_bypass && head.bypassed >= max_bypass))
            return head_ready ? 0 : -1;
        ptrdiff_t best = -1;
        uint32_t best_n = 0;
        for (size_t q = 0; q < std::min(queue.size(), ORDER_WINDOW); ++q) {
            const request &r = reqs[queue[q]];
            if (r.build && !r.build->ready.load(std::memory_order_acquire)) continue;
            if (best < 0 || r.n < best_n) {
                best = (ptrdiff_t)q;
                best_n = r.n;
            }
        }
        return best;
    }
    // remove queue position q; the requests ahead of it were overtaken
    void take(ptrdiff_t q)
    {
        for (ptrdiff_t k = 0; k < q; ++k) reqs[queue[k]].bypassed++;
        queue.erase(queue.begin() + q);
    }

    void build_loop()
    {
        for (;;) {
            std::shared_ptr<build_job> j;
            {
                std::unique_lock<std::mutex> lk(bmu);
                bcv.wait(lk, [&] { return bstop || !bq.empty(); });
                if (bstop) return;
                j = bq.front();
                bq.pop_front();
            }
            build_nn(*j);
        }
    }
    void enqueue_build(std::shared_ptr<build_job> j)
    {
        std::lock_guard<std::mutex> lk(bmu);
        if (!builder.joinable()) builder = std::thread([this] { build_loop(); });
        bq.push_back(std::move(j));
        bcv.notify_one();
    }
    ~logo2opt_server()
    {
        {
            std::lock_guard<std::mutex> lk(bmu);
            bstop = true;
        }
        bcv.notify_all();
        if (builder.joinable()) builder.join();
    }
};

extern "C" {

int logo2opt_server_create(int device, uint64_t hbm_budget_bytes, uint32_t max_slots, uint32_t queue_cap,
                           logo2opt_server **out)
{
    if (!out || max_slots == 0) return LOGO2OPT_EINVAL;
    logo2opt_server *s = new (std::nothrow) logo2opt_server();
    if (!s) return LOGO2OPT_ENOMEM;
    int rc = logo2opt_engine_create(device, hbm_budget_bytes, max_slots, &s->e);
    if (rc) {
        delete s;
        return rc;
    }
    s->queue_cap = queue_cap ? queue_cap : 1u << 20;
    *out = s;
    return LOGO2OPT_OK;
}

int logo2opt_server_destroy(logo2opt_server *s)
{
    if (!s) return LOGO2OPT_EINVAL;
    if (s->e) logo2opt_engine_destroy(s->e);
    delete s;
    return LOGO2OPT_OK;
}

int logo2opt_server_submit(logo2opt_server *s, uint32_t n, const logo2opt_city_coords *coords, const uint32_t *route,
                           const logo2opt_job *job, int64_t *id)
{
    if (!s || !coords || !id || n < 3 || n > LOGO2OPT_MAX_N) return LOGO2OPT_EINVAL;
    s->n_submitted++;
    if (s->queue.size() >= s->queue_cap) {
        s->n_rejected++;
        *id = -1;
        return LOGO2OPT_EAGAIN;
    }
    request r;
    r.id = s->next_id++;
    r.n = n;
    r.coords.assign(coords, coords + n);
    if (route) {
        r.route.assign(route, route + n);
    } else {
        r.build = std::make_shared<build_job>();
        r.build->coords = r.coords;
    }
    if (job) r.job = *job;
    r.t_submit = now_s();
    *id = r.id;
    s->queue.push_back(r.id);
    std::shared_ptr<build_job> j = r.build;
    s->reqs.emplace(r.id, std::move(r));
    if (j) s->enqueue_build(std::move(j));
    return LOGO2OPT_OK;
}

// admit -> run up to `steps` batched steps -> harvest
int logo2opt_server_pump(logo2opt_server *s, uint32_t steps, uint32_t sync_every, uint32_t *running,
                         uint32_t *queued)
{
    if (!s) return LOGO2OPT_EINVAL;
    while (!s->queue.empty()) {
        if (s->max_running && s->slot_req.size() >= s->max_running) break;
        const ptrdiff_t pos = s->pick();
        if (pos < 0) break;  // its start tour is not built yet
        request &r = s->reqs[s->queue[pos]];
        const uint64_t pairs = r.n >= 4 ? (uint64_t)(r.n - 2) * (r.n - 3) / 2 : 0;
        if (s->max_inflight_pairs && !s->slot_req.empty() && s->inflight_pairs + pairs > s->max_inflight_pairs)
            break;
        int32_t slot = -1;
        int rc = LOGO2OPT_OK;
        if (r.build) {
            rc = r.build->rc;
            r.route.swap(r.build->route);
            r.t_construct = r.build->seconds;
            r.build.reset();
        }
        if (rc == LOGO2OPT_OK) rc = logo2opt_engine_add(s->e, r.n, r.coords.data(), r.route.data(), &r.job, &slot);
        if (rc == LOGO2OPT_EAGAIN) break;  // the pick does not fit yet
        s->take(pos);
        if (rc) {  // invalid, or larger than the whole budget: finished with an error
            r.state = Q_REJECTED;
            r.err = rc;
            r.t_done = now_s();
            s->n_rejected++;
            s->completed.push_back(r.id);
            std::vector<logo2opt_city_coords>().swap(r.coords);
            std::vector<uint32_t>().swap(r.route);
            continue;
        }
        r.slot = slot;
        r.state = Q_RUNNING;
        r.t_admit = now_s();
        s->slot_req[slot] = r.id;
        s->n_admitted++;
        s->inflight_pairs += pairs;
        // host copies are no longer needed: the instance is resident in HBM
        std::vector<logo2opt_city_coords>().swap(r.coords);
    }
    uint32_t live = 0;
    if (!s->slot_req.empty()) {
        const int rc = logo2opt_engine_run(s->e, steps, sync_every, &live);
        if (rc) return rc;
    }
    int32_t done[256];
    for (;;) {
        uint32_t cnt = 0;
        int rc = logo2opt_engine_poll_done(s->e, done, 256, &cnt);
        if (rc) return rc;
        if (cnt == 0) break;
        const double t = now_s();
        for (uint32_t q = 0; q < cnt; ++q) {
            auto it = s->slot_req.find(done[q]);
            if (it == s->slot_req.end()) continue;
            request &r = s->reqs[it->second];
            rc = logo2opt_engine_result(s->e, done[q], r.route.data(), &r.res, nullptr);
            if (rc) return rc;
            rc = logo2opt_engine_remove(s->e, done[q]);
            if (rc) return rc;
            r.state = Q_DONE;
            r.t_done = t;
            r.slot = -1;
            s->inflight_pairs -= r.n >= 4 ? (uint64_t)(r.n - 2) * (r.n - 3) / 2 : 0;
            s->slot_req.erase(it);
            s->completed.push_back(r.id);
            s->n_done++;
        }
    }
    if (running) *running = (uint32_t)s->slot_req.size();
    if (queued) *queued = (uint32_t)s->queue.size();
    return LOGO2OPT_OK;
}

int logo2opt_server_poll(logo2opt_server *s, int64_t *ids, uint32_t cap, uint32_t *count)
{
    if (!s || !count) return LOGO2OPT_EINVAL;
    const uint32_t c = (uint32_t)std::min<size_t>(cap, s->completed.size());
    for (uint32_t q = 0; q < c; ++q) ids[q] = s->completed[q];
    s->completed.erase(s->completed.begin(), s->completed.begin() + c);
    *count = c;
    return LOGO2OPT_OK;
}

int logo2opt_server_result(logo2opt_server *s, int64_t id, uint32_t *route_out, logo2opt_result *res,
                           double *latency_s, double *queue_s)
{
    if (!s) return LOGO2OPT_EINVAL;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return LOGO2OPT_ENOTFOUND;
    request &r = it->second;
    if (r.state == Q_REJECTED) {
        if (latency_s) *latency_s = r.t_done - r.t_submit;
        if (queue_s) *queue_s = -1.0;
        return r.err;
    }
    if (r.state != Q_DONE) return LOGO2OPT_ESTATE;
    if (route_out) std::memcpy(route_out, r.route.data(), 4ull * r.n);
    if (res) *res = r.res;
    if (latency_s) *latency_s = r.t_done - r.t_submit;
    if (queue_s) *queue_s = r.t_admit - r.t_submit;
    return LOGO2OPT_OK;
}

int logo2opt_server_release(logo2opt_server *s, int64_t id)
{
    if (!s) return LOGO2OPT_EINVAL;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return LOGO2OPT_ENOTFOUND;
    if (it->second.state == Q_RUNNING) {
        const int rc = logo2opt_engine_remove(s->e, it->second.slot);
        if (rc) return rc;
        s->slot_req.erase(it->second.slot);
        const uint32_t n = it->second.n;
        s->inflight_pairs -= n >= 4 ? (uint64_t)(n - 2) * (n - 3) / 2 : 0;
    } else if (it->second.state == Q_QUEUED) {
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), id));
    }
    s->reqs.erase(it);
    return LOGO2OPT_OK;
}

int logo2opt_server_set_limits(logo2opt_server *s, uint32_t max_running, uint64_t max_inflight_pairs)
{
    if (!s) return LOGO2OPT_EINVAL;
    s->max_running = max_running;
    s->max_inflight_pairs = max_inflight_pairs;
    return LOGO2OPT_OK;
}

int logo2opt_server_set_order(logo2opt_server *s, int order, uint32_t max_bypass)
{
    if (!s || (order != LOGO2OPT_ORDER_FIFO && order != LOGO2OPT_ORDER_SHORTEST)) return LOGO2OPT_EINVAL;
    s->order = order;
    s->max_bypass = max_bypass;
    return LOGO2OPT_OK;
}

int logo2opt_server_set_pruning(logo2opt_server *s, int mode)
{
    if (!s) return LOGO2OPT_EINVAL;
    return logo2opt_engine_set_pruning(s->e, mode);
}

int logo2opt_server_counts(logo2opt_server *s, uint64_t *counts)
{
    if (!s || !counts) return LOGO2OPT_EINVAL;
    counts[0] = s->n_submitted;
    counts[1] = s->n_rejected;
    counts[2] = s->n_admitted;
    counts[3] = s->n_done;
    counts[4] = s->queue.size();
    counts[5] = s->slot_req.size();
    return LOGO2OPT_OK;
}

// Open-loop replay: request q arrives arrival_s[q] seconds after the start
// (ascending) and is submitted then; the loop pumps `steps` steps at a time
// until every request finished or `horizon_s` elapsed.  latency_out[q] is the
// time from the scheduled arrival to completion (-1: rejected or unfinished),
// queue_out[q] the time from arrival to admission.
int logo2opt_server_replay(logo2opt_server *s, uint32_t count, const double *arrival_s, const uint32_t *n,
                           const logo2opt_city_coords *const *coords, const uint32_t *const *routes,
                           const logo2opt_job *jobs, double horizon_s, uint32_t steps, uint32_t sync_every,
                           double *latency_out, double *queue_out, double *elapsed_s)
{
    if (!s || !arrival_s || !n || !coords || !routes || !jobs || !latency_out) return LOGO2OPT_EINVAL;
    std::vector<int64_t> ids(count, -1);
    std::map<int64_t, uint32_t> idx;
    for (uint32_t q = 0; q < count; ++q) {
        latency_out[q] = -1.0;
        if (queue_out) queue_out[q] = -1.0;
    }
    const double t0 = now_s();
    uint32_t next = 0, finished = 0;
    while (true) {
        const double t = now_s() - t0;
        if (t > horizon_s) break;
        while (next < count && arrival_s[next] <= t) {
            int64_t id = -1;
            const int rc = logo2opt_server_submit(s, n[next], coords[next], routes[next], &jobs[next], &id);
            if (rc == LOGO2OPT_OK) {
                ids[next] = id;
                idx[id] = next;
            } else if (rc != LOGO2OPT_EAGAIN) {
                return rc;
            } else {
                ++finished;  // rejected at the door (queue full)
            }
            ++next;
        }
        uint32_t running = 0, queued = 0;
        int rc = logo2opt_server_pump(s, steps, sync_every, &running, &queued);
        if (rc) return rc;
        int64_t done[256];
        for (;;) {
            uint32_t c = 0;
            rc = logo2opt_server_poll(s, done, 256, &c);
            if (rc) return rc;
            if (!c) break;
            for (uint32_t k = 0; k < c; ++k) {
                auto it = idx.find(done[k]);
                if (it == idx.end()) continue;
                const request &r = s->reqs[done[k]];
                if (r.state == Q_REJECTED) {  // never admitted: latency stays -1
                    ++finished;
                    continue;
                }
                latency_out[it->second] = (r.t_done - t0) - arrival_s[it->second];
                if (queue_out) queue_out[it->second] = (r.t_admit - t0) - arrival_s[it->second];
                ++finished;
            }
        }
        if (finished >= count && next >= count) break;
        if (running == 0 && queued == 0 && next < count) {
            const double wait = arrival_s[next] - (now_s() - t0);
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
        } else if (running == 0 && queued > 0) {
            // the head of the queue waits for its start tour
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
    }
    if (elapsed_s) *elapsed_s = now_s() - t0;
    return LOGO2OPT_OK;
}

}  // extern "C"
