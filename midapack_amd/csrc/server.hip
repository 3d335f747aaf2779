// server.hip -- map-making request loop with HBM-budget admission and
// continuous batching.
//
// Requests (mgpu_batch_item, host arrays owned by the caller until wait
// returns) enter a FIFO queue.  A dispatcher thread admits the queue head as
// soon as a slot is free and its estimated device footprint fits the HBM
// budget that running solves leave; with it, it admits every later queued
// request that still fits, up to max_batch, whatever their sizes, bands,
// grids, tolerances and iteration caps.  An admitted group runs as one
// segmented batch solve (batch.hip) on a host thread and stream of its own;
// each request is answered as soon as its own segment has stopped (the rest of
// its group iterates on), and when the group finishes its memory and slot
// return to the pool and the queue is scanned again.  One group runs on the
// device at a time (max_slots must be 1): with several concurrent groups the
// loop stopped answering under sustained load (cause not found), so that mode
// is refused.  Each result is bit-identical to the
// request's solo solve: batching changes neither the kernels nor the
// summation order of an item.
//
// At creation the server returns the library's cached blocks to the device and
// reserves 80% of the free device memory as one pool; every solve allocates
// from it, so no worker thread allocates or frees device memory through the
// HIP runtime (whose frees synchronise the device) while other solves run.
// The admission budget is the pool, or less when the caller sets one.
//
// Footprints are estimated per sample and per grid pixel (a bound from the
// allocations of a solve) and, once groups have finished, raised by the
// largest ratio of what a group actually held (its arena's high-water mark)
// to its estimate, plus 5%, when that exceeds 1.  A group whose allocations fail while other
// solves hold memory goes back to the queue head; such requests are later
// admitted one at a time and only onto an idle device, where a second failure
// is final.
#include "mgpu_internal.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace mgpu {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    long long ticket = 0;
    mgpu_batch_item *item = nullptr;
    double tol = 1e-8;
    int maxiter = 500;
    int grid = 0;
    int64_t est = 0;    // footprint estimate before the learned correction
    int64_t bytes = 0;  // bytes charged to the budget while admitted
    int retries = 0;    // times sent back to the queue after an allocation failure
    Clock::time_point t_submit, t_start, t_done;
    bool done = false;
    int status = MGPU_OK;
    std::string err;
};

struct Worker {
    std::thread th;
    bool finished = false;
};

}  // namespace

struct Server {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request *> pending;
    std::unordered_map<long long, std::unique_ptr<Request>> reqs;
    std::list<Worker> workers;
    int64_t budget = 0, used = 0;
    double correction = 0.0;  // max over finished groups of 1.05 x (measured peak / estimate); 0: none yet
                              // (it only raises the estimates: below 1 they stay as they are)
    int max_slots = 1, max_batch = 16, running = 0;
    int device = 0;  // the device current when the server was created
    DevPool *pool = nullptr;  // the HBM budget, reserved at creation; every solve allocates from it
    long long next_ticket = 1;
    bool stop = false;
    mgpu_options opt{};
    std::thread dispatcher;
    // statistics
    long long admitted = 0, groups = 0, requeued = 0;
    int64_t peak_used = 0;
    int peak_running = 0;

    void run_dispatcher();
    void run_group(std::vector<Request *> grp, Worker *self);
};

// Device bytes of one solve: inputs staged on the device (pixels, weights,
// signal, noise), pointing tables (q, Q/U weights, sliced ELL with its sort
// keys and scan temporaries during the build), TOD work vectors in both
// precisions, and the map work vectors and 3x3 blocks per pixel of the grid.
static int64_t request_bytes(const mgpu_batch_item &it, int grid, bool mixed) {
    const int64_t per_sample = 160 + (mixed ? 32 : 0);
    const int64_t per_pixel = 8 * 3 * 9 + 8 * 9 + 8 + 4 + 4 * 4;
    return (int64_t) it.m * per_sample + (int64_t) std::min<int64_t>(grid, it.m) * per_pixel + (64 << 20);
}

// Limits of one merged group: the batch solve indexes samples and map values
// with int32 (each item padded to whole chunks).
constexpr int64_t kGroupSamples = (int64_t) 1 << 30;
constexpr int64_t kGroupPixels = (int64_t) 1 << 28;

void Server::run_group(std::vector<Request *> grp, Worker *self) {
    const int n = (int) grp.size();
    std::vector<mgpu_batch_item> items((size_t) n);
    std::vector<int> all((size_t) n), grid((size_t) n), mx((size_t) n);
    std::vector<double> tl((size_t) n);
    for (int i = 0; i < n; i++) {
        items[i] = *grp[i]->item;
        all[i] = i;
        grid[i] = grp[i]->grid;
        tl[i] = grp[i]->tol;
        mx[i] = grp[i]->maxiter;
    }
    const Clock::time_point t0 = Clock::now();
    int64_t group_bytes = 0, group_est = 0;  // read now: answered requests may be freed by their waiters
    for (Request *r : grp) {
        r->t_start = t0;
        group_bytes += r->bytes;
        group_est += r->est;
    }
    // A request is answered as soon as its own segment has stopped.  Once
    // answered, its waiter may collect it and free it: from then on this
    // thread never touches that request again (answered[i] is this thread's own record).
    std::vector<char> answered((size_t) n, 0);
    const std::function<void(int)> done = [&](int i) {
        const Clock::time_point t = Clock::now();
        std::lock_guard<std::mutex> g(mu);
        Request *r = grp[i];
        r->item->rep = items[i].rep;
        r->status = MGPU_OK;
        r->t_done = t;
        r->done = true;
        answered[i] = 1;
        cv.notify_all();
    };
    int rc = MGPU_OK;
    struct Bind {  // this thread's device allocations come from the server's pool
        explicit Bind(DevPool *pl) {
            dev_pool_bind(pl);
            serving_bind(true);
        }
        ~Bind() {
            dev_pool_bind(nullptr);
            serving_bind(false);
        }
    } bind(pool);
    if (hipSetDevice(device) != hipSuccess) {
        rc = MGPU_EHIP;
        set_error("hipSetDevice(%d) failed in a server worker", device);
    } else {
        for (int i = 0; i < n && rc == MGPU_OK; i++) rc = check_item(items[i], i);
        if (rc == MGPU_OK) rc = pcg_batch_impl(items.data(), all, grid, tl, mx, false, &opt, &done);
    }
    std::string err = rc == MGPU_OK ? std::string() : std::string(mappraiser_gpu_last_error());
    const Clock::time_point t1 = Clock::now();
    std::lock_guard<std::mutex> g(mu);
    // an allocation failure beside other solves, or in a merged group, is retried
    // later on its own; alone on an idle device it is final
    const bool retry = rc == MGPU_ENOMEM && (running > 1 || n > 1);
    std::vector<Request *> back;
    for (int i = 0; i < n; i++) {
        if (answered[i]) continue;
        Request *r = grp[i];  // not answered yet: its waiter still blocks, the request is alive
        if (retry && r->retries < 2) {
            r->retries++;
            back.push_back(r);
            continue;
        }
        r->item->rep = items[i].rep;
        r->status = rc;
        r->err = err;
        r->t_done = t1;
        r->done = true;
    }
    pending.insert(pending.begin(), back.begin(), back.end());
    requeued += (long long) back.size();
    if (rc == MGPU_OK && n > 0) {
        const double peak = (double) items[0].rep.dev_peak_bytes;
        if (group_est > 0 && peak > 0)
            correction = std::min(4.0, std::max(correction, 1.05 * peak / (double) group_est));
    }
    used -= group_bytes;
    running--;
    self->finished = true;
    cv.notify_all();
}

void Server::run_dispatcher() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        // reap finished worker threads
        for (auto it = workers.begin(); it != workers.end();) {
            if (it->finished) {
                it->th.join();
                it = workers.erase(it);
            } else {
                ++it;
            }
        }
        auto cost = [&](const Request *r) { return (int64_t) ((double) r->est * std::max(1.0, correction)); };
        auto fits = [&](const Request *r) {
            if (r->retries > 0) return running == 0;  // failed beside others: run alone
            return running == 0 || used + cost(r) <= budget;
        };
        if (stop && pending.empty() && running == 0) break;
        if (pending.empty() || running >= max_slots || !fits(pending.front())) {
            cv.wait(lk);
            continue;
        }
        // admit the head and every later request that fits
        std::vector<Request *> grp;
        Request *head = pending.front();
        pending.pop_front();
        head->bytes = cost(head);
        grp.push_back(head);
        used += head->bytes;
        int64_t samples = head->item->m + kChunk, pixels = head->grid;
        for (auto it = pending.begin(); head->retries == 0 && it != pending.end() && (int) grp.size() < max_batch;) {
            Request *r = *it;
            const int64_t b = cost(r);
            if (r->retries == 0 && used + b <= budget && samples + r->item->m + kChunk <= kGroupSamples &&
                pixels + r->grid <= kGroupPixels) {
                r->bytes = b;
                grp.push_back(r);
                used += b;
                samples += r->item->m + kChunk;
                pixels += r->grid;
                it = pending.erase(it);
            } else {
                ++it;
            }
        }
        running++;
        admitted += (long long) grp.size();
        groups++;
        peak_used = std::max(peak_used, used);
        peak_running = std::max(peak_running, running);
        workers.emplace_back();
        Worker *w = &workers.back();
        w->th = std::thread(&Server::run_group, this, grp, w);
    }
    for (auto &w : workers) w.th.join();
    workers.clear();
}

}  // namespace mgpu

using namespace mgpu;

extern "C" {

int mappraiser_gpu_server_create(mgpu_server **out, double hbm_budget_gb, int max_slots, int max_batch,
                                 const mgpu_options *opt) {
    MGPU_ARG(out, "NULL argument");
    MGPU_ARG(max_batch >= 1, "max_batch=%d must be >= 1", max_batch);
    MGPU_ARG(max_slots == 1,
             "max_slots=%d: concurrent groups are disabled (with several groups on the device the request loop "
             "stopped answering under sustained load, cause not found); use max_slots=1",
             max_slots);
    MGPU_ARG(!opt || opt->precond == 0, "precond=%d: the request loop solves with block Jacobi", opt->precond);
    std::unique_ptr<Server> s(new Server());
    mappraiser_gpu_options_default(&s->opt);
    if (opt) s->opt = *opt;
    s->opt.profile = 0;
    // solves of concurrent workers allocate while others run: no graph capture here
    s->opt.graph = 0;
    MGPU_HIP(hipGetDevice(&s->device));
    // the library's cached blocks go back to the device before the budget is reserved
    dev_cache_trim(s->device);
    size_t free_b = 0, total_b = 0;
    MGPU_HIP(hipMemGetInfo(&free_b, &total_b));
    const int64_t reserve = (int64_t) (0.8 * (double) free_b);
    s->budget = hbm_budget_gb > 0 ? std::min(reserve, (int64_t) (hbm_budget_gb * 1e9)) : reserve;
    MGPU_RET(dev_pool_create(&s->pool, (size_t) reserve, s->device));
    s->max_slots = max_slots;
    s->max_batch = max_batch;
    Server *raw = s.get();
    s->dispatcher = std::thread(&Server::run_dispatcher, raw);
    *out = reinterpret_cast<mgpu_server *>(s.release());
    return MGPU_OK;
}

int mappraiser_gpu_server_submit(mgpu_server *h, mgpu_batch_item *item, int grid_npix, double tol, int maxiter,
                                 long long *ticket) {
    MGPU_ARG(h && item && ticket, "NULL argument");
    MGPU_ARG(grid_npix > 0 && maxiter >= 1, "grid_npix=%d / maxiter=%d", grid_npix, maxiter);
    MGPU_ARG(item->local_blocks_sizes && item->nb > 0 && item->m > 0, "invalid request");
    Server &s = *reinterpret_cast<Server *>(h);
    std::unique_ptr<Request> r(new Request());
    r->item = item;
    r->tol = tol;
    r->maxiter = maxiter;
    r->grid = grid_npix;
    r->est = request_bytes(*item, grid_npix, s.opt.precision == MGPU_MIXED);
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> g(s.mu);
    MGPU_ARG(!s.stop, "server is shutting down");
    r->ticket = s.next_ticket++;
    *ticket = r->ticket;
    s.pending.push_back(r.get());
    s.reqs[r->ticket] = std::move(r);
    s.cv.notify_all();
    return MGPU_OK;
}

int mappraiser_gpu_server_wait(mgpu_server *h, long long ticket, double *latency_ms, double *queue_ms) {
    MGPU_ARG(h, "NULL handle");
    Server &s = *reinterpret_cast<Server *>(h);
    std::unique_lock<std::mutex> lk(s.mu);
    auto it = s.reqs.find(ticket);
    MGPU_ARG(it != s.reqs.end(), "unknown ticket %lld", ticket);
    Request *r = it->second.get();
    s.cv.wait(lk, [&] { return r->done; });
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (queue_ms) *queue_ms = std::chrono::duration<double, std::milli>(r->t_start - r->t_submit).count();
    const int rc = r->status;
    const std::string err = r->err;
    s.reqs.erase(it);
    if (rc != MGPU_OK) set_error("%s", err.c_str());
    return rc;
}

int mappraiser_gpu_server_stats(mgpu_server *h, long long *admitted, long long *groups, double *peak_gb,
                                int *peak_running, long long *requeued) {
    MGPU_ARG(h, "NULL handle");
    Server &s = *reinterpret_cast<Server *>(h);
    std::lock_guard<std::mutex> g(s.mu);
    if (admitted) *admitted = s.admitted;
    if (groups) *groups = s.groups;
    if (peak_gb) *peak_gb = (double) s.peak_used / 1e9;
    if (peak_running) *peak_running = s.peak_running;
    if (requeued) *requeued = s.requeued;
    return MGPU_OK;
}

int mappraiser_gpu_server_destroy(mgpu_server *h) {
    if (!h) return MGPU_OK;
    Server *s = reinterpret_cast<Server *>(h);
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    s->dispatcher.join();
    dev_pool_destroy(s->pool);
    delete s;
    return MGPU_OK;
}

}  // extern "C"
