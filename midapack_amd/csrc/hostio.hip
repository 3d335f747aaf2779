// hostio.hip -- large copies between the caller's (pageable) host arrays and
// the device.
//
// MLmap's arrays (pointing, weights, signal, noise) are ordinary pageable host
// memory; a plain hipMemcpy from pageable memory runs at a fraction of the PCIe
// rate.  Here a pool of pinned staging buffers is shared by a few host worker
// threads: each worker copies its chunks of the caller's array into its own
// pinned buffers (two in flight) and queues the DMA on its own stream, so the
// host copies of one worker overlap the DMA of the others.  The pool is
// allocated once per process (hipHostMalloc; freed by mappraiser_gpu_finalize)
// and used by one transfer at a time (others wait for it).  A staged transfer
// is complete when the call returns: the caller's stream never waits on the
// pool's events, which the next transfer (possibly another host thread's)
// records again as soon as the pool is released.
#include "mgpu_internal.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <thread>

namespace mgpu {

namespace {

constexpr size_t kStageChunk = (size_t) 12 << 20;  // bytes per chunk (a multiple of 8, 12 and 16)
constexpr size_t kStageMin = (size_t) 48 << 20;    // smaller copies go through hipMemcpyAsync
constexpr int kMaxWorkers = 8;
constexpr int kMaxDevices = 16;

struct StagePool {
    std::mutex busy;       // held for the whole of one staged transfer
    std::mutex init_mu;
    int workers = 0;
    unsigned char *pinned = nullptr;  // workers * 2 * kStageChunk
    hipStream_t ws[kMaxDevices][kMaxWorkers] = {};
    hipEvent_t ev[kMaxDevices][kMaxWorkers][2] = {};
    hipEvent_t done[kMaxDevices][kMaxWorkers] = {};
    bool dev_ready[kMaxDevices] = {};
};

StagePool g_pool;

unsigned char *buffer(int w, int b) { return g_pool.pinned + ((size_t) w * 2 + b) * kStageChunk; }

// pinned buffers (once per process) and the device's streams and events (once per device)
int pool_ready(int device) {
    std::lock_guard<std::mutex> g(g_pool.init_mu);
    if (!g_pool.pinned) {
        const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
        g_pool.workers = (int) std::min<unsigned>(kMaxWorkers, hc);
        void *p = nullptr;
        MGPU_HIP(hipHostMalloc(&p, (size_t) g_pool.workers * 2 * kStageChunk, hipHostMallocPortable));
        g_pool.pinned = (unsigned char *) p;
    }
    MGPU_ARG(device >= 0 && device < kMaxDevices, "device %d: host staging supports %d devices", device,
             kMaxDevices);
    if (!g_pool.dev_ready[device]) {
        for (int w = 0; w < g_pool.workers; w++) {
            MGPU_HIP(hipStreamCreateWithFlags(&g_pool.ws[device][w], hipStreamNonBlocking));
            for (int b = 0; b < 2; b++)
                MGPU_HIP(hipEventCreateWithFlags(&g_pool.ev[device][w][b], hipEventDisableTiming));
            MGPU_HIP(hipEventCreateWithFlags(&g_pool.done[device][w], hipEventDisableTiming));
        }
        g_pool.dev_ready[device] = true;
    }
    return MGPU_OK;
}

// run fn(w) on `workers` threads (w = 0 on the calling thread); the first
// non-zero status is returned (with its error message)
template <typename F>
int run_workers(int device, int workers, F fn) {
    std::vector<int> rc((size_t) workers, MGPU_OK);
    std::vector<std::string> err((size_t) workers);
    std::vector<std::thread> th;
    for (int w = 1; w < workers; w++)
        th.emplace_back([&, w] {
            if (hipSetDevice(device) != hipSuccess) {
                rc[w] = MGPU_EHIP;
                err[w] = "hipSetDevice failed in a staging thread";
                return;
            }
            rc[w] = fn(w);
            if (rc[w]) err[w] = mappraiser_gpu_last_error();
        });
    rc[0] = fn(0);
    if (rc[0]) err[0] = mappraiser_gpu_last_error();
    for (auto &t : th) t.join();
    for (int w = 0; w < workers; w++)
        if (rc[w]) {
            set_error("%s", err[w].c_str());
            return rc[w];
        }
    return MGPU_OK;
}

}  // namespace

int h2d_staged(int device, void *d, const void *h, size_t bytes, hipStream_t s) {
    if (bytes < kStageMin) {
        MGPU_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
        return MGPU_OK;
    }
    std::lock_guard<std::mutex> lk(g_pool.busy);
    MGPU_RET(pool_ready(device));
    const int W = g_pool.workers;
    const size_t nch = (bytes + kStageChunk - 1) / kStageChunk;
    unsigned char *dd = (unsigned char *) d;
    const unsigned char *hh = (const unsigned char *) h;
    MGPU_RET(run_workers(device, W, [&](int w) -> int {
        hipStream_t st = g_pool.ws[device][w];
        int j = 0;
        for (size_t k = (size_t) w; k < nch; k += (size_t) W, j++) {
            const int b = j & 1;
            const size_t off = k * kStageChunk, len = std::min(kStageChunk, bytes - off);
            // the buffer's previous DMA (of this or an earlier transfer) has finished
            MGPU_HIP(hipEventSynchronize(g_pool.ev[device][w][b]));
            std::memcpy(buffer(w, b), hh + off, len);
            MGPU_HIP(hipMemcpyAsync(dd + off, buffer(w, b), len, hipMemcpyHostToDevice, st));
            MGPU_HIP(hipEventRecord(g_pool.ev[device][w][b], st));
        }
        MGPU_HIP(hipEventRecord(g_pool.done[device][w], st));
        return MGPU_OK;
    }));
    // complete before the pool is released (and its events recorded again)
    for (int w = 0; w < W; w++) MGPU_HIP(hipEventSynchronize(g_pool.done[device][w]));
    return MGPU_OK;
}

int h2d_staged_fill(int device, void *d, size_t bytes, const StageProducer &fill) {
    std::lock_guard<std::mutex> lk(g_pool.busy);
    MGPU_RET(pool_ready(device));
    const int W = g_pool.workers;
    const size_t nch = (bytes + kStageChunk - 1) / kStageChunk;
    unsigned char *dd = (unsigned char *) d;
    MGPU_RET(run_workers(device, W, [&](int w) -> int {
        hipStream_t st = g_pool.ws[device][w];
        int j = 0;
        for (size_t k = (size_t) w; k < nch; k += (size_t) W, j++) {
            const int b = j & 1;
            const size_t off = k * kStageChunk, len = std::min(kStageChunk, bytes - off);
            MGPU_HIP(hipEventSynchronize(g_pool.ev[device][w][b]));
            fill(off, buffer(w, b), len);
            MGPU_HIP(hipMemcpyAsync(dd + off, buffer(w, b), len, hipMemcpyHostToDevice, st));
            MGPU_HIP(hipEventRecord(g_pool.ev[device][w][b], st));
        }
        MGPU_HIP(hipEventRecord(g_pool.done[device][w], st));
        return MGPU_OK;
    }));
    for (int w = 0; w < W; w++) MGPU_HIP(hipEventSynchronize(g_pool.done[device][w]));
    return MGPU_OK;
}

int d2h_staged(int device, void *h, const void *d, size_t bytes, hipEvent_t after, const StageConsumer *consume) {
    std::unique_lock<std::mutex> lk(g_pool.busy, std::defer_lock);
    if (bytes >= kStageMin) lk.try_lock();
    if (!lk.owns_lock()) {
        // small or pool busy: one copy, then the consumer over the whole array
        MGPU_HIP(hipEventSynchronize(after));
        if (consume) {
            std::vector<unsigned char> tmp(bytes);
            MGPU_HIP(hipMemcpy(tmp.data(), d, bytes, hipMemcpyDeviceToHost));
            (*consume)(0, tmp.data(), bytes);
        } else {
            MGPU_HIP(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
        }
        return MGPU_OK;
    }
    MGPU_RET(pool_ready(device));
    const int W = g_pool.workers;
    const size_t nch = (bytes + kStageChunk - 1) / kStageChunk;
    const unsigned char *dd = (const unsigned char *) d;
    unsigned char *hh = (unsigned char *) h;
    MGPU_HIP(hipEventSynchronize(after));  // the pool's streams wait on no other stream
    return run_workers(device, W, [&](int w) -> int {
        hipStream_t st = g_pool.ws[device][w];
        auto issue = [&](size_t k, int b) -> int {
            const size_t off = k * kStageChunk, len = std::min(kStageChunk, bytes - off);
            MGPU_HIP(hipEventSynchronize(g_pool.ev[device][w][b]));
            MGPU_HIP(hipMemcpyAsync(buffer(w, b), dd + off, len, hipMemcpyDeviceToHost, st));
            MGPU_HIP(hipEventRecord(g_pool.ev[device][w][b], st));
            return MGPU_OK;
        };
        size_t k = (size_t) w;
        int j = 0;
        if (k < nch) MGPU_RET(issue(k, 0));
        for (; k < nch; k += (size_t) W, j++) {
            const int b = j & 1;
            if (k + W < nch) MGPU_RET(issue(k + W, b ^ 1));  // next chunk in flight meanwhile
            MGPU_HIP(hipEventSynchronize(g_pool.ev[device][w][b]));
            const size_t off = k * kStageChunk, len = std::min(kStageChunk, bytes - off);
            if (consume)
                (*consume)(off, buffer(w, b), len);
            else
                std::memcpy(hh + off, buffer(w, b), len);
        }
        return MGPU_OK;
    });
}

void staging_free() {
    std::lock_guard<std::mutex> b(g_pool.busy);
    std::lock_guard<std::mutex> g(g_pool.init_mu);
    int cur = 0;
    (void) hipGetDevice(&cur);
    for (int dv = 0; dv < kMaxDevices; dv++) {
        if (!g_pool.dev_ready[dv]) continue;
        if (hipSetDevice(dv) != hipSuccess) continue;
        for (int w = 0; w < g_pool.workers; w++) {
            (void) hipStreamSynchronize(g_pool.ws[dv][w]);
            (void) hipStreamDestroy(g_pool.ws[dv][w]);
            for (int b = 0; b < 2; b++) (void) hipEventDestroy(g_pool.ev[dv][w][b]);
            (void) hipEventDestroy(g_pool.done[dv][w]);
        }
        g_pool.dev_ready[dv] = false;
    }
    (void) hipSetDevice(cur);
    if (g_pool.pinned) (void) hipHostFree(g_pool.pinned);
    g_pool.pinned = nullptr;
}

}  // namespace mgpu
