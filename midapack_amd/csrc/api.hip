// api.hip -- the extern "C" boundary of libmappraiser_gpu.so.
//
// Each entry point validates its inputs, runs on the handle's own HIP stream
// and returns an int status; the message of the last failure on the calling
// thread is kept for mappraiser_gpu_last_error().  Nothing here calls exit()
// (the reference exits on allocation or LAPACK failures, mappraiser.c:120-123,
// precond.c:485-488, 577).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "mgpu_internal.h"

#include <thread>

namespace mgpu {

static thread_local char g_err[1024] = "";
// kernel-class times of the drop-in solves run with opt.profile (process-wide)
static std::mutex g_kt_mu;
static double g_kt_ms[MGPU_NKT] = {0};
static long long g_kt_cnt[MGPU_NKT] = {0};

// adds a finished solve's kernel-class times (opt.profile) to the process totals
int kernel_times_accumulate(Problem &pb) {
    if (!pb.prof.on) return MGPU_OK;
    MGPU_HIP(hipStreamSynchronize(pb.stream));
    pb.prof.collect();
    std::lock_guard<std::mutex> g(g_kt_mu);
    for (int i = 0; i < MGPU_NKT; i++) {
        g_kt_ms[i] += pb.prof.ms[i];
        g_kt_cnt[i] += pb.prof.cnt[i];
    }
    return MGPU_OK;
}
static int g_device = 0;

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// ---------------------------------------------------------------------------
// device block cache
// ---------------------------------------------------------------------------
namespace {
std::mutex g_cache_mu;
std::map<int, std::multimap<size_t, void *>> g_cache;  // device -> (class bytes, block)
}  // namespace

static size_t size_class(size_t bytes) {
    const size_t mib = (size_t) 1 << 20;
    if (bytes <= mib) return (bytes + 511) / 512 * 512;
    return (bytes + 2 * mib - 1) / (2 * mib) * (2 * mib);
}

// ---------------------------------------------------------------------------
// reserved device pools
// ---------------------------------------------------------------------------
struct DevPool {
    unsigned char *base = nullptr;
    size_t size = 0;
    int device = 0;
    std::mutex mu;
    std::map<size_t, size_t> free_;  // offset -> bytes, ordered by address
};

namespace {
std::mutex g_pools_mu;
std::vector<DevPool *> g_pools;
thread_local DevPool *t_pool = nullptr;
constexpr size_t kPoolAlign = 256;
}  // namespace

int dev_pool_create(DevPool **out, size_t bytes, int device) {
    bytes = (bytes + kPoolAlign - 1) / kPoolAlign * kPoolAlign;
    std::unique_ptr<DevPool> pl(new DevPool());
    void *b = nullptr;
    const hipError_t e = hipMalloc(&b, bytes);
    if (e != hipSuccess) {
        (void) hipGetLastError();
        set_error("reserving %zu device bytes failed: %s", bytes, hipGetErrorString(e));
        return MGPU_ENOMEM;
    }
    pl->base = (unsigned char *) b;
    pl->size = bytes;
    pl->device = device;
    pl->free_[0] = bytes;
    std::lock_guard<std::mutex> g(g_pools_mu);
    g_pools.push_back(pl.get());
    *out = pl.release();
    return MGPU_OK;
}

void dev_pool_destroy(DevPool *pl) {
    if (!pl) return;
    {
        std::lock_guard<std::mutex> g(g_pools_mu);
        g_pools.erase(std::remove(g_pools.begin(), g_pools.end(), pl), g_pools.end());
    }
    (void) hipFree(pl->base);
    delete pl;
}

void dev_pool_bind(DevPool *pl) { t_pool = pl; }

namespace {
thread_local bool t_serving = false;
}
void serving_bind(bool on) { t_serving = on; }
bool serving_thread() { return t_serving; }

// first fit in address order
static int pool_alloc(DevPool &pl, void **p, size_t bytes, size_t *cls) {
    const size_t need = std::max(kPoolAlign, (bytes + kPoolAlign - 1) / kPoolAlign * kPoolAlign);
    std::lock_guard<std::mutex> g(pl.mu);
    for (auto it = pl.free_.begin(); it != pl.free_.end(); ++it) {
        if (it->second < need) continue;
        const size_t off = it->first, rest = it->second - need;
        pl.free_.erase(it);
        if (rest) pl.free_[off + need] = rest;
        *p = pl.base + off;
        *cls = need;
        return MGPU_OK;
    }
    *p = nullptr;
    set_error("device pool of %zu bytes: no free range of %zu bytes", pl.size, need);
    return MGPU_ENOMEM;
}

// returns false if p is not in a pool
static bool pool_free(void *p, size_t cls) {
    std::lock_guard<std::mutex> g(g_pools_mu);
    for (DevPool *pl : g_pools) {
        unsigned char *q = (unsigned char *) p;
        if (q < pl->base || q >= pl->base + pl->size) continue;
        std::lock_guard<std::mutex> h(pl->mu);
        size_t off = (size_t) (q - pl->base), len = cls;
        auto next = pl->free_.lower_bound(off);
        if (next != pl->free_.end() && off + len == next->first) {  // merge with the following range
            len += next->second;
            next = pl->free_.erase(next);
        }
        if (next != pl->free_.begin()) {  // and with the preceding one
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += len;
                return true;
            }
        }
        pl->free_[off] = len;
        return true;
    }
    return false;
}

int dev_alloc(void **p, size_t bytes, int device, size_t *cls) {
    if (t_pool) return pool_alloc(*t_pool, p, bytes, cls);
    const size_t c = size_class(bytes);
    *cls = c;
    {
        // smallest cached block of at least c bytes, if no more than 1/8 larger
        // (blocks keep their own size class)
        std::lock_guard<std::mutex> g(g_cache_mu);
        auto &fl = g_cache[device];
        auto it = fl.lower_bound(c);
        if (it != fl.end() && it->first <= c + c / 8) {
            *cls = it->first;
            *p = it->second;
            fl.erase(it);
            return MGPU_OK;
        }
    }
    hipError_t e = hipMalloc(p, c);
    if (e != hipSuccess) {
        (void) hipGetLastError();
        dev_cache_trim(device);
        e = hipMalloc(p, c);
    }
    if (e != hipSuccess) {
        (void) hipGetLastError();
        *p = nullptr;
        set_error("hipMalloc of %zu bytes failed: %s", c, hipGetErrorString(e));
        return MGPU_ENOMEM;
    }
    return MGPU_OK;
}

void dev_free(void *p, size_t cls, int device) {
    if (!p) return;
    if (pool_free(p, cls)) return;
    std::lock_guard<std::mutex> g(g_cache_mu);
    g_cache[device].emplace(cls, p);
}

void dev_cache_trim(int device) {
    std::vector<void *> blocks;
    {
        std::lock_guard<std::mutex> g(g_cache_mu);
        for (auto &kv : g_cache)
            if (device < 0 || kv.first == device) {
                for (auto &b : kv.second) blocks.push_back(b.second);
                kv.second.clear();
            }
    }
    for (void *b : blocks) (void) hipFree(b);
}

Problem::~Problem() {
    if (ev0) (void) hipEventDestroy(ev0);
    if (ev1) (void) hipEventDestroy(ev1);
    if (stream) {
        (void) hipStreamSynchronize(stream);
        (void) hipStreamDestroy(stream);
    }
}

template <typename T>
static int upload(Problem &pb, T **d, const T *h, size_t n) {
    MGPU_RET(pb.arena.alloc(d, n));
    return h2d_staged(pb.device, *d, h, n * sizeof(T), pb.stream);
}

int new_problem(std::unique_ptr<Problem> &pb, const mgpu_options *opt) {
    pb.reset(new Problem());
    pb->device = g_device;
    pb->arena.device = g_device;
    MGPU_HIP(hipSetDevice(g_device));
    MGPU_HIP(hipStreamCreateWithFlags(&pb->stream, hipStreamNonBlocking));
    pb->arena.stream = pb->stream;
    pb->cub_tmp.arena = &pb->arena;
    mappraiser_gpu_options_default(&pb->opt);
    if (opt) pb->opt = *opt;
    pb->prof.on = pb->opt.profile != 0;
    MGPU_HIP(hipEventCreate(&pb->ev0));
    MGPU_HIP(hipEventCreate(&pb->ev1));
    MGPU_ARG(pb->opt.precision == MGPU_FP64 || pb->opt.precision == MGPU_MIXED, "unknown precision %d",
             pb->opt.precision);
    return MGPU_OK;
}

static int elapsed_ms(Problem &pb, double *ms) {
    MGPU_HIP(hipEventSynchronize(pb.ev1));
    float t = 0.f;
    MGPU_HIP(hipEventElapsedTime(&t, pb.ev0, pb.ev1));
    *ms = t;
    return MGPU_OK;
}

int problem_setup(Problem &pb, const int *d_pix, const double *d_w, int64_t m, int nnz, const int *block_sizes,
                  int nb, int lambda, const double *d_invtt, const double *h_invtt) {
    MGPU_HIP(hipEventRecord(pb.ev0, pb.stream));
    MGPU_RET(pointing_build(pb, d_pix, d_w, m, nnz));
    MGPU_RET(toeplitz_plan(pb, block_sizes, nb, lambda, d_invtt, h_invtt));
    MGPU_HIP(hipEventRecord(pb.ev1, pb.stream));
    return elapsed_ms(pb, &pb.t_create_ms);
}

int problem_setup_packed(Problem &pb, const int *d_p, double2 *d_w12, int64_t m, int pmin, int pmax,
                         bool any_flagged, const int *block_sizes, int nb, int lambda, const double *h_invtt) {
    MGPU_HIP(hipEventRecord(pb.ev0, pb.stream));
    MGPU_RET(pointing_build_packed(pb, d_p, d_w12, m, pmin, pmax, any_flagged));
    MGPU_RET(toeplitz_plan(pb, block_sizes, nb, lambda, nullptr, h_invtt));
    MGPU_HIP(hipEventRecord(pb.ev1, pb.stream));
    return elapsed_ms(pb, &pb.t_create_ms);
}

// What the host learned while packing the caller's pointing for upload.
struct PackStats {
    int pmin = INT32_MAX, pmax = INT32_MIN;
    bool bad = false;      // some sample not encoded 3p+k / -3,-2,-1
    bool flagged = false;  // some sample flagged
    bool w0 = false;       // some I weight is not 1
};

// The pointing uploaded packed: per sample its global pixel p (-1 flagged,
// 4 bytes instead of the 12 of pix[3t..3t+2]) and its (Q, U) weights (16
// bytes instead of 24), built in the staging buffers while the encoding and
// the I weights are checked.  Nothing is usable when ps->bad or ps->w0.
static int upload_packed(int dev, const int *pix, const double *w, size_t m, int *d_p, double2 *d_w12,
                         PackStats *ps) {
    std::mutex mu;
    MGPU_RET(h2d_staged_fill(dev, d_p, sizeof(int) * m, [&](size_t off, void *chunk, size_t len) {
        const size_t t0 = off / sizeof(int), n = len / sizeof(int);
        int *o = static_cast<int *>(chunk);
        int lo = INT32_MAX, hi = INT32_MIN;
        bool bad = false, fl = false;
        for (size_t i = 0; i < n; i++) {
            const int *q = pix + 3 * (t0 + i);
            const int a = q[0];
            int p = -1;
            if (a >= 0) {
                if (a % 3 == 0 && q[1] == a + 1 && q[2] == a + 2) {
                    p = a / 3;
                    lo = std::min(lo, p);
                    hi = std::max(hi, p);
                } else {
                    bad = true;
                }
            } else if (a == -3 && q[1] == -2 && q[2] == -1) {
                fl = true;
            } else {
                bad = true;
            }
            o[i] = p;
        }
        std::lock_guard<std::mutex> g(mu);
        ps->pmin = std::min(ps->pmin, lo);
        ps->pmax = std::max(ps->pmax, hi);
        ps->bad = ps->bad || bad;
        ps->flagged = ps->flagged || fl;
    }));
    if (ps->bad) return MGPU_OK;
    return h2d_staged_fill(dev, d_w12, sizeof(double2) * m, [&](size_t off, void *chunk, size_t len) {
        const size_t t0 = off / sizeof(double2), n = len / sizeof(double2);
        double *o = static_cast<double *>(chunk);
        bool w0 = false;
        for (size_t i = 0; i < n; i++) {
            const double *q = w + 3 * (t0 + i);
            w0 = w0 || q[0] != 1.0;
            o[2 * i] = q[1];
            o[2 * i + 1] = q[2];
        }
        if (w0) {
            std::lock_guard<std::mutex> g(mu);
            ps->w0 = true;
        }
    });
}

int precond_build_timed(Problem &pb, double *d_signal) {
    TimerRange tr("Preconditioner computation time");  // pcg_true.c:48-56
    MGPU_HIP(hipEventRecord(pb.ev0, pb.stream));
    MGPU_RET(precond_build(pb, d_signal));
    MGPU_HIP(hipEventRecord(pb.ev1, pb.stream));
    return elapsed_ms(pb, &pb.t_precond_ms);
}

int check_common(int m, int nnz, int nb, const int *bs, int lambda) {
    MGPU_ARG(m > 0, "m=%d must be positive", m);
    MGPU_ARG(nnz == 3, "nnz=%d: only the I,Q,U pointing (nnz=3) is supported", nnz);
    MGPU_ARG(nb > 0 && bs, "nb_blocks=%d / local_blocks_sizes missing", nb);
    MGPU_ARG(lambda >= 1, "lambda=%d must be >= 1", lambda);
    return MGPU_OK;
}

}  // namespace mgpu

using namespace mgpu;

#define PB(h) (reinterpret_cast<Problem *>(h))

extern "C" {

int mappraiser_gpu_init(int device) {
    int n = 0;
    MGPU_HIP(hipGetDeviceCount(&n));
    MGPU_ARG(device >= 0 && device < n, "device %d out of range (%d devices)", device, n);
    MGPU_HIP(hipSetDevice(device));
    g_device = device;
    // address of the allocation-free products' scratch array on this device
    MGPU_RET(noalloc_scratch_init(device));
    return MGPU_OK;
}

int mappraiser_gpu_finalize(void) {
    dev_cache_trim(-1);
    staging_free();
    return MGPU_OK;
}

const char *mappraiser_gpu_last_error(void) { return g_err; }

void mappraiser_gpu_options_default(mgpu_options *opt) {
    if (!opt) return;
    std::memset(opt, 0, sizeof(*opt));
    opt->precision = MGPU_FP64;
    opt->refresh_every = 10;
    opt->tile_log2 = 0;
    opt->poll_every = 8;
}

int mappraiser_gpu_problem_create(mgpu_problem **out, int m, int nnz, const int *pix, const double *pixweights,
                                  int nb, const int *local_blocks_sizes, int lambda, const double *invtt,
                                  const mgpu_options *opt) {
    g_err[0] = 0;
    MGPU_ARG(out && pix && pixweights && invtt, "NULL argument");
    MGPU_RET(check_common(m, nnz, nb, local_blocks_sizes, lambda));
    std::unique_ptr<Problem> pb;
    MGPU_RET(new_problem(pb, opt));
    int *d_pix;
    double *d_w;
    MGPU_RET(upload(*pb, &d_pix, pix, (size_t) m * nnz));
    MGPU_RET(upload(*pb, &d_w, pixweights, (size_t) m * nnz));
    MGPU_RET(problem_setup(*pb, d_pix, d_w, m, nnz, local_blocks_sizes, nb, lambda, nullptr, invtt));
    pb->arena.release(d_pix);
    pb->arena.release(d_w);
    *out = reinterpret_cast<mgpu_problem *>(pb.release());
    return MGPU_OK;
}

int mappraiser_gpu_problem_create_device(mgpu_problem **out, int m, int nnz, const int *d_pix,
                                         const double *d_pixweights, int nb, const int *local_blocks_sizes,
                                         int lambda, const double *d_invtt, const mgpu_options *opt) {
    g_err[0] = 0;
    MGPU_ARG(out && d_pix && d_pixweights && d_invtt, "NULL argument");
    MGPU_RET(check_common(m, nnz, nb, local_blocks_sizes, lambda));
    std::unique_ptr<Problem> pb;
    MGPU_RET(new_problem(pb, opt));
    MGPU_RET(problem_setup(*pb, d_pix, d_pixweights, m, nnz, local_blocks_sizes, nb, lambda, d_invtt, nullptr));
    *out = reinterpret_cast<mgpu_problem *>(pb.release());
    return MGPU_OK;
}

int mappraiser_gpu_problem_destroy(mgpu_problem *h) {
    if (h) {
        (void) hipSetDevice(PB(h)->device);
        delete PB(h);
    }
    return MGPU_OK;
}

int mappraiser_gpu_local_shape(mgpu_problem *h, int *lcount, int *trash_pix, int *lindices, int *indices) {
    MGPU_ARG(h, "NULL handle");
    Problem &pb = *PB(h);
    MGPU_HIP(hipSetDevice(pb.device));
    const int lc = 3 * pb.pt.npix + 3 * pb.pt.trash;
    if (lcount) *lcount = lc;
    if (trash_pix) *trash_pix = pb.pt.trash;
    int *d;
    if (lindices) {
        MGPU_RET(pb.arena.alloc(&d, (size_t) lc));
        MGPU_RET(pointing_lindices(pb, d));
        MGPU_HIP(hipMemcpyAsync(lindices, d, sizeof(int) * lc, hipMemcpyDeviceToHost, pb.stream));
        MGPU_HIP(hipStreamSynchronize(pb.stream));
        pb.arena.release(d);
    }
    if (indices) {
        MGPU_RET(pb.arena.alloc(&d, (size_t) pb.pt.m * 3));
        MGPU_RET(pointing_local_indices(pb, d));
        MGPU_HIP(hipMemcpyAsync(indices, d, sizeof(int) * pb.pt.m * 3, hipMemcpyDeviceToHost, pb.stream));
        MGPU_HIP(hipStreamSynchronize(pb.stream));
        pb.arena.release(d);
    }
    return MGPU_OK;
}

int mappraiser_gpu_matvec(mgpu_problem *h, const double *x, double *y) {
    MGPU_ARG(h && x && y, "NULL argument");
    Problem &pb = *PB(h);
    MGPU_HIP(hipSetDevice(pb.device));
    MGPU_RET(ensure_work(pb));
    const int n = 3 * pb.pt.npix;
    MGPU_HIP(hipMemcpyAsync(pb.d_map[0], x, sizeof(double) * n, hipMemcpyHostToDevice, pb.stream));
    MGPU_RET(launch_matvec(pb, pb.d_map[0], pb.d_tod, false));
    MGPU_HIP(hipMemcpyAsync(y, pb.d_tod, sizeof(double) * pb.pt.m, hipMemcpyDeviceToHost, pb.stream));
    MGPU_HIP(hipStreamSynchronize(pb.stream));
    return MGPU_OK;
}

int mappraiser_gpu_trmatvec(mgpu_problem *h, const double *y, double *x) {
    MGPU_ARG(h && x && y, "NULL argument");
    Problem &pb = *PB(h);
    MGPU_HIP(hipSetDevice(pb.device));
    MGPU_RET(ensure_work(pb));
    const int n = 3 * pb.pt.npix;
    MGPU_HIP(hipMemcpyAsync(pb.d_tod, y, sizeof(double) * pb.pt.m, hipMemcpyHostToDevice, pb.stream));
    MGPU_RET(launch_trmatvec(pb, pb.d_tod, pb.d_map[0], nullptr, nullptr, nullptr));
    MGPU_HIP(hipMemcpyAsync(x, pb.d_map[0], sizeof(double) * n, hipMemcpyDeviceToHost, pb.stream));
    MGPU_HIP(hipStreamSynchronize(pb.stream));
    return MGPU_OK;
}

int mappraiser_gpu_apply_weights(mgpu_problem *h, double *tod) {
    MGPU_ARG(h && tod, "NULL argument");
    Problem &pb = *PB(h);
    MGPU_HIP(hipSetDevice(pb.device));
    MGPU_RET(ensure_work(pb));
    MGPU_HIP(hipMemcpyAsync(pb.d_tod, tod, sizeof(double) * pb.pt.m, hipMemcpyHostToDevice, pb.stream));
    MGPU_RET(toeplitz_apply(pb, pb.d_tod, pb.d_tod2, true));
    MGPU_HIP(hipMemcpyAsync(tod, pb.d_tod2, sizeof(double) * pb.pt.m, hipMemcpyDeviceToHost, pb.stream));
    MGPU_HIP(hipStreamSynchronize(pb.stream));
    return MGPU_OK;
}

int mappraiser_gpu_apply_weights_noalloc(mgpu_problem *h, const double *d_in, double *d_out, void *stream) {
    MGPU_ARG(h && d_in && d_out, "NULL argument");
    Problem &pb = *PB(h);
    MGPU_HIP(hipSetDevice(pb.device));
    return toeplitz_apply_noalloc(pb, d_in, d_out, static_cast<hipStream_t>(stream));
}

int mappraiser_gpu_precond_build(mgpu_problem *h, double *signal, double *cond, int *lhits, int *n_map) {
    MGPU_ARG(h, "NULL handle");
    Problem &pb = *PB(h);
    MGPU_HIP(hipSetDevice(pb.device));
    double *d_sig = nullptr;
    if (signal) {
        if (!pb.d_sig) MGPU_RET(pb.arena.alloc(&pb.d_sig, (size_t) pb.pt.m));
        d_sig = pb.d_sig;
        MGPU_HIP(hipMemcpyAsync(d_sig, signal, sizeof(double) * pb.pt.m, hipMemcpyHostToDevice, pb.stream));
    }
    MGPU_RET(precond_build_timed(pb, d_sig));
    const int np = pb.pt.npix;
    if (signal)
        MGPU_HIP(hipMemcpyAsync(signal, d_sig, sizeof(double) * pb.pt.m, hipMemcpyDeviceToHost, pb.stream));
    if (cond) MGPU_HIP(hipMemcpyAsync(cond, pb.pc.d_cond, sizeof(double) * np, hipMemcpyDeviceToHost, pb.stream));
    if (lhits) MGPU_HIP(hipMemcpyAsync(lhits, pb.pc.d_hits, sizeof(int) * np, hipMemcpyDeviceToHost, pb.stream));
    MGPU_HIP(hipStreamSynchronize(pb.stream));
    if (n_map) *n_map = 3 * np;
    return MGPU_OK;
}

int mappraiser_gpu_precond_apply(mgpu_problem *h, const double *g, double *Cg) {
    MGPU_ARG(h && g && Cg, "NULL argument");
    Problem &pb = *PB(h);
    MGPU_HIP(hipSetDevice(pb.device));
    MGPU_ARG(pb.pc.built, "precond_apply before precond_build");
    MGPU_RET(ensure_work(pb));
    const int n = 3 * pb.pt.npix;
    MGPU_HIP(hipMemcpyAsync(pb.d_map[0], g, sizeof(double) * n, hipMemcpyHostToDevice, pb.stream));
    MGPU_RET(launch_precond_apply(pb, pb.d_map[0], pb.d_map[1]));
    MGPU_HIP(hipMemcpyAsync(Cg, pb.d_map[1], sizeof(double) * n, hipMemcpyDeviceToHost, pb.stream));
    MGPU_HIP(hipStreamSynchronize(pb.stream));
    return MGPU_OK;
}

int mappraiser_gpu_pcg(mgpu_problem *h, const double *signal, const double *noise, double tol, int maxiter,
                       int stop_after, double *x, double *res_hist, mgpu_report *rep) {
    MGPU_ARG(h && signal && noise, "NULL argument");
    Problem &pb = *PB(h);
    MGPU_HIP(hipSetDevice(pb.device));
    MGPU_ARG(pb.pc.built, "pcg before precond_build");
    if (!pb.d_sig) MGPU_RET(pb.arena.alloc(&pb.d_sig, (size_t) pb.pt.m));
    if (!pb.d_noise) MGPU_RET(pb.arena.alloc(&pb.d_noise, (size_t) pb.pt.m));
    MGPU_HIP(hipMemcpyAsync(pb.d_sig, signal, sizeof(double) * pb.pt.m, hipMemcpyHostToDevice, pb.stream));
    MGPU_HIP(hipMemcpyAsync(pb.d_noise, noise, sizeof(double) * pb.pt.m, hipMemcpyHostToDevice, pb.stream));
    mgpu_report r{};
    MGPU_RET(pcg_run(pb, pb.d_sig, pb.d_noise, tol, maxiter, stop_after, x, res_hist, &r));
    if (rep) *rep = r;
    return MGPU_OK;
}

int mappraiser_gpu_solve_device(mgpu_problem *h, double *d_signal, const double *d_noise, double tol, int maxiter,
                                double *x, double *res_hist, double *cond, int *lhits, mgpu_report *rep) {
    MGPU_ARG(h && d_signal && d_noise, "NULL argument");
    Problem &pb = *PB(h);
    MGPU_HIP(hipSetDevice(pb.device));
    MGPU_RET(ensure_work(pb));
    MGPU_RET(precond_build_timed(pb, d_signal));
    mgpu_report r{};
    MGPU_RET(pcg_run(pb, d_signal, d_noise, tol, maxiter, 0, x, res_hist, &r));
    const int np = pb.pt.npix;
    if (cond) MGPU_HIP(hipMemcpyAsync(cond, pb.pc.d_cond, sizeof(double) * np, hipMemcpyDeviceToHost, pb.stream));
    if (lhits) MGPU_HIP(hipMemcpyAsync(lhits, pb.pc.d_hits, sizeof(int) * np, hipMemcpyDeviceToHost, pb.stream));
    MGPU_HIP(hipStreamSynchronize(pb.stream));
    if (rep) *rep = r;
    return MGPU_OK;
}

int mappraiser_gpu_kernel_times(mgpu_problem *h, double *ms, long long *counts, int reset) {
    if (!h) {  // totals of the drop-in solves run with opt.profile
        std::lock_guard<std::mutex> g(g_kt_mu);
        for (int i = 0; i < MGPU_NKT; i++) {
            if (ms) ms[i] = g_kt_ms[i];
            if (counts) counts[i] = g_kt_cnt[i];
            if (reset) {
                g_kt_ms[i] = 0.0;
                g_kt_cnt[i] = 0;
            }
        }
        return MGPU_OK;
    }
    Problem &pb = *PB(h);
    MGPU_HIP(hipSetDevice(pb.device));
    MGPU_HIP(hipStreamSynchronize(pb.stream));
    pb.prof.collect();
    for (int i = 0; i < MGPU_NKT; i++) {
        if (ms) ms[i] = pb.prof.ms[i];
        if (counts) counts[i] = pb.prof.cnt[i];
        if (reset) {
            pb.prof.ms[i] = 0.0;
            pb.prof.cnt[i] = 0;
        }
    }
    return MGPU_OK;
}

int mappraiser_gpu_noise_invtt(const double *noise, int nb, const int *local_blocks_sizes, int lambda, double fsamp,
                               int apod_window, int on_device, double *invtt, double *fit_params) {
    g_err[0] = 0;
    MGPU_ARG(noise && local_blocks_sizes && invtt && nb > 0, "NULL argument");
    std::unique_ptr<Problem> pb;
    MGPU_RET(new_problem(pb, nullptr));
    int64_t m = 0;
    for (int b = 0; b < nb; b++) m += local_blocks_sizes[b];
    const double *d_noise = noise;
    double *d_inv = invtt;
    if (!on_device) {
        double *dn, *di;
        MGPU_RET(upload(*pb, &dn, noise, (size_t) m));
        MGPU_RET(pb->arena.alloc(&di, (size_t) nb * lambda));
        d_noise = dn;
        d_inv = di;
    }
    MGPU_RET(noise_invtt(*pb, d_noise, nb, local_blocks_sizes, lambda, fsamp, apod_window, d_inv, fit_params));
    if (!on_device)
        MGPU_HIP(hipMemcpyAsync(invtt, d_inv, sizeof(double) * (size_t) nb * lambda, hipMemcpyDeviceToHost,
                                pb->stream));
    MGPU_HIP(hipStreamSynchronize(pb->stream));
    return MGPU_OK;
}

int mappraiser_gpu_tile_variants(int precision, int tile_log2, int variant, int *count, char *desc, int desc_len) {
    MGPU_ARG(count, "NULL count");
    MGPU_ARG(precision == MGPU_FP64 || precision == MGPU_MIXED, "unknown precision %d", precision);
    const bool fp32 = precision == MGPU_MIXED;
    *count = tile_variant_count(fp32, tile_log2);
    if (desc && desc_len > 0) {
        MGPU_ARG(variant >= 1 && variant <= *count, "variant %d out of range 1..%d", variant, *count);
        std::snprintf(desc, (size_t) desc_len, "%s", tile_variant_desc(fp32, tile_log2, variant));
    }
    return MGPU_OK;
}

int mappraiser_gpu_pcg_gls(const char *outpath, const char *ref, int m, int nnz, int *pix, double *pixweights,
                           double *signal, const double *noise, int nb, const int *local_blocks_sizes, int lambda,
                           const double *invtt, double tol, int maxiter, double *x, double *cond, int *lhits,
                           int *lindices, const mgpu_options *opt, mgpu_report *rep) {
    return mgpu::pcg_gls_impl(outpath, ref, m, nnz, pix, pixweights, signal, noise, nb, local_blocks_sizes, lambda,
                              invtt, tol, maxiter, x, cond, lhits, lindices, opt, rep, nullptr);
}

}  // extern "C"

namespace mgpu {

int pcg_gls_impl(const char *outpath, const char *ref, int m, int nnz, int *pix, double *pixweights, double *signal,
                 const double *noise, int nb, const int *local_blocks_sizes, int lambda, const double *invtt,
                 double tol, int maxiter, double *x, double *cond, int *lhits, int *lindices,
                 const mgpu_options *opt, mgpu_report *rep, const mgpu_templates *tm) {
    g_err[0] = 0;
    MGPU_ARG(pix && pixweights && signal && noise && x, "NULL argument");
    MGPU_RET(check_common(m, nnz, nb, local_blocks_sizes, lambda));
    std::unique_ptr<Problem> guard;
    MGPU_RET(new_problem(guard, opt));
    Problem &pb = *guard;
    hipStream_t s = pb.stream;
    const int dev = pb.device;
    const size_t mm = (size_t) m;
    int *d_pix = nullptr, *d_p = nullptr;
    double *d_w = nullptr;
    double2 *d_w12 = nullptr;
    MGPU_RET(pb.arena.alloc(&pb.d_sig, mm));
    MGPU_RET(pb.arena.alloc(&pb.d_noise, mm));
    PackStats ps;
    bool packed = false;
    {
        TimerRange tr("Initializing pointing matrix");  // mappraiser.c:92-103 (with the Toeplitz set-up)
        // packed when the pointing is encoded 3p+k and every I weight is 1 (the
        // configurations of MLmap); otherwise the arrays go up as they are
        MGPU_RET(pb.arena.alloc(&d_p, mm));
        MGPU_RET(pb.arena.alloc(&d_w12, mm));
        MGPU_RET(upload_packed(dev, pix, pixweights, mm, d_p, d_w12, &ps));
        packed = !ps.bad && !ps.w0;
        if (!packed) {
            pb.arena.release(d_p);
            pb.arena.release(d_w12);
            d_p = nullptr;
            d_w12 = nullptr;
            MGPU_RET(pb.arena.alloc(&d_pix, mm * nnz));
            MGPU_RET(pb.arena.alloc(&d_w, mm * nnz));
            MGPU_RET(h2d_staged(dev, d_pix, pix, sizeof(int) * mm * nnz, s));
            MGPU_RET(h2d_staged(dev, d_w, pixweights, sizeof(double) * mm * nnz, s));
        }
        // signal and noise go up on a stream of their own (another host thread)
        // while the device builds the pointing tables and the Toeplitz plan; the
        // preconditioner build cuts the device signal, which the solve then uses
        hipStream_t s_up = nullptr;
        MGPU_HIP(hipStreamCreateWithFlags(&s_up, hipStreamNonBlocking));
        struct UpGuard {
            hipStream_t &st;
            ~UpGuard() {
                if (st) (void) hipStreamDestroy(st);
            }
        } ug{s_up};
        int rc_up = MGPU_OK;
        std::string err_up;
        std::thread up([&] {
            if (hipSetDevice(dev) != hipSuccess) {
                rc_up = MGPU_EHIP;
                err_up = "hipSetDevice failed in the upload thread";
                return;
            }
            rc_up = h2d_staged(dev, pb.d_sig, signal, sizeof(double) * mm, s_up);
            if (rc_up == MGPU_OK) rc_up = h2d_staged(dev, pb.d_noise, noise, sizeof(double) * mm, s_up);
            if (rc_up != MGPU_OK) err_up = mappraiser_gpu_last_error();
        });
        const int rc = packed ? problem_setup_packed(pb, d_p, d_w12, m, ps.pmin, ps.pmax, ps.flagged,
                                                     local_blocks_sizes, nb, lambda, invtt)
                              : problem_setup(pb, d_pix, d_w, m, nnz, local_blocks_sizes, nb, lambda, nullptr, invtt);
        up.join();
        MGPU_RET(rc);
        if (rc_up != MGPU_OK) {
            set_error("%s", err_up.c_str());
            return rc_up;
        }
        MGPU_HIP(hipStreamSynchronize(s_up));  // signal and noise are on the device
    }
    // the packed Q/U weights now belong to the pointing tables
    for (void *p : {(void *) d_pix, (void *) d_w, (void *) d_p})
        if (p) pb.arena.release(p);
    TimerRange total("Total PCG time");  // mappraiser.c:208-231
    if (tm) MGPU_RET(templates_build(pb, tm));
    MGPU_RET(precond_build_timed(pb, pb.d_sig));
    const int np = pb.pt.npix;
    const int trash = pb.pt.trash;
    const int lcount = 3 * np + 3 * trash;
    if (cond) MGPU_HIP(hipMemcpyAsync(cond, pb.pc.d_cond, sizeof(double) * np, hipMemcpyDeviceToHost, s));
    if (lhits) MGPU_HIP(hipMemcpyAsync(lhits, pb.pc.d_hits, sizeof(int) * np, hipMemcpyDeviceToHost, s));
    if (lindices) {
        int *d_lind;
        MGPU_RET(pb.arena.alloc(&d_lind, (size_t) lcount));
        MGPU_RET(pointing_lindices(pb, d_lind));
        MGPU_HIP(hipMemcpyAsync(lindices, d_lind, sizeof(int) * lcount, hipMemcpyDeviceToHost, s));
        MGPU_HIP(hipStreamSynchronize(s));
        pb.arena.release(d_lind);
    }
    // Side effects of the reference on the caller's arrays, delivered by another
    // host thread while the device solves: local indices of the rebuilt matrix
    // into pix; weights and signal of cut samples (valid before, trash rows
    // now) zeroed (precond.c:601-616).
    int *d_loc;
    MGPU_RET(pb.arena.alloc(&d_loc, mm * 3));
    MGPU_RET(pointing_local_indices(pb, d_loc));
    hipEvent_t e_loc = nullptr;
    MGPU_HIP(hipEventCreateWithFlags(&e_loc, hipEventDisableTiming));
    struct EvGuard {
        hipEvent_t e;
        ~EvGuard() { (void) hipEventDestroy(e); }
    } eg{e_loc};
    MGPU_HIP(hipEventRecord(e_loc, s));
    const StageConsumer apply_side_effects = [&](size_t off, const void *chunk, size_t bytes) {
        const int *loc = static_cast<const int *>(chunk);
        const int64_t t0 = (int64_t) (off / (3 * sizeof(int))), nt = (int64_t) (bytes / (3 * sizeof(int)));
        for (int64_t i = 0; i < nt; i++) {
            const int64_t t = t0 + i;
            if (trash && loc[3 * i] == 0 && pix[3 * t] >= 0) {
                for (int c = 0; c < 3; c++) pixweights[3 * t + c] = 0.0;
                signal[t] = 0.0;
            }
        }
        std::memcpy(pix + 3 * t0, loc, bytes);
    };
    int rc_dn = MGPU_OK;
    std::string err_dn;
    std::thread down([&] {
        if (hipSetDevice(dev) != hipSuccess) {
            rc_dn = MGPU_EHIP;
            err_dn = "hipSetDevice failed in the side-effect thread";
            return;
        }
        rc_dn = d2h_staged(dev, nullptr, d_loc, sizeof(int) * mm * 3, e_loc, &apply_side_effects);
        if (rc_dn != MGPU_OK) err_dn = mappraiser_gpu_last_error();
    });
    std::vector<double> hist(maxiter > 0 ? maxiter : 1);
    mgpu_report r{};
    const int rc = pcg_run(pb, pb.d_sig, pb.d_noise, tol, maxiter, 0, x, hist.data(), &r);
    down.join();
    MGPU_RET(rc);
    if (rc_dn != MGPU_OK) {
        set_error("%s", err_dn.c_str());
        return rc_dn;
    }
    r.lcount = lcount;
    r.trash_pix = trash;
    MGPU_RET(kernel_times_accumulate(pb));
    if (outpath && ref) {
        char fn[4096];
        std::snprintf(fn, sizeof(fn), "%s/pcg_residuals_%s.dat", outpath, ref);
        FILE *fp = std::fopen(fn, "wb");
        if (fp) {
            const int nh = std::min(maxiter, r.iterations + 1);
            std::fwrite(hist.data(), sizeof(double), (size_t) nh, fp);
            std::fclose(fp);
        }
    }
    if (rep) *rep = r;
    return MGPU_OK;
}

}  // namespace mgpu
