// mgpu_internal.h -- shared declarations of libmappraiser_gpu.so (gfx950).
//
// Device-resident state of one rank's GLS problem.  Layout in HBM (m samples,
// npix valid pixels, nh = hits on valid pixels, nb Toeplitz blocks):
//   pointing, sample-major  q[m] int32 (valid pixel number, -1 = trash row),
//                           w1[m], w2[m] (and w0[m] when the I weight is not 1)
//   pointing, pixel-major   off[npix+1] int32, ct[nh] int32 (sample index, t
//                           increasing within a pixel), cw1[nh], cw2[nh] (cw0)
//   map vectors             3*npix doubles, x[3q+k] (local map order of the
//                           reference, mapmat.c:496-521)
//   TOD vectors             m elements (double, or float in MGPU_MIXED)
//   Toeplitz                per-block start offsets, tile descriptors, spectral
//                           coefficients per distinct filter, twiddle table
#pragma once

#include <algorithm>
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <memory>
#include <vector>

#include "../../include/mappraiser_gpu.h"

#include <rocprofiler-sdk-roctx/roctx.h>

namespace mgpu {

// roctx range over a scope, named after the reference's timer of the same span
// (MPI_Wtime pairs in mappraiser.c / pcg_true.c), so that a rocprofv3
// --marker-trace shows the reference's phases on the device timeline
struct TimerRange {
    explicit TimerRange(const char *name) { roctxRangePush(name); }
    ~TimerRange() { roctxRangePop(); }
    TimerRange(const TimerRange &) = delete;
    TimerRange &operator=(const TimerRange &) = delete;
};

void set_error(const char *fmt, ...);

#define MGPU_HIP(expr)                                                                         \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            ::mgpu::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr,               \
                              hipGetErrorString(e_));                                          \
            return MGPU_EHIP;                                                                  \
        }                                                                                      \
    } while (0)

#define MGPU_ARG(cond, ...)                                                                    \
    do {                                                                                       \
        if (!(cond)) {                                                                         \
            ::mgpu::set_error(__VA_ARGS__);                                                    \
            return MGPU_EINVAL;                                                                \
        }                                                                                      \
    } while (0)

#define MGPU_RET(expr)                                                                         \
    do {                                                                                       \
        int r_ = (expr);                                                                       \
        if (r_) return r_;                                                                     \
    } while (0)

// Launch check: kernel launches report configuration errors through
// hipGetLastError.
// kernel launches queued by this thread (MGPU_LAUNCHED counts them): lets the
// profiler share one event between two ranges with no launch between them
inline unsigned long long &launch_count() {
    static thread_local unsigned long long n = 0;
    return n;
}
#define MGPU_LAUNCHED()        \
    do {                       \
        ++launch_count();      \
        MGPU_HIP(hipGetLastError()); \
    } while (0)

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
// Process-wide cache of device blocks: hipFree synchronises the device and
// hipMalloc of large blocks costs milliseconds, so blocks released by a
// problem are kept per device (size classes of 2 MiB above 1 MiB) and handed
// to later allocations of at most 1/8 fewer bytes (*cls returns the block's
// class).  On an allocation failure the cache of the device is emptied and
// the allocation retried.
// mappraiser_gpu_finalize() empties every cache.
int dev_alloc(void **p, size_t bytes, int device, size_t *cls);
void dev_free(void *p, size_t cls, int device);
void dev_cache_trim(int device);  // device < 0: every device
// A reserved device region (one hipMalloc) that serves the allocations of the
// threads bound to it, first fit in address order, ranges coalesced on free:
// the request loop reserves its HBM budget once, so no allocation or free
// reaches the HIP runtime while solves run.  dev_free finds a pool's blocks
// by address, from any thread.
struct DevPool;
int dev_pool_create(DevPool **out, size_t bytes, int device);
void dev_pool_destroy(DevPool *pl);  // after every block is freed
void dev_pool_bind(DevPool *pl);     // this thread's dev_alloc from pl (nullptr: the cache)
// Solves run by the request loop's workers (set per thread): their segmented
// PCG keeps the full work lists and refreshes every running segment, the
// behaviour the request loop was verified with under concurrent solves
void serving_bind(bool on);
bool serving_thread();

// Copies between the caller's pageable host arrays and the device through a
// process-wide pool of pinned staging buffers (hostio.hip).  h2d_staged: when
// it returns, the host array may be reused, and work queued on s afterwards
// sees the data (copies of 48 MiB and more are complete on return; smaller
// ones are queued on s).  d2h_staged: the copies start once the event `after` has
// completed; when it returns the data has been delivered, either copied to h or, with
// a consumer, handed to consume(offset, chunk, bytes) chunk by chunk (called
// from several threads at once on disjoint chunks; offsets are multiples of
// 12 MiB).
using StageConsumer = std::function<void(size_t off, const void *chunk, size_t bytes)>;
// h2d_staged_fill: `bytes` of device data produced on the host chunk by chunk:
// fill(off, chunk, len) writes bytes [off, off + len) of the data into the
// pinned chunk (called from several threads at once on disjoint chunks;
// offsets are multiples of 12 MiB).  Complete when the call returns.
using StageProducer = std::function<void(size_t off, void *chunk, size_t bytes)>;
int h2d_staged_fill(int device, void *d, size_t bytes, const StageProducer &fill);
int h2d_staged(int device, void *d, const void *h, size_t bytes, hipStream_t s);
int d2h_staged(int device, void *h, const void *d, size_t bytes, hipEvent_t after,
               const StageConsumer *consume = nullptr);
void staging_free();

// buffers owned by a problem; returned to the cache in its destructor (after
// the problem's stream has drained) or by release(), which first waits for
// the stream so that no pending kernel of this problem still uses the block
// when another problem receives it
struct Arena {
    struct Blk {
        void *p;
        size_t cls;
    };
    std::vector<Blk> blks;
    int device = 0;
    hipStream_t stream = nullptr;
    size_t live = 0, peak = 0;  // bytes held (whole blocks) now and at the high-water mark
    template <typename T>
    int alloc(T **p, size_t n) {
        *p = nullptr;
        if (n == 0) n = 1;
        void *q = nullptr;
        size_t cls = 0;
        const int rc = dev_alloc(&q, n * sizeof(T), device, &cls);
        if (rc != MGPU_OK) return rc;
        blks.push_back({q, cls});
        live += cls;
        peak = std::max(peak, live);
        *p = (T *) q;
        return MGPU_OK;
    }
    void release(void *p) {
        for (auto &b : blks)
            if (b.p == p) {
                if (stream) (void) hipStreamSynchronize(stream);
                dev_free(b.p, b.cls, device);
                live -= b.cls;
                b.p = nullptr;
            }
    }
    ~Arena() {
        for (auto &b : blks)
            if (b.p) dev_free(b.p, b.cls, device);
    }
};

// Scratch that grows on demand (hipcub temporaries), held in the problem's arena.
struct Scratch {
    Arena *arena = nullptr;
    unsigned char *p = nullptr;
    size_t n = 0;
    int ensure(size_t bytes) {
        if (bytes <= n) return MGPU_OK;
        if (p) arena->release(p);
        p = nullptr;
        n = 0;
        MGPU_RET(arena->alloc(&p, bytes));
        n = bytes;
        return MGPU_OK;
    }
};

// ---------------------------------------------------------------------------
// Toeplitz plan
// ---------------------------------------------------------------------------
struct TileDesc {
    int64_t blk_lo;   // first sample of the block
    int64_t blk_hi;   // one past the last sample of the block
    int64_t out_lo;   // first output sample of this tile (blk_lo + k*L)
    int32_t block;    // block index (mean slot)
    int32_t filter;   // spectral coefficient set
};

constexpr int kChunk = 1024;  // samples per partial-sum chunk (P epilogue / partials kernel)

// One class of FFT tiles: the blocks of one band lambda that use one tile
// length (a solo problem has one class; a batch of problems with different
// bands or block-length profiles has one per (lambda, tile length)).
struct TileClass {
    int lambda = 0;
    int log2n = 0;                   // real tile length N = 2^log2n
    int L = 0;                       // useful outputs per tile, N - 2(lambda-1)
    int ntiles = 0;
    TileDesc *d_tiles = nullptr;     // TileDesc.filter indexes this class's coefficient sets
    int nfilters = 0;
    double *d_coef64 = nullptr;      // nfilters * NC * 2  (a_k, b_k) interleaved, includes 1/NC
    float *d_coef32 = nullptr;
    double *d_tw64 = nullptr;        // NC complex: exp(-2 pi i m / NC)
    float *d_tw32 = nullptr;
    int variant[2] = {0, 0};         // tile kernel variant used for FP64 [0] and FP32 [1] (0 = not chosen)
    // N = 32768 only: the coefficients in the lane order of the wave-local tile kernel
    double *d_coefW64 = nullptr;
    float *d_coefW32 = nullptr;
    // FP32 products (MGPU_MIXED) with a longer tile than the FP64 one, when that
    // pays off (less overlap per useful sample); log2n32 == 0: FP32 uses the FP64 tiles
    int log2n32 = 0, L32 = 0, ntiles32 = 0;
    bool fp32_long = false;          // FP32 uses the long tiles (only when the autotuner picks them)
    TileDesc *d_tiles32 = nullptr;
    float *d_coef32L = nullptr;      // nfilters * (N32/2) (a_k, b_k)
    float *d_tw32L = nullptr;
    // host copies of both tile lists in sample order, and the tiles of the
    // segments still running during a segmented PCG (-1: every tile)
    std::vector<TileDesc> h_tiles, h_tiles32;
    TileDesc *d_tiles_run = nullptr, *d_tiles32_run = nullptr;
    int ntiles_run = -1, ntiles32_run = -1;
};

struct ToeplitzPlan {
    int lambda = 0;                  // the band of every block (the largest one when bands differ)
    bool mixed_lambda = false;       // blocks carry different bands (batched problems)
    int nb = 0;
    int64_t m = 0;
    int64_t covered = 0;             // sum of block sizes (samples past it are left unchanged)
    std::vector<int64_t> h_start;    // nb+1 block starts
    int64_t *d_start = nullptr;      // nb+1
    double *d_T0 = nullptr;          // nb, T_b[0]
    int *d_block_lambda = nullptr;   // nb, band of each block
    // FFT tiles, one set per class
    std::vector<TileClass> cls;
    int nfilters = 0;                // distinct filters over all classes
    std::vector<int> h_block_filter; // nb: global filter of each block
    // blocks shorter than lambda-1 are only mean-subtracted (stmm_main:912-913)
    std::vector<int> h_short_blocks;
    int *d_short_blocks = nullptr;
    // bands too wide for an LDS tile: direct product over these blocks
    std::vector<int> h_direct_blocks;
    int *d_direct_blocks = nullptr;
    double *d_direct_invtt = nullptr; // rows of every block, row b at d_inv_off[b]
    int64_t *d_inv_off = nullptr;     // nb: offset of block b's band in d_direct_invtt
    // partial sums for the per-block mean (toeplitz_wizard.c:86-95)
    int64_t nchunks = 0;
    double *d_partials = nullptr;    // nchunks + nb
    double *d_mean = nullptr;        // nb: input mean per block
    // MGPU_OFFSET_SYMMETRIC: mean of T (v - mean) per block, from the block-edge terms
    int offset_mode = MGPU_OFFSET_REFERENCE;
    int *d_chunk_block = nullptr;    // nchunks: first block overlapping chunk c (partial-sum epilogues)
    double *d_edge_part = nullptr;   // nb * 8 slice sums of the symmetric mode's output mean
    double *d_tail = nullptr;        // tails of every filter: tail_f[j] = sum_{k>j} T_f[k], j < lambda_f
    int64_t *d_tail_off = nullptr;   // nfilters: offset of filter f's tail in d_tail
    int *d_block_filter = nullptr;   // nb
    int *d_block_kind = nullptr;     // nb: 0 Toeplitz product, 1 short block (identity)
    // tile variant of the products, for reports: class 0's (0 = none chosen)
    int variant(int slot) const { return cls.empty() ? 0 : cls[0].variant[slot]; }
};

// ---------------------------------------------------------------------------
// Pointing
// ---------------------------------------------------------------------------
struct Pointing {
    int64_t m = 0;
    int nnz = 3;
    int npix = 0;          // valid pixels (current)
    int trash = 0;         // Mat.trash_pix (current)
    int npix0 = 0;         // valid pixels of the first MatInit (before the cut)
    int trash0 = 0;
    bool unit_w0 = true;   // I weight identically 1
    int64_t nh = 0;        // sliced-ELL length (hits + padding)
    int *d_q = nullptr;    // [m]
    double *d_w0 = nullptr;   // [m] I weight (only when not identically 1)
    double2 *d_w12 = nullptr; // [m] (Q, U) weights interleaved: one 16-byte load per sample
    float *d_w0f = nullptr;   // FP32 copies (mixed precision)
    float2 *d_w12f = nullptr;
    int64_t nvalid = 0;    // samples on valid pixels (CSR length)
    int *d_off = nullptr;  // [npix+1] CSR offsets (hits of pixel q: off[q]..off[q+1])
    int *d_cst = nullptr;  // [nvalid] CSR sample indices, increasing within a pixel
    int64_t *d_ell_base = nullptr; // [nslice+1] sliced-ELL slice offsets (elements)
    int *d_ct = nullptr;   // [nh] sliced-ELL sample index (-1 = padding)
    double *d_cw0 = nullptr;    // [nh] I weight in ELL order (only when not identically 1)
    double2 *d_cw12 = nullptr;  // [nh] (Q, U) weights in ELL order
    float *d_cw0f = nullptr;
    float2 *d_cw12f = nullptr;
    int *d_pixg = nullptr; // [npix0] global pixel number p (valid pixels, increasing)
    int *d_chunk_perm = nullptr; // [ceil(m/kChunk)] chunk processed by P's workgroup b
    std::vector<int> h_chunk_order; // the chunks in P's order before they are dealt to the XCDs
    int *d_perm_run = nullptr;   // chunks of the segments still running (segmented PCG)
    int nch_run = -1;            // their count (-1: every chunk, d_chunk_perm)
};

// ---------------------------------------------------------------------------
// Preconditioner
// ---------------------------------------------------------------------------
struct Precond {
    bool built = false;
    int n_cut = 0;
    double *d_inv = nullptr;   // [npix * 9] row-major inverse blocks
    double *d_cond = nullptr;  // [npix0]
    int *d_hits = nullptr;     // [npix0]
};

// ---------------------------------------------------------------------------
// Kernel timing with HIP events on the problem's stream (opt.profile)
// ---------------------------------------------------------------------------
struct Profiler {
    bool on = false;
    struct Rec {
        int cat;
        int tag;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> free_ev;
    double ms[MGPU_NKT] = {0};
    long long cnt[MGPU_NKT] = {0};
    int open_cat = -1;
    int tag = 0;  // PCG iteration the next launches belong to (0 = outside the loop)
    hipEvent_t open_a = nullptr;
    // end event of the last range: the next range starts from it when nothing
    // was launched in between on the same stream (half the events in the PCG
    // loop, whose ranges follow each other)
    hipEvent_t last_end = nullptr;
    hipStream_t last_stream = nullptr;
    unsigned long long last_launch = 0;

    hipEvent_t get() {
        if (!free_ev.empty()) {
            hipEvent_t e = free_ev.back();
            free_ev.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void) hipEventCreate(&e);
        return e;
    }
    void begin(int cat, hipStream_t s) {
        if (!on) return;
        open_cat = cat;
        if (last_end && last_stream == s && last_launch == launch_count()) {
            open_a = last_end;
        } else {
            open_a = get();
            (void) hipEventRecord(open_a, s);
        }
    }
    void end(hipStream_t s) {
        if (!on || open_cat < 0) return;
        hipEvent_t b = get();
        (void) hipEventRecord(b, s);
        pending.push_back(Rec{open_cat, tag, open_a, b});
        open_cat = -1;
        last_end = b;
        last_stream = s;
        last_launch = launch_count();
    }
    // call after the stream has been synchronised; launches of iterations past
    // `last_iter` were queued after convergence and returned at once: not counted
    void collect(int last_iter = 0x7fffffff) {
        for (size_t i = 0; i < pending.size(); i++) {
            const Rec &r = pending[i];
            float t = 0.f;
            if (r.tag <= last_iter && hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
                ms[r.cat] += t;
                cnt[r.cat] += 1;
            }
            if (i == 0 || pending[i - 1].b != r.a) free_ev.push_back(r.a);  // a shared start is the previous end
            free_ev.push_back(r.b);
        }
        pending.clear();
        last_end = nullptr;  // returned to the free list
    }
    ~Profiler() {
        for (size_t i = 0; i < pending.size(); i++) {
            if (i == 0 || pending[i - 1].b != pending[i].a) (void) hipEventDestroy(pending[i].a);
            (void) hipEventDestroy(pending[i].b);
        }
        for (auto e : free_ev) (void) hipEventDestroy(e);
    }
};

// ---------------------------------------------------------------------------
// Problem
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Map segments: independent problems solved together (batched realizations).
// Segment r owns the local pixels [h_q[r], h_q[r+1]).  Every map-domain
// workgroup (kMapWG pixels) lies inside one segment and segment r's workgroups
// are [h_wg0[r], h_wg0[r+1]), so per-workgroup dot partials never mix
// segments and each segment's sums are grouped exactly as in a solo solve.
// A solo problem is one segment.
// ---------------------------------------------------------------------------
constexpr int kMapWG = 256;

struct Segments {
    int nseg = 1;
    std::vector<int64_t> h_pix0;  // nseg+1: segment r owns global pixels [h_pix0[r], h_pix0[r+1]); empty: one segment
    int npix = -1;            // pixel count the tables describe (-1: not built)
    std::vector<int> h_q;     // nseg+1 first local pixel of each segment
    std::vector<int> h_wg0;   // nseg+1 first workgroup of each segment
    int nwg = 0;              // map-domain workgroups
    int *d_q = nullptr;       // nseg+1
    int *d_wg0 = nullptr;     // nseg+1
    int *d_wg_seg = nullptr;  // nwg: segment of each workgroup
    int *d_wg_q0 = nullptr;   // nwg: first local pixel of each workgroup
    int *d_state = nullptr;   // nseg: PCG state (0 running, 1 converged, 2 stopped)
    int *d_rstate = nullptr;  // nseg: nonzero for the segments an FP64 refresh of pending segments skips
    // segment of every P chunk and Toeplitz block (-1: none, e.g. padding), so
    // that P, N^-1 and P^T skip the work of stopped segments
    int *d_chunk_seg = nullptr;
    int *d_block_seg = nullptr;
    std::vector<int> h_chunk_seg, h_block_seg;  // host copies
};

// Work lists of a segmented PCG restricted to its running segments: P's chunks
// (dealt to the XCDs again) and every tile class's tiles.  While the PCG loop
// runs (Problem::live_state set) the P and tile launches use them, so stopped
// segments cost no workgroups; run_lists_clear restores the full lists.
struct Problem;
int run_lists_update(Problem &pb, const std::vector<int> &state);
void run_lists_clear(Problem &pb);
// XCD x runs workgroups b = x, x + 8, ...: position b of the result holds element
// off[b % 8] + b / 8 of `order`, so each XCD gets a contiguous run of it
template <typename T>
std::vector<T> deal_to_xcds(const std::vector<T> &order) {
    constexpr int kX = 8;
    const int n = (int) order.size();
    int off[kX + 1] = {0};
    for (int x = 0; x < kX; x++) off[x + 1] = off[x] + (x < n ? (n - 1 - x) / kX + 1 : 0);
    std::vector<T> o(order.size());
    for (int b = 0; b < n; b++) o[b] = order[off[b % kX] + b / kX];
    return o;
}

// ---------------------------------------------------------------------------
// Two-level preconditioner (build_precond precond = 1, 2): deflation space Z,
// its image A Z and pinv(E), E[i][j] = <Z_j, A Z_i>
// ---------------------------------------------------------------------------
constexpr int kMaxZn = 256;
struct TwoLevel {
    int zn = 0;                 // columns of Z (0: block Jacobi only)
    int ld = 0;                 // column stride (3 * npix of the solve)
    double *d_Z = nullptr;      // [zn * ld], column j at j * ld
    double *d_AZ = nullptr;     // [zn * ld]
    double *d_Em1 = nullptr;    // [zn * zn] row-major
    double *d_qtx = nullptr;    // [zn] Em1 Z^T g of the last apply
    double *d_part = nullptr;   // [zn * nchunk] partials of Z^T g
    int nchunk = 0;
    double build_ms = 0.0;      // device time of the last build
};

// ---------------------------------------------------------------------------
// Template marginalisation (MTmap / PCG_GLS_templates): T, T^T, the per-block
// kernels B_b = w_b T_b^T T_b and their pseudo-inverses.  Blocks are CES-major
// (block c*ndet + j) and span their CES, so every table indexed by a CES
// sample is shared by the CES's detectors.
// ---------------------------------------------------------------------------
struct TItem {           // one workgroup of T^T: a sweep, an azimuth bin or a time bin of a block
    int b, kind, lo, hi; // block; kind 0 sweep, 1 azimuth bin (range in the azimuth order), 2 time bin
    int64_t rel;         // first output amplitude, relative to the block's amplitudes
    int stride, pad;     // distance between the class outputs (sweeps or time bins)
};
struct BTask {           // one kernel entry (and its mirror) summed in sample order
    int kind;            // 0 samples [lo, hi); 1 weighted, 2 unweighted over perm[lo, hi)
    int row, col;        // entry within the block
    int lo, hi;
    int ka, kb;          // weight slots (0..15 Legendre order, 16.. HWP slot)
    int perm;            // kind 1/2: 0 azimuth order, 1 (sweep, az), 2 (time bin, az)
    int sym;             // also write (col, row)
    int pad;
};

struct Templates {
    bool on = false;
    bool live = false;          // applied after N^-1 inside the PCG loop and its right-hand side
    int npoly = 0, ground = 0, nsss = 0, nhwp = 0;
    double fsdt = 0.0;          // sampling_freq * delta_t
    int nces = 0, ndet = 0;
    int64_t ncs = 0;            // entries of a per-CES table (all CES, workgroup-aligned)
    int64_t ntot = 0, nB = 0;   // amplitudes; entries of all kernels
    std::vector<int> h_nt, h_rank;
    std::vector<int64_t> h_off, h_boff;
    int *d_ns = nullptr, *d_nh = nullptr;   // [nces] sweeps, time bins
    int64_t *d_c0 = nullptr;                // [nces+1] first entry of each CES in the per-CES tables
                                            // (starts rounded up to whole 256-sample workgroups)
    int *d_len = nullptr;                   // [nces] CES length
    int *d_ces_of = nullptr;                // [nwg_samples] CES of each sample-kernel workgroup
    int nwg_samples = 0;
    int *d_sweep = nullptr;                 // [ncs] sweep of the sample (-1: none)
    int *d_az = nullptr;                    // [ncs] azimuth bin
    double *d_leg = nullptr;                // [npoly * ncs] Legendre weights, order-major
    double *d_hc = nullptr;                 // [2 nhwp * ncs] cos_1, sin_1, cos_2, ...
    int *d_perm = nullptr;                  // [3 * ncs] sample orders: azimuth, (sweep, az), (time bin, az)
    int64_t *d_off = nullptr;               // [nb+1] first amplitude of each block
    int *d_row_block = nullptr;             // [ntot] block of each amplitude
    int64_t *d_boff = nullptr;              // [nb+1] first kernel entry of each block
    double *d_B = nullptr, *d_Binv = nullptr; // [nB] row-major nt_b x nt_b blocks
    double *d_a1 = nullptr, *d_a2 = nullptr;  // [ntot]
    TItem *d_items = nullptr;
    int nitems = 0;
    double build_ms = 0.0;
};

struct Problem {
    Profiler prof;
    int device = 0;
    hipStream_t stream = nullptr;
    mgpu_options opt{};
    Arena arena;
    Scratch cub_tmp;
    Pointing pt;
    ToeplitzPlan tp;
    Precond pc;
    Segments sg;
    TwoLevel tl;
    Templates tm;
    // PCG state per segment while the iteration loop runs (P, N^-1 and P^T
    // skip stopped segments); nullptr outside it, where every segment is computed
    const int *live_state = nullptr;
    // N^-1 products follow stbmmProd rather than apply_weights: with lambda 1
    // they remove the block means too (the two-level products, precond.c:1017-1022)
    bool stbmm_offsets = false;
    // working TOD / map vectors (allocated lazily)
    double *d_tod = nullptr, *d_tod2 = nullptr;   // [m] FP64
    float *d_todf = nullptr, *d_tod2f = nullptr;  // [m] FP32 (mixed)
    double *d_sig = nullptr;                      // [m] signal (device copy)
    double *d_noise = nullptr;                    // [m]
    double *d_map[8] = {nullptr};                 // map work vectors (3*npix0 each)
    float *d_mapf = nullptr;                      // FP32 copy of a map vector (mixed P)
    double *d_red = nullptr;                      // reduction partials
    double *d_scal = nullptr;                     // device scalars of PCG (per segment)
    int64_t nscal = 0;
    int *d_flag = nullptr;                        // PCG stop flag
    int64_t nred = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // device time of the set-up steps (mgpu_report.t_setup_ms = their sum):
    // pointing build + Toeplitz plan of the constructor, preconditioner build
    double t_create_ms = 0.0, t_precond_ms = 0.0;
    ~Problem();
};

// ---------------------------------------------------------------------------
// entry points between translation units
// ---------------------------------------------------------------------------
int pointing_build(Problem &pb, const int *d_pix, const double *d_w, int64_t m, int nnz);
// the same from a packed upload: d_p = each sample's global pixel (-1 flagged),
// d_w12 = its (Q, U) weights with the I weight 1 (the arena block becomes P's
// w12); pmin/pmax over the valid samples, any_flagged if some sample is flagged
int pointing_build_packed(Problem &pb, const int *d_p, double2 *d_w12, int64_t m, int pmin, int pmax,
                          bool any_flagged);
int pointing_build_from_p(Problem &pb, const int *d_p, int pmin, int pmax, bool any_flagged, const double *d_wts,
                          double2 *d_w12);
int pointing_local_indices(Problem &pb, int *d_out);   // MatLocalShape indices [m*3]
int pointing_lindices(Problem &pb, int *d_out);        // lindices [lcount]
int pointing_make_fp32(Problem &pb);
// remove pixels with kept[q]==0: renumber q, compact CSR, rebuild the ELL
int pointing_compact(Problem &pb, const unsigned *d_kept);

// y = P x (FP64); optional fused per-block partials when tp has blocks
int launch_matvec(Problem &pb, const double *d_x, double *d_y, bool partials);
int launch_matvec_f32(Problem &pb, const float *d_x, float *d_y, bool partials);
// x = P^T y; optional fused partial dots of <x, hvec> into d_dot (one per workgroup)
int launch_trmatvec(Problem &pb, const double *d_y, double *d_x, const double *d_hvec,
                    double *d_dotpart, int *nparts);
int launch_trmatvec_f32(Problem &pb, const float *d_y, double *d_x, const double *d_hvec,
                        double *d_dotpart, int *nparts);
int trmatvec_nparts(const Problem &pb);
// smallest valid local pixel of every P chunk (INT32_MAX: none), on the host
int pointing_chunk_keys(Problem &pb, std::vector<int> &key);

// Block definitions and the Toeplitz set-up (defineBlocks_avg, tpltz_init).
// invtt holds the band of every block, block b's band at the running sum of
// the earlier blocks' bands (b * lambda when every band is lambda long).
// block_lambda (optional): the band of each block (default: lambda for all);
// block_log2n (optional): the FP64 tile length of each block (default: the
// length choose_log2n picks for its band and longest block of that band).
int toeplitz_plan(Problem &pb, const int *block_sizes, int nb, int lambda, const double *d_invtt,
                  const double *h_invtt, const int *block_lambda = nullptr, const int *block_log2n = nullptr);
// N^-1 on d_in (its per-block partials must already be in tp.d_partials unless
// compute_partials) -> d_out (may not alias d_in)
int toeplitz_apply(Problem &pb, const double *d_in, double *d_out, bool compute_partials);
int toeplitz_apply_f32(Problem &pb, const float *d_in, float *d_out, bool compute_partials);
// scratch of one N^-1 product: partial sums [nchunks + nb], means [nb], slice
// sums of the symmetric mode's output mean [nb * 8]
struct ToeplitzScratch {
    double *partials, *mean, *edge_part;
};
size_t toeplitz_scratch_doubles(const ToeplitzPlan &T);
// N^-1 (FP64, partial sums computed) with the scratch laid out in d_scratch
// (toeplitz_scratch_doubles(T) doubles) instead of the plan's buffers
int toeplitz_apply_scratch(Problem &pb, const double *d_in, double *d_out, double *d_scratch);
// N^-1 with the scratch in the library's static device array (no allocation),
// launched on `stream` (nullptr: the handle's stream); noalloc_scratch_init
// looks up the array's address on the current device (once per device)
int noalloc_scratch_init(int device);
int toeplitz_apply_noalloc(Problem &pb, const double *d_in, double *d_out, hipStream_t stream);
int tile_variant_count(bool fp32, int log2n);
// real tile length 2^k the plan of a lambda / longest block would use (0: no tiles)
int toeplitz_tile_log2(int lambda, int64_t nmax, int forced);
const char *tile_variant_desc(bool fp32, int log2n, int variant);

int precond_build(Problem &pb, double *d_signal);
// precond_build timed with the problem's events (t_precond_ms)
int precond_build_timed(Problem &pb, double *d_signal);
// pointing build + Toeplitz plan of a new problem, timed (t_create_ms)
int problem_setup(Problem &pb, const int *d_pix, const double *d_w, int64_t m, int nnz, const int *block_sizes,
                  int nb, int lambda, const double *d_invtt, const double *h_invtt);
// A v = P^T N^-1 P v in FP64 (map -> map), optional fused <v, Av> partials
int apply_A64(Problem &pb, const double *d_v, double *d_out, double *d_dot);
// two-level preconditioner: build Z, A Z, pinv(E) for opt.precond (1 a priori,
// 2 a posteriori from signal + noise); release frees it
int twolvl_build(Problem &pb, const double *d_signal, const double *d_noise);
void twolvl_release(Problem &pb);
// Cg = BJ (g - AZ Qtx) + Z Qtx, Qtx = Em1 Z^T g, with the PCG dot partials:
// init: h = Cg, part = [<Cg,g> | <g,g>]; else part = [<Cg,g> | <Cg,g-gp> | <g,g>]
// (one entry per map workgroup of each dot, the layout of the block-Jacobi kernels)
int twolvl_apply(Problem &pb, const double *d_g, const double *d_gp, double *d_Cg, double *d_h, double *d_part,
                 bool init, const int *d_state, const int *d_flag);
int launch_precond_apply(Problem &pb, const double *d_g, double *d_Cg);
// template marginalisation (MTmap): tables, kernels and pseudo-inverses
int templates_build(Problem &pb, const mgpu_templates *t);
// y (a TOD already weighted by N^-1) -= W T Binv T^T y
int templates_filter(Problem &pb, double *d_y);
// y = W q - W T Binv T^T W q: apply_weights (lambda 1) and the template part in one pass
int templates_noise(Problem &pb, const double *d_q, double *d_y);
// out = T a (TVecProd) or out = T^T y (TrTVecProd) on device buffers
int templates_tvec(Problem &pb, const double *d_a, double *d_out);
int templates_trtvec(Problem &pb, const double *d_y, double *d_out);
// mappraiser_gpu_pcg_gls, with template classes attached when tm is not NULL (MTmap)
int pcg_gls_impl(const char *outpath, const char *ref, int m, int nnz, int *pix, double *pixweights, double *signal,
                 const double *noise, int nb, const int *local_blocks_sizes, int lambda, const double *invtt,
                 double tol, int maxiter, double *x, double *cond, int *lhits, int *lindices,
                 const mgpu_options *opt, mgpu_report *rep, const mgpu_templates *tm);

// per-segment PCG outcome (batched problems)
struct SegReport {
    int iterations = 0, converged = 0;
    double res0 = 0.0, res = 0.0, true_res_rel = -1.0;
    int q0 = 0, q1 = 0;  // local pixel range of the segment
};
// h_hist: maxiter entries per segment (segment-major), may be NULL.  seg_tol /
// seg_maxiter (optional, one entry per segment) replace tol / maxiter per
// segment; maxiter is then their largest value.  seg_done (optional) is called
// on the calling thread once per segment, at the first poll that finds the
// segment stopped (its map d_x[3 q0 .. 3 q1) final, the stream idle), with its
// report (true_res_rel: -1, the FP64 true residual comes after the loop).
using SegDone = std::function<int(int seg, const SegReport &rep, const double *d_x)>;
int pcg_run(Problem &pb, const double *d_signal, const double *d_noise, double tol, int maxiter,
            int stop_after, double *h_x, double *h_hist, mgpu_report *rep,
            std::vector<SegReport> *segs = nullptr, const std::vector<double> *seg_tol = nullptr,
            const std::vector<int> *seg_maxiter = nullptr, const SegDone *seg_done = nullptr);

int ensure_work(Problem &pb);
// batched solve (batch.hip) of items[all[k]], each with its grid size,
// tolerance and iteration cap; done (optional) is called with the item index
// as soon as that item's results are in place (the others may still iterate);
// check_item validates one item
int pcg_batch_impl(mgpu_batch_item *items, const std::vector<int> &all, const std::vector<int> &grid_of,
                   const std::vector<double> &tol_of, const std::vector<int> &maxiter_of, bool on_dev,
                   const mgpu_options *opt, const std::function<void(int)> *done);
int check_item(const mgpu_batch_item &it, int i);
// adds a finished solve's kernel-class times (opt.profile) to the process totals
// that mappraiser_gpu_kernel_times(NULL, ...) reports
int kernel_times_accumulate(Problem &pb);
// invtt rows [nb*lambda] (device) from the noise timestream (device), see mappraiser_gpu_noise_invtt
int noise_invtt(Problem &pb, const double *d_noise, int nb, const int *block_sizes, int lambda, double fsamp,
                int window_kind, double *d_invtt, double *h_params);
// a problem on the current device with its own stream and the given options
int new_problem(std::unique_ptr<Problem> &pb, const mgpu_options *opt);
// common argument checks of the problem constructors
int check_common(int m, int nnz, int nb, const int *bs, int lambda);
// (re)build the segment tables for the current pixel set (no-op when current)
int segments_ensure(Problem &pb);

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
inline unsigned grid_for(int64_t n, int block) { return (unsigned) ((n + block - 1) / block); }

}  // namespace mgpu
