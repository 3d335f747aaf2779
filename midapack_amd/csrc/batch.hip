// batch.hip -- many independent GLS problems solved as one segmented problem.
//
// Every item is placed in one problem (items with diagonal noise, lambda = 1,
// form a problem of their own, solved after the others):
//   * samples: item r occupies [off_r, off_r + m_r) of the TOD, off_r a multiple
//     of kChunk, followed by padding samples up to the next multiple.  Padding
//     samples are flagged (-3,-2,-1), carry weights (1,0,0) and zero signal and
//     noise, and form one extra Toeplitz block of their own, so every TOD value
//     they take is exactly zero.  Chunk alignment makes the fused per-block sums
//     of P group the same samples as in the item's solo solve.
//   * pixels: item r's pixel p becomes pix0_r + p, pix0_r the sum of the earlier
//     items' grid sizes, so items never share a pixel; after the pointing build
//     the local pixels of item r are one contiguous range, a map segment of the
//     PCG (pcg.hip) with its own tolerance and iteration cap.
//   * Toeplitz blocks keep their item's band and the tile length the item would
//     use alone: the plan holds one tile class per (band, tile length), and an
//     N^-1 product launches one tile kernel per class, all in the same launch
//     sequence of each PCG iteration.
// P, N^-1 and P^T are block diagonal over the items, the preconditioner and
// the cut are per pixel and the PCG scalars are per segment, so each item
// follows exactly the arithmetic of its solo solve.
#include "mgpu_internal.h"

#include <algorithm>
#include <map>
#include <string>

namespace mgpu {

// shift item pixels into the item's segment: 3p+k -> 3(p + r G)+k; flag p >= G
__global__ void k_batch_pix(int *__restrict__ pix, int64_t n3, int add3, int g3, int *__restrict__ err) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const int v = pix[i];
    if (v >= 0) {
        if (v >= g3)
            atomicOr(err, 1);
        else
            pix[i] = v + add3;
    }
}

// padding samples [t0, t1): flagged pointing, weights (1,0,0), zero signal and noise
__global__ void k_batch_pad(int *__restrict__ pix, double *__restrict__ w, double *__restrict__ sig,
                            double *__restrict__ noise, int64_t t0, int64_t t1) {
    const int64_t t = t0 + (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= t1) return;
    for (int k = 0; k < 3; k++) {
        pix[3 * t + k] = k - 3;
        w[3 * t + k] = k == 0 ? 1.0 : 0.0;
    }
    sig[t] = 0.0;
    noise[t] = 0.0;
}

namespace {

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void) hipEventDestroy(a);
        if (b) (void) hipEventDestroy(b);
    }
};

}  // namespace

// One batched solve: the items idx[r] of the caller's array, each with its
// grid size, tolerance and iteration cap.
int solve_items(mgpu_batch_item *items, const std::vector<int> &idx, const std::vector<int> &grid,
                const std::vector<double> &tol, const std::vector<int> &maxiter, bool on_dev,
                const mgpu_options *opt, const std::function<void(int)> *done) {
    const int R = (int) idx.size();
    const hipMemcpyKind kind = on_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    mgpu_options gopt;
    mappraiser_gpu_options_default(&gopt);
    if (opt) gopt = *opt;
    // layout
    std::vector<int64_t> off((size_t) R + 1, 0), pix0((size_t) R + 1, 0), boff((size_t) R + 1, 0);
    std::vector<int> l2((size_t) R, 0);
    int nb_tot = 0, lam_max = 1, maxit = 1;
    double tol_max = 0.0;
    int64_t ninv = 0;
    for (int r = 0; r < R; r++) {
        const mgpu_batch_item &it = items[idx[r]];
        const int64_t padded = (it.m + kChunk - 1) / kChunk * kChunk;
        off[r + 1] = off[r] + padded;
        pix0[r + 1] = pix0[r] + grid[r];
        const int nbr = it.nb + (padded > it.m ? 1 : 0);
        boff[r + 1] = boff[r] + nbr;
        nb_tot += nbr;
        ninv += (int64_t) nbr * it.lambda;
        lam_max = std::max(lam_max, it.lambda);
        maxit = std::max(maxit, maxiter[r]);
        tol_max = std::max(tol_max, tol[r]);
        int64_t nmax = 0;
        for (int b = 0; b < it.nb; b++) nmax = std::max<int64_t>(nmax, it.local_blocks_sizes[b]);
        // the tile length the item would use alone (its own block-length profile)
        l2[r] = toeplitz_tile_log2(it.lambda, nmax, gopt.tile_log2);
    }
    const int64_t M = off[R];
    MGPU_ARG(M < (int64_t) INT32_MAX, "batch of %lld padded samples exceeds 2^31-1: split it", (long long) M);
    MGPU_ARG(3 * pix0[R] < (int64_t) INT32_MAX - 3,
             "%d items with %lld grid pixels in all exceed the int32 pixel index range: split the batch", R,
             (long long) pix0[R]);

    std::unique_ptr<Problem> pb;
    MGPU_RET(new_problem(pb, &gopt));
    Problem &P = *pb;
    hipStream_t s = P.stream;
    EventPair ev;
    MGPU_HIP(hipEventCreate(&ev.a));
    MGPU_HIP(hipEventCreate(&ev.b));
    int *d_pix, *d_err;
    double *d_w, *d_inv;
    MGPU_RET(P.arena.alloc(&d_pix, (size_t) (3 * M)));
    MGPU_RET(P.arena.alloc(&d_w, (size_t) (3 * M)));
    MGPU_RET(P.arena.alloc(&P.d_sig, (size_t) M));
    MGPU_RET(P.arena.alloc(&P.d_noise, (size_t) M));
    MGPU_RET(P.arena.alloc(&d_inv, (size_t) ninv));
    MGPU_RET(P.arena.alloc(&d_err, 1));
    MGPU_HIP(hipMemsetAsync(d_err, 0, sizeof(int), s));
    // host arrays go through the pinned staging pool (large ones), device arrays are copied on s
    auto upload = [&](void *dst, const void *src, size_t bytes) -> int {
        if (!on_dev) return h2d_staged(P.device, dst, src, bytes, s);
        MGPU_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
        return MGPU_OK;
    };
    std::vector<int> sizes, blam, bl2;
    sizes.reserve((size_t) nb_tot);
    blam.reserve((size_t) nb_tot);
    bl2.reserve((size_t) nb_tot);
    int64_t irow = 0;  // running offset of the bands in d_inv
    for (int r = 0; r < R; r++) {
        const mgpu_batch_item &it = items[idx[r]];
        const int64_t o = off[r];
        MGPU_RET(upload(d_pix + 3 * o, it.pix, sizeof(int) * 3 * (size_t) it.m));
        MGPU_RET(upload(d_w + 3 * o, it.pixweights, sizeof(double) * 3 * (size_t) it.m));
        MGPU_RET(upload(P.d_sig + o, it.signal, sizeof(double) * (size_t) it.m));
        MGPU_RET(upload(P.d_noise + o, it.noise, sizeof(double) * (size_t) it.m));
        MGPU_RET(upload(d_inv + irow, it.invtt, sizeof(double) * (size_t) it.nb * it.lambda));
        irow += (int64_t) it.nb * it.lambda;
        k_batch_pix<<<grid_for(3 * (int64_t) it.m, 256), 256, 0, s>>>(d_pix + 3 * o, 3 * (int64_t) it.m,
                                                                      (int) (3 * pix0[r]), 3 * grid[r], d_err);
        MGPU_LAUNCHED();
        for (int b = 0; b < it.nb; b++) {
            sizes.push_back(it.local_blocks_sizes[b]);
            blam.push_back(it.lambda);
            bl2.push_back(l2[r]);
        }
        const int64_t pad = off[r + 1] - o - it.m;
        if (pad > 0) {
            k_batch_pad<<<grid_for(pad, 256), 256, 0, s>>>(d_pix, d_w, P.d_sig, P.d_noise, o + it.m, off[r + 1]);
            MGPU_LAUNCHED();
            // the padding block reuses the item's first filter (zero data: any band gives zero)
            MGPU_HIP(hipMemcpyAsync(d_inv + irow, it.invtt, sizeof(double) * it.lambda, kind, s));
            irow += it.lambda;
            sizes.push_back((int) pad);
            blam.push_back(it.lambda);
            bl2.push_back(l2[r]);
        }
    }
    int herr = 0;
    MGPU_HIP(hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipStreamSynchronize(s));
    MGPU_ARG(!herr, "a pixel index is >= its item's grid_npix");

    // set-up time: the span of a solo solve's t_setup_ms (pointing build, Toeplitz
    // plan, preconditioner), without the input uploads
    MGPU_HIP(hipEventRecord(ev.a, s));
    MGPU_RET(pointing_build(P, d_pix, d_w, M, 3));
    P.arena.release(d_pix);
    P.arena.release(d_w);
    MGPU_RET(toeplitz_plan(P, sizes.data(), (int) sizes.size(), lam_max, d_inv, nullptr, blam.data(), bl2.data()));
    P.arena.release(d_inv);
    // one segment per item; pixels seen per item before the cut
    P.sg.h_pix0 = pix0;
    P.sg.npix = -1;
    MGPU_RET(segments_ensure(P));
    const std::vector<int> seen0 = P.sg.h_q;
    MGPU_RET(ensure_work(P));
    MGPU_RET(precond_build(P, P.d_sig));
    MGPU_HIP(hipEventRecord(ev.b, s));
    MGPU_HIP(hipEventSynchronize(ev.b));
    float setup_ms = 0.f;
    MGPU_HIP(hipEventElapsedTime(&setup_ms, ev.a, ev.b));

    // results per item, delivered as soon as the item's segment stops (the
    // others keep iterating); done(i) tells the caller that item i is complete
    const int npix = P.pt.npix;
    std::vector<int> pixg((size_t) std::max(npix, 1));
    MGPU_HIP(hipMemcpyAsync(pixg.data(), P.pt.d_pixg, sizeof(int) * npix, hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipStreamSynchronize(s));
    const SegDone finish = [&](int r, const SegReport &o, const double *d_x) -> int {
        mgpu_batch_item &it = items[idx[r]];
        const int np = o.q1 - o.q0;
        MGPU_ARG(!it.x || it.x_cap >= 3 * np, "item %d: x_cap=%d < %d map values", idx[r], it.x_cap, 3 * np);
        if (it.x)
            MGPU_HIP(hipMemcpyAsync(it.x, d_x + 3 * (int64_t) o.q0, sizeof(double) * 3 * (size_t) np,
                                    hipMemcpyDeviceToHost, s));
        if (it.cond)
            MGPU_HIP(hipMemcpyAsync(it.cond, P.pc.d_cond + o.q0, sizeof(double) * (size_t) np, hipMemcpyDeviceToHost,
                                    s));
        if (it.lhits)
            MGPU_HIP(hipMemcpyAsync(it.lhits, P.pc.d_hits + o.q0, sizeof(int) * (size_t) np, hipMemcpyDeviceToHost,
                                    s));
        if (it.pixels)
            for (int i = 0; i < np; i++) it.pixels[i] = (int) (pixg[(size_t) o.q0 + i] - pix0[r]);
        mgpu_report &ir = it.rep;
        ir = mgpu_report{};
        ir.iterations = o.iterations;
        ir.converged = o.converged;
        ir.res0 = o.res0;
        ir.res = o.res;
        ir.n_map = 3 * np;
        ir.lcount = 3 * np;
        ir.trash_pix = 0;
        ir.n_cut = (seen0[r + 1] - seen0[r]) - np;
        ir.t_setup_ms = setup_ms;
        ir.true_res_rel = o.true_res_rel;
        ir.tile_variant = P.tp.variant(P.opt.precision == MGPU_MIXED ? 1 : 0);
        ir.dev_peak_bytes = (long long) P.arena.peak;
        MGPU_HIP(hipStreamSynchronize(s));
        if (done) (*done)(idx[r]);
        return MGPU_OK;
    };
    mgpu_report rep{};
    std::vector<SegReport> segs;
    MGPU_RET(pcg_run(P, P.d_sig, P.d_noise, tol_max, maxit, 0, nullptr, nullptr, &rep, &segs, &tol, &maxiter,
                     &finish));
    for (int r = 0; r < R; r++) {
        mgpu_report &ir = items[idx[r]].rep;
        ir.t_solve_ms = rep.t_solve_ms;
        ir.refreshes = rep.refreshes;
        ir.true_res_rel = segs[r].true_res_rel;
        ir.dev_peak_bytes = (long long) P.arena.peak;
    }
    MGPU_RET(kernel_times_accumulate(P));
    return MGPU_OK;
}

// check one item; its blocks must cover it exactly
int check_item(const mgpu_batch_item &it, int i) {
    MGPU_ARG(it.pix && it.pixweights && it.signal && it.noise && it.invtt, "item %d: NULL input", i);
    MGPU_RET(check_common(it.m, 3, it.nb, it.local_blocks_sizes, it.lambda));
    int64_t tot = 0;
    for (int b = 0; b < it.nb; b++) {
        MGPU_ARG(it.local_blocks_sizes[b] > 0, "item %d: block %d has size %d", i, b, it.local_blocks_sizes[b]);
        tot += it.local_blocks_sizes[b];
    }
    MGPU_ARG(tot == it.m, "item %d: block sizes sum to %lld, not m = %d", i, (long long) tot, it.m);
    return MGPU_OK;
}

}  // namespace mgpu

using namespace mgpu;

namespace mgpu {

int pcg_batch_impl(mgpu_batch_item *items, const std::vector<int> &all, const std::vector<int> &grid_of,
                   const std::vector<double> &tol_of, const std::vector<int> &maxiter_of, bool on_dev,
                   const mgpu_options *opt, const std::function<void(int)> *done) {
    // one problem for the items with a banded noise model whose tiles fit
    // (all bands and tile lengths together); diagonal noise (lambda = 1), and
    // each band too wide for a tile (direct product), on their own
    std::map<int, std::vector<int>> groups;  // key: 0 = tiled, 1 = diagonal, lambda + 1 = direct
    for (int k = 0; k < (int) all.size(); k++) {
        const mgpu_batch_item &it = items[all[k]];
        int64_t nmax = 0;
        for (int b = 0; b < it.nb; b++) nmax = std::max<int64_t>(nmax, it.local_blocks_sizes[b]);
        const int l2 = it.lambda > 1 ? toeplitz_tile_log2(it.lambda, nmax, opt ? opt->tile_log2 : 0) : 0;
        const int key = it.lambda == 1 ? 1 : (l2 > 0 ? 0 : it.lambda + 1);
        groups[key].push_back(k);
    }
    // groups one after the other: each fills the device on its own, and no
    // graph capture ever overlaps allocations made by another host thread
    for (auto &kv : groups) {
        std::vector<int> idx, grid, mx;
        std::vector<double> tl;
        for (int k : kv.second) {
            idx.push_back(all[k]);
            grid.push_back(grid_of[k]);
            tl.push_back(tol_of[k]);
            mx.push_back(maxiter_of[k]);
        }
        const int rc = solve_items(items, idx, grid, tl, mx, on_dev, opt, done);
        if (rc != MGPU_OK) {
            const std::string err = mappraiser_gpu_last_error();
            set_error("batch of %zu item(s) (%s): %s", idx.size(),
                      kv.first == 0 ? "banded noise" : kv.first == 1 ? "diagonal noise" : "direct product",
                      err.c_str());
            return rc;
        }
    }
    return MGPU_OK;
}

}  // namespace mgpu

extern "C" int mappraiser_gpu_pcg_batch(int nitems, mgpu_batch_item *items, int grid_npix, int inputs_on_device,
                                        double tol, int maxiter, const mgpu_options *opt) {
    MGPU_ARG(nitems > 0 && items, "empty batch");
    MGPU_ARG(grid_npix > 0, "grid_npix=%d must be positive", grid_npix);
    MGPU_ARG(maxiter >= 1, "maxiter=%d must be >= 1", maxiter);
    MGPU_ARG(!opt || opt->precond == 0,
             "precond=%d: batches use block Jacobi (two-level preconditioners solve one problem at a time)",
             opt ? opt->precond : 0);
    std::vector<int> all((size_t) nitems), grid((size_t) nitems, grid_npix), mx((size_t) nitems, maxiter);
    std::vector<double> tl((size_t) nitems, tol);
    for (int i = 0; i < nitems; i++) {
        MGPU_RET(check_item(items[i], i));
        all[i] = i;
    }
    return pcg_batch_impl(items, all, grid, tl, mx, inputs_on_device != 0, opt, nullptr);
}
