// batch.hip -- many independent GLS problems solved as one segmented problem.
//
// Items of equal lambda and equal tile length (the one each would use alone)
// are concatenated into one problem:
//   * samples: item r occupies [off_r, off_r + m_r) of the TOD, off_r a multiple
//     of kChunk, followed by padding samples up to the next multiple.  Padding
//     samples are flagged (-3,-2,-1), carry weights (1,0,0) and zero signal and
//     noise, and form one extra Toeplitz block of their own, so every TOD value
//     they take is exactly zero.  Chunk alignment makes the fused per-block sums
//     of P group the same samples as in the item's solo solve.
//   * pixels: item r's pixel p becomes r * grid_npix + p, so items never share
//     a pixel; after the pointing build the local pixels of item r are one
//     contiguous range, a map segment of the PCG (pcg.hip).
// P, N^-1 and P^T are block diagonal over the items, the preconditioner and
// the cut are per pixel and the PCG scalars are per segment, so each item
// follows exactly the arithmetic of its solo solve.  Groups (different lambda
// or tile length) are solved one after the other.
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

struct Group {
    int lambda = 0;
    int tile_log2 = 0;       // tile length every item of the group uses alone
    std::vector<int> items;  // indices into the caller's array
    int rc = MGPU_OK;
    std::string err;
};

int solve_group(Group &gr, mgpu_batch_item *items, int grid_npix, bool on_dev, double tol, int maxiter,
                const mgpu_options *opt) {
    const int R = (int) gr.items.size();
    const int lam = gr.lambda;
    const hipMemcpyKind kind = on_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    // layout
    std::vector<int64_t> off((size_t) R + 1, 0);
    int nb_tot = 0;
    for (int r = 0; r < R; r++) {
        const mgpu_batch_item &it = items[gr.items[r]];
        const int64_t padded = (it.m + kChunk - 1) / kChunk * kChunk;
        off[r + 1] = off[r] + padded;
        nb_tot += it.nb + (padded > it.m ? 1 : 0);
    }
    const int64_t M = off[R];
    MGPU_ARG(M < (int64_t) INT32_MAX, "batch of %lld padded samples exceeds 2^31-1: split it", (long long) M);
    MGPU_ARG(3 * (int64_t) grid_npix * R < (int64_t) INT32_MAX - 3,
             "%d items x grid_npix %d exceed the int32 pixel index range: split the batch", R, grid_npix);

    // the tile length each item would use alone (same block-length profile class)
    mgpu_options gopt;
    mappraiser_gpu_options_default(&gopt);
    if (opt) gopt = *opt;
    gopt.tile_log2 = gr.tile_log2;
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
    MGPU_RET(P.arena.alloc(&d_inv, (size_t) nb_tot * lam));
    MGPU_RET(P.arena.alloc(&d_err, 1));
    MGPU_HIP(hipMemsetAsync(d_err, 0, sizeof(int), s));
    std::vector<int> sizes;
    sizes.reserve((size_t) nb_tot);
    int64_t brow = 0;
    for (int r = 0; r < R; r++) {
        const mgpu_batch_item &it = items[gr.items[r]];
        const int64_t o = off[r];
        MGPU_HIP(hipMemcpyAsync(d_pix + 3 * o, it.pix, sizeof(int) * 3 * (size_t) it.m, kind, s));
        MGPU_HIP(hipMemcpyAsync(d_w + 3 * o, it.pixweights, sizeof(double) * 3 * (size_t) it.m, kind, s));
        MGPU_HIP(hipMemcpyAsync(P.d_sig + o, it.signal, sizeof(double) * (size_t) it.m, kind, s));
        MGPU_HIP(hipMemcpyAsync(P.d_noise + o, it.noise, sizeof(double) * (size_t) it.m, kind, s));
        MGPU_HIP(hipMemcpyAsync(d_inv + brow * lam, it.invtt, sizeof(double) * (size_t) it.nb * lam, kind, s));
        k_batch_pix<<<grid_for(3 * (int64_t) it.m, 256), 256, 0, s>>>(d_pix + 3 * o, 3 * (int64_t) it.m,
                                                                      3 * r * grid_npix, 3 * grid_npix, d_err);
        MGPU_LAUNCHED();
        for (int b = 0; b < it.nb; b++) sizes.push_back(it.local_blocks_sizes[b]);
        brow += it.nb;
        const int64_t pad = off[r + 1] - o - it.m;
        if (pad > 0) {
            k_batch_pad<<<grid_for(pad, 256), 256, 0, s>>>(d_pix, d_w, P.d_sig, P.d_noise, o + it.m, off[r + 1]);
            MGPU_LAUNCHED();
            // the padding block reuses the item's first filter (zero data: any band gives zero)
            MGPU_HIP(hipMemcpyAsync(d_inv + brow * lam, it.invtt, sizeof(double) * lam, kind, s));
            sizes.push_back((int) pad);
            brow++;
        }
    }
    int herr = 0;
    MGPU_HIP(hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipStreamSynchronize(s));
    MGPU_ARG(!herr, "a pixel index is >= grid_npix (%d)", grid_npix);

    // set-up time: the span of a solo solve's t_setup_ms (pointing build, Toeplitz
    // plan, preconditioner), without the input uploads
    MGPU_HIP(hipEventRecord(ev.a, s));
    MGPU_RET(pointing_build(P, d_pix, d_w, M, 3));
    P.arena.release(d_pix);
    P.arena.release(d_w);
    MGPU_RET(toeplitz_plan(P, sizes.data(), (int) sizes.size(), lam, d_inv, nullptr));
    P.arena.release(d_inv);
    // one segment per item; pixels seen per item before the cut
    P.sg.stride = grid_npix;
    P.sg.nseg = R;
    P.sg.npix = -1;
    MGPU_RET(segments_ensure(P));
    const std::vector<int> seen0 = P.sg.h_q;
    MGPU_RET(ensure_work(P));
    MGPU_RET(precond_build(P, P.d_sig));
    MGPU_HIP(hipEventRecord(ev.b, s));
    MGPU_HIP(hipEventSynchronize(ev.b));
    float setup_ms = 0.f;
    MGPU_HIP(hipEventElapsedTime(&setup_ms, ev.a, ev.b));

    mgpu_report rep{};
    std::vector<SegReport> segs;
    MGPU_RET(pcg_run(P, P.d_sig, P.d_noise, tol, maxiter, 0, nullptr, nullptr, &rep, &segs));

    // results per item
    const int npix = P.pt.npix;
    std::vector<int> pixg((size_t) std::max(npix, 1));
    MGPU_HIP(hipMemcpyAsync(pixg.data(), P.pt.d_pixg, sizeof(int) * npix, hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipStreamSynchronize(s));
    for (int r = 0; r < R; r++) {
        mgpu_batch_item &it = items[gr.items[r]];
        const SegReport &o = segs[r];
        const int np = o.q1 - o.q0;
        MGPU_ARG(!it.x || it.x_cap >= 3 * np, "item %d: x_cap=%d < %d map values", gr.items[r], it.x_cap, 3 * np);
        if (it.x)
            MGPU_HIP(hipMemcpyAsync(it.x, P.d_map[0] + 3 * (int64_t) o.q0, sizeof(double) * 3 * (size_t) np,
                                    hipMemcpyDeviceToHost, s));
        if (it.cond)
            MGPU_HIP(hipMemcpyAsync(it.cond, P.pc.d_cond + o.q0, sizeof(double) * (size_t) np, hipMemcpyDeviceToHost,
                                    s));
        if (it.lhits)
            MGPU_HIP(hipMemcpyAsync(it.lhits, P.pc.d_hits + o.q0, sizeof(int) * (size_t) np, hipMemcpyDeviceToHost,
                                    s));
        if (it.pixels)
            for (int i = 0; i < np; i++) it.pixels[i] = pixg[(size_t) o.q0 + i] - r * grid_npix;
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
        ir.t_solve_ms = rep.t_solve_ms;
        ir.true_res_rel = o.true_res_rel;
        ir.refreshes = rep.refreshes;
        ir.tile_variant = rep.tile_variant;
    }
    MGPU_HIP(hipStreamSynchronize(s));
    return MGPU_OK;
}

}  // namespace
}  // namespace mgpu

using namespace mgpu;

extern "C" int mappraiser_gpu_pcg_batch(int nitems, mgpu_batch_item *items, int grid_npix, int inputs_on_device,
                                        double tol, int maxiter, const mgpu_options *opt) {
    MGPU_ARG(nitems > 0 && items, "empty batch");
    MGPU_ARG(grid_npix > 0, "grid_npix=%d must be positive", grid_npix);
    MGPU_ARG(maxiter >= 1, "maxiter=%d must be >= 1", maxiter);
    MGPU_ARG(!opt || opt->precond == 0,
             "precond=%d: batches use block Jacobi (two-level preconditioners solve one problem at a time)",
             opt ? opt->precond : 0);
    std::map<std::pair<int, int>, Group> groups;
    for (int i = 0; i < nitems; i++) {
        const mgpu_batch_item &it = items[i];
        MGPU_ARG(it.pix && it.pixweights && it.signal && it.noise && it.invtt, "item %d: NULL input", i);
        MGPU_RET(check_common(it.m, 3, it.nb, it.local_blocks_sizes, it.lambda));
        int64_t tot = 0, nmax = 0;
        for (int b = 0; b < it.nb; b++) {
            MGPU_ARG(it.local_blocks_sizes[b] > 0, "item %d: block %d has size %d", i, b, it.local_blocks_sizes[b]);
            tot += it.local_blocks_sizes[b];
            nmax = std::max<int64_t>(nmax, it.local_blocks_sizes[b]);
        }
        MGPU_ARG(tot == it.m, "item %d: block sizes sum to %lld, not m = %d", i, (long long) tot, it.m);
        // items are grouped by lambda and by the tile length they would use alone
        const int l2 = toeplitz_tile_log2(it.lambda, nmax, opt ? opt->tile_log2 : 0);
        Group &g = groups[std::make_pair(it.lambda, l2)];
        g.lambda = it.lambda;
        g.tile_log2 = l2;
        g.items.push_back(i);
    }
    std::vector<Group *> gl;
    for (auto &kv : groups) gl.push_back(&kv.second);
    auto run = [&](Group *g) {
        g->rc = solve_group(*g, items, grid_npix, inputs_on_device != 0, tol, maxiter, opt);
        if (g->rc != MGPU_OK) g->err = mappraiser_gpu_last_error();
    };
    // groups one after the other: each fills the device on its own, and no
    // graph capture ever overlaps allocations made by another host thread
    for (Group *g : gl) run(g);
    for (Group *g : gl)
        if (g->rc != MGPU_OK) {
            set_error("lambda=%d, tile 2^%d group: %s", g->lambda, g->tile_log2, g->err.c_str());
            return g->rc;
        }
    return MGPU_OK;
}
