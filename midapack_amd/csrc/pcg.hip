// pcg.hip -- PCG_GLS_true on the device (gfx950).
//
// Reference: pcg_true.c:20-241 (Polak-Ribiere PCG, x0 = 0, block-Jacobi M).
// Per iteration k:
//   swap(g, gp); Ah = P h; N^-1 Ah; q = P^T N^-1 A h
//   coeff = <h,q>; ro = g2pix/coeff; x += ro h; g = gp - ro q; Cg = M g
//   g2pixp = g2pix; g2pix = <Cg,g>; g2pix_polak = <Cg,g-gp>; res = <g,g>
//   stop if res <= tol^2 res0 (pcg_true.c:200); gamma = g2pix_polak/g2pixp; h = gamma h + Cg
// (the absolute test at k=0, pcg_true.c:127, is overridden by the loop header
// and has no effect; the loop therefore always runs.)
//
// Device design.  All vectors stay in HBM; the host only launches.  Every
// kernel of an iteration first reads a device stop flag and returns when it is
// set, so the host queues a block of iterations at a time and reads the flag
// once per block.  Dot products are fused into the producing kernels as
// per-workgroup partials and reduced by one workgroup per segment in a fixed
// order: results are deterministic and independent of what else runs on the
// GPU.  With opt.graph the block of iterations is captured once into a HIP
// graph and replayed (the kernels and their order are unchanged, so the
// results are bit-identical to the eager launches).
//
// Segments (batched problems): the map is split into independent segments,
// each with its own scalars (ro, gamma, residuals), stop test and iteration
// count; a converged segment is frozen while the others continue.  Every
// map-domain workgroup lies in one segment (Segments in mgpu_internal.h), so
// a segment's dot products are summed exactly as in a solo solve.
//
// MGPU_MIXED: P, N^-1 and P^T run in FP32 on FP32 copies of the pointing
// weights and TOD; map vectors, dot products and the 3x3 blocks stay FP64,
// and every `refresh_every` iterations g is recomputed in FP64 as b - A x.
#include "pcg_device.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace mgpu {


// a segment is finished: count it, and raise the global stop flag with the last one
__device__ __forceinline__ void segment_done(int *flag, int nseg) {
    if (atomicAdd(&flag[F_NDONE], 1) + 1 == nseg) atomicExch(&flag[F_STOP], 1);
}

// tod = s + n (the first residual with x0 = 0, pcg_true.c:78-80)
__global__ __launch_bounds__(256) void k_rhs_tod(const double *__restrict__ s, const double *__restrict__ n,
                                                 double *__restrict__ tod, int64_t m) {
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) tod[t] = (s[t] + n[t]) - 0.0;
}

__global__ __launch_bounds__(256) void k_map_to_f32(const double *__restrict__ a, float *__restrict__ b, int n,
                                                    const int *__restrict__ flag) {
    if (flag && stopped(flag)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float) a[i];
}

// init: Cg = M g, h = Cg, partials of <Cg,g> and <g,g>
__global__ __launch_bounds__(kMB) void k_init_dots(const double *__restrict__ inv, const double *__restrict__ g,
                                                   double *__restrict__ Cg, double *__restrict__ h,
                                                   const int *__restrict__ wg_seg, const int *__restrict__ wg_q0,
                                                   const int *__restrict__ seg_q, double *__restrict__ part) {
    __shared__ double red[kMB / 64];
    const SegPix sp = seg_pixel(wg_seg, wg_q0, seg_q);
    const int q = sp.q;
    double a = 0.0, c = 0.0;
    if (q < sp.qend) {
        const double g0 = g[3 * q], g1 = g[3 * q + 1], g2 = g[3 * q + 2];
        const double *M = inv + 9 * (int64_t) q;
        for (int r = 0; r < 3; r++) {
            double y = 0.0;
            y += M[3 * r] * g0;
            y += M[3 * r + 1] * g1;
            y += M[3 * r + 2] * g2;
            Cg[3 * q + r] = y;
            h[3 * q + r] = y;
            const double gr = g[3 * q + r];
            a += y * gr * 1.0;
            c += gr * gr * 1.0;
        }
    }
    a = block_sum(a, red);
    c = block_sum(c, red);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[gridDim.x + blockIdx.x] = c;
    }
}

// one workgroup per segment
__global__ __launch_bounds__(kRB) void k_init_scalars(const double *__restrict__ part, int nwg,
                                                      const int *__restrict__ wg0, double *__restrict__ S,
                                                      double *__restrict__ hist, int maxiter,
                                                      int *__restrict__ state, int nseg, int *__restrict__ flag) {
    __shared__ double red[kRB / 64];
    const int r = blockIdx.x;
    const int p0 = wg0[r], np = wg0[r + 1] - p0;
    const double g2pix = reduce_parts(part + p0, np, red);
    __syncthreads();
    const double res = reduce_parts(part + nwg + p0, np, red);
    if (threadIdx.x == 0) {
        double *Sr = S + (int64_t) r * S_NSCAL;
        Sr[S_G2PIX] = g2pix;
        Sr[S_RES] = res;
        Sr[S_RES0] = res;
        Sr[S_TOL2] = Sr[S_TOL] * Sr[S_TOL] * res;
        Sr[S_K] = 0.0;
        Sr[S_RREF] = res;
        Sr[S_KREF] = 0.0;
        if (hist) hist[(int64_t) r * maxiter] = sqrt(res) / sqrt(res);
        if (np == 0) {  // empty segment (every pixel cut): nothing to solve
            state[r] = 2;
            segment_done(flag, nseg);
        } else {
            state[r] = 0;
        }
    }
}

// ro = g2pix / coeff, one workgroup per segment
__global__ __launch_bounds__(kRB) void k_coeff(const double *__restrict__ part, const int *__restrict__ wg0,
                                               double *__restrict__ S, const int *__restrict__ state,
                                               const int *__restrict__ flag) {
    if (stopped(flag)) return;
    const int r = blockIdx.x;
    if (state[r]) return;
    __shared__ double red[kRB / 64];
    const double coeff = reduce_parts(part + wg0[r], wg0[r + 1] - wg0[r], red);
    if (threadIdx.x == 0) {
        double *Sr = S + (int64_t) r * S_NSCAL;
        Sr[S_COEFF] = coeff;
        Sr[S_RO] = Sr[S_G2PIX] / coeff;
    }
}

// x += ro h; g = gp - ro q; Cg = M g; partials <Cg,g>, <Cg,g-gp>, <g,g>
__global__ __launch_bounds__(kMB) void k_update(double *__restrict__ x, const double *__restrict__ h,
                                                double *__restrict__ g, const double *__restrict__ gp,
                                                const double *__restrict__ qv, const double *__restrict__ inv,
                                                double *__restrict__ Cg, const int *__restrict__ wg_seg,
                                                const int *__restrict__ wg_q0, const int *__restrict__ seg_q,
                                                const double *__restrict__ S, const int *__restrict__ state,
                                                double *__restrict__ part, const int *__restrict__ flag) {
    if (stopped(flag)) return;
    const SegPix sp = seg_pixel(wg_seg, wg_q0, seg_q);
    if (state[sp.r]) return;  // frozen segment (uniform over the workgroup)
    __shared__ double red[kMB / 64];
    const int q = sp.q;
    double a = 0.0, b = 0.0, c = 0.0;
    if (q < sp.qend) {
        const double ro = S[(int64_t) sp.r * S_NSCAL + S_RO];
        double gn[3], gpv[3];
        for (int r = 0; r < 3; r++) {
            const int i = 3 * q + r;
            x[i] = x[i] + ro * h[i];
            gpv[r] = gp[i];
            gn[r] = gpv[r] - ro * qv[i];
            g[i] = gn[r];
        }
        const double *M = inv + 9 * (int64_t) q;
        for (int r = 0; r < 3; r++) {
            double y = 0.0;
            y += M[3 * r] * gn[0];
            y += M[3 * r + 1] * gn[1];
            y += M[3 * r + 2] * gn[2];
            Cg[3 * q + r] = y;
            a += y * gn[r] * 1.0;
            b += y * (gn[r] - gpv[r]) * 1.0;
            c += gn[r] * gn[r] * 1.0;
        }
    }
    a = block_sum(a, red);
    b = block_sum(b, red);
    c = block_sum(c, red);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[gridDim.x + blockIdx.x] = b;
        part[2 * gridDim.x + blockIdx.x] = c;
    }
}

// x += ro h; g = gp - ro q (the two-level preconditioner applies M and the dots afterwards)
__global__ __launch_bounds__(kMB) void k_update_xg(double *__restrict__ x, const double *__restrict__ h,
                                                   double *__restrict__ g, const double *__restrict__ gp,
                                                   const double *__restrict__ qv, const int *__restrict__ wg_seg,
                                                   const int *__restrict__ wg_q0, const int *__restrict__ seg_q,
                                                   const double *__restrict__ S, const int *__restrict__ state,
                                                   const int *__restrict__ flag) {
    if (stopped(flag)) return;
    const SegPix sp = seg_pixel(wg_seg, wg_q0, seg_q);
    if (state[sp.r] || sp.q >= sp.qend) return;
    const double ro = S[(int64_t) sp.r * S_NSCAL + S_RO];
    for (int r = 0; r < 3; r++) {
        const int i = 3 * sp.q + r;
        x[i] = x[i] + ro * h[i];
        g[i] = gp[i] - ro * qv[i];
    }
}

// dots -> g2pix, polak, res; history; stop test; gamma.  One workgroup per segment.
// kind (IT_*): where a segment's g was updated in FP32 this iteration (mixed
// precision, the segment not refreshed), a passing stop test does not stop it:
// it sets S_PEND (counted in F_NPEND) and the first iteration of the next block
// refreshes that segment's residual in FP64, whose test decides.  Whether a
// segment refreshes depends on its own state and the iteration number only,
// so a segment runs the same iterations alone and in a batch.
__global__ __launch_bounds__(kRB) void k_scalars(const double *__restrict__ part, int nwg,
                                                 const int *__restrict__ wg0, double *__restrict__ S,
                                                 double *__restrict__ hist, int maxiter, int *__restrict__ state,
                                                 int nseg, int *__restrict__ flag, int kind, int adaptive) {
    if (stopped(flag)) return;
    const int r = blockIdx.x;
    if (state[r]) return;
    __shared__ double red[kRB / 64];
    const int p0 = wg0[r], np = wg0[r + 1] - p0;
    const double g2pix = reduce_parts(part + p0, np, red);
    __syncthreads();
    const double polak = reduce_parts(part + nwg + p0, np, red);
    __syncthreads();
    const double res = reduce_parts(part + 2 * nwg + p0, np, red);
    if (threadIdx.x == 0) {
        double *Sr = S + (int64_t) r * S_NSCAL;
        const double k = Sr[S_K] + 1.0;
        Sr[S_K] = k;
        Sr[S_G2PIXP] = Sr[S_G2PIX];
        Sr[S_G2PIX] = g2pix;
        Sr[S_POLAK] = polak;
        Sr[S_RES] = res;
        const int ki = (int) k;
        if (hist && ki < maxiter) hist[(int64_t) r * maxiter + ki] = sqrt(res) / sqrt(Sr[S_RES0]);
        atomicMax(&flag[F_KMAX], ki);
        const bool pend = Sr[S_PEND] != 0.0;
        const bool refreshed = kind == IT_REFRESH_ALL || (kind == IT_REFRESH_PENDING && pend);
        const bool verify = kind != IT_FP64 && !refreshed;
        bool pend_now = pend && !refreshed;
        if (refreshed) {
            Sr[S_RREF] = res;
            Sr[S_KREF] = k;
        }
        int st = 0;
        if (res <= Sr[S_TOL2] && !verify)
            st = 1;
        else if ((Sr[S_STOPK] > 0.0 && k >= Sr[S_STOPK]) || k + 1.0 >= Sr[S_MAXK])
            st = 2;
        if (res <= Sr[S_TOL2] && verify) pend_now = true;
        if (adaptive && verify && (res <= kAdaptDrop * Sr[S_RREF] || k - Sr[S_KREF] >= kAdaptGap)) pend_now = true;
        if (st) pend_now = false;
        if (pend_now != pend) {
            Sr[S_PEND] = pend_now ? 1.0 : 0.0;
            atomicAdd(&flag[F_NPEND], pend_now ? 1 : -1);
        }
        if (st) {
            state[r] = st;
            if (r == 0 && st == 1) flag[F_CONV0] = 1;
            segment_done(flag, nseg);
        }
        Sr[S_GAMMA] = polak / Sr[S_G2PIXP];
    }
}

// h = h gamma + Cg
__global__ __launch_bounds__(kMB) void k_hupdate(double *__restrict__ h, const double *__restrict__ Cg,
                                                 const int *__restrict__ wg_seg, const int *__restrict__ wg_q0,
                                                 const int *__restrict__ seg_q, const double *__restrict__ S,
                                                 const int *__restrict__ state, const int *__restrict__ flag) {
    if (stopped(flag)) return;
    const SegPix sp = seg_pixel(wg_seg, wg_q0, seg_q);
    if (state[sp.r] || sp.q >= sp.qend) return;
    const double gamma = S[(int64_t) sp.r * S_NSCAL + S_GAMMA];
    for (int c = 0; c < 3; c++) {
        const int i = 3 * sp.q + c;
        h[i] = h[i] * gamma + Cg[i];
    }
}

// segment r skips the refresh of an IT_REFRESH_PENDING iteration unless it is pending
__device__ __forceinline__ bool skip_refresh(const double *__restrict__ S, int kind, int r) {
    return kind == IT_REFRESH_PENDING && S[(int64_t) r * S_NSCAL + S_PEND] == 0.0;
}

// segments an IT_REFRESH_PENDING iteration's FP64 operator skips: stopped or not pending
__global__ void k_refresh_mask(const int *__restrict__ state, const double *__restrict__ S, int nseg,
                               int *__restrict__ rstate) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nseg) rstate[r] = state[r] != 0 || S[(int64_t) r * S_NSCAL + S_PEND] == 0.0;
}

// FP64 refresh of the residual: g = b - Ax (mixed precision); state == nullptr
// computes every segment (final true residual)
__global__ __launch_bounds__(kMB) void k_refresh(const double *__restrict__ b, const double *__restrict__ Ax,
                                                 double *__restrict__ g, const int *__restrict__ wg_seg,
                                                 const int *__restrict__ wg_q0, const int *__restrict__ seg_q,
                                                 const int *__restrict__ state, const int *__restrict__ flag,
                                                 int *__restrict__ nref, const double *__restrict__ S, int kind) {
    if (stopped(flag)) return;
    if (nref && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(nref, 1);
    const SegPix sp = seg_pixel(wg_seg, wg_q0, seg_q);
    if ((state && state[sp.r]) || sp.q >= sp.qend) return;
    if (S && skip_refresh(S, kind, sp.r)) return;
    for (int c = 0; c < 3; c++) {
        const int i = 3 * sp.q + c;
        g[i] = b[i] - Ax[i];
    }
}

// refreshed residual: Cg = M g and the three dots <Cg,g>, <Cg,g-gp>, <g,g>
__global__ __launch_bounds__(kMB) void k_refresh_three(const double *__restrict__ inv, const double *__restrict__ g,
                                                       const double *__restrict__ gp, double *__restrict__ Cg,
                                                       const int *__restrict__ wg_seg, const int *__restrict__ wg_q0,
                                                       const int *__restrict__ seg_q, const int *__restrict__ state,
                                                       double *__restrict__ part, const int *__restrict__ flag,
                                                       const double *__restrict__ S, int kind) {
    if (stopped(flag)) return;
    const SegPix sp = seg_pixel(wg_seg, wg_q0, seg_q);
    // segments not refreshed keep the partials of k_update (same g, same values)
    if (state[sp.r] || skip_refresh(S, kind, sp.r)) return;
    __shared__ double red[kMB / 64];
    const int q = sp.q;
    double a = 0.0, b = 0.0, c = 0.0;
    if (q < sp.qend) {
        const double g0 = g[3 * q], g1 = g[3 * q + 1], g2 = g[3 * q + 2];
        const double *M = inv + 9 * (int64_t) q;
        for (int r = 0; r < 3; r++) {
            double y = 0.0;
            y += M[3 * r] * g0;
            y += M[3 * r + 1] * g1;
            y += M[3 * r + 2] * g2;
            Cg[3 * q + r] = y;
            const double gr = g[3 * q + r];
            a += y * gr * 1.0;
            b += y * (gr - gp[3 * q + r]) * 1.0;
            c += gr * gr * 1.0;
        }
    }
    a = block_sum(a, red);
    b = block_sum(b, red);
    c = block_sum(c, red);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[gridDim.x + blockIdx.x] = b;
        part[2 * gridDim.x + blockIdx.x] = c;
    }
}

// partials of <r,r> and <b,b> (true residual of the returned map)
__global__ __launch_bounds__(kMB) void k_two_norms(const double *__restrict__ rv, const double *__restrict__ b,
                                                   const int *__restrict__ wg_seg, const int *__restrict__ wg_q0,
                                                   const int *__restrict__ seg_q, double *__restrict__ part) {
    __shared__ double red[kMB / 64];
    const SegPix sp = seg_pixel(wg_seg, wg_q0, seg_q);
    double a = 0.0, c = 0.0;
    if (sp.q < sp.qend)
        for (int k = 0; k < 3; k++) {
            const int i = 3 * sp.q + k;
            a += rv[i] * rv[i];
            c += b[i] * b[i];
        }
    a = block_sum(a, red);
    c = block_sum(c, red);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[gridDim.x + blockIdx.x] = c;
    }
}

__global__ __launch_bounds__(kRB) void k_seg_norms(const double *__restrict__ part, int nwg,
                                                   const int *__restrict__ wg0, double *__restrict__ S) {
    __shared__ double red[kRB / 64];
    const int r = blockIdx.x;
    const int p0 = wg0[r], np = wg0[r + 1] - p0;
    const double a = reduce_parts(part + p0, np, red);
    __syncthreads();
    const double c = reduce_parts(part + nwg + p0, np, red);
    if (threadIdx.x == 0) {
        S[(int64_t) r * S_NSCAL + S_TRUE_R2] = a;
        S[(int64_t) r * S_NSCAL + S_TRUE_B2] = c;
    }
}

// first local pixel of every segment: lower bound of its first global pixel
// (pix0[r]) in the sorted pixel ids
__global__ void k_seg_bounds(const int *__restrict__ pixg, int npix, const int64_t *__restrict__ pix0, int nseg,
                             int *__restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nseg) return;
    if (r == nseg) {
        out[r] = npix;
        return;
    }
    const int64_t key = pix0[r];
    int lo = 0, hi = npix;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int64_t) pixg[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    out[r] = lo;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int segments_ensure(Problem &pb) {
    Segments &G = pb.sg;
    const int npix = pb.pt.npix;
    if (G.npix == npix && G.d_wg_seg) return MGPU_OK;
    hipStream_t s = pb.stream;
    const int nseg = G.h_pix0.size() > 1 ? (int) G.h_pix0.size() - 1 : 1;
    G.h_q.assign((size_t) nseg + 1, 0);
    if (nseg == 1) {
        G.h_q[1] = npix;
    } else {
        int *d_b;
        int64_t *d_k;
        MGPU_RET(pb.arena.alloc(&d_b, (size_t) nseg + 1));
        MGPU_RET(pb.arena.alloc(&d_k, (size_t) nseg));
        MGPU_HIP(hipMemcpyAsync(d_k, G.h_pix0.data(), sizeof(int64_t) * nseg, hipMemcpyHostToDevice, s));
        k_seg_bounds<<<grid_for(nseg + 1, 64), 64, 0, s>>>(pb.pt.d_pixg, npix, d_k, nseg, d_b);
        MGPU_LAUNCHED();
        MGPU_HIP(hipMemcpyAsync(G.h_q.data(), d_b, sizeof(int) * (nseg + 1), hipMemcpyDeviceToHost, s));
        MGPU_HIP(hipStreamSynchronize(s));
        pb.arena.release(d_b);
        pb.arena.release(d_k);
    }
    G.h_wg0.assign((size_t) nseg + 1, 0);
    std::vector<int> wseg, wq0;
    for (int r = 0; r < nseg; r++) {
        G.h_wg0[r] = (int) wseg.size();
        for (int a = G.h_q[r]; a < G.h_q[r + 1]; a += kMapWG) {
            wseg.push_back(r);
            wq0.push_back(a);
        }
    }
    G.h_wg0[nseg] = (int) wseg.size();
    G.nwg = (int) wseg.size();
    for (int **p : {&G.d_q, &G.d_wg0, &G.d_wg_seg, &G.d_wg_q0})
        if (*p) {
            pb.arena.release(*p);
            *p = nullptr;
        }
    if (G.d_state && G.nseg != nseg) {
        pb.arena.release(G.d_state);
        G.d_state = nullptr;
        if (G.d_rstate) pb.arena.release(G.d_rstate);
        G.d_rstate = nullptr;
    }
    G.nseg = nseg;
    const size_t nw = std::max<size_t>(wseg.size(), 1);
    MGPU_RET(pb.arena.alloc(&G.d_q, (size_t) nseg + 1));
    MGPU_RET(pb.arena.alloc(&G.d_wg0, (size_t) nseg + 1));
    MGPU_RET(pb.arena.alloc(&G.d_wg_seg, nw));
    MGPU_RET(pb.arena.alloc(&G.d_wg_q0, nw));
    if (!G.d_state) MGPU_RET(pb.arena.alloc(&G.d_state, (size_t) nseg));
    if (!G.d_rstate) MGPU_RET(pb.arena.alloc(&G.d_rstate, (size_t) nseg));
    MGPU_HIP(hipMemcpyAsync(G.d_q, G.h_q.data(), sizeof(int) * (nseg + 1), hipMemcpyHostToDevice, s));
    MGPU_HIP(hipMemcpyAsync(G.d_wg0, G.h_wg0.data(), sizeof(int) * (nseg + 1), hipMemcpyHostToDevice, s));
    if (!wseg.empty()) {
        MGPU_HIP(hipMemcpyAsync(G.d_wg_seg, wseg.data(), sizeof(int) * wseg.size(), hipMemcpyHostToDevice, s));
        MGPU_HIP(hipMemcpyAsync(G.d_wg_q0, wq0.data(), sizeof(int) * wq0.size(), hipMemcpyHostToDevice, s));
    }
    MGPU_HIP(hipStreamSynchronize(s));
    // segment of every P chunk (by its smallest pixel) and Toeplitz block (by its
    // first chunk): batch items are padded to whole chunks, so neither straddles two
    std::vector<int> key;
    MGPU_RET(pointing_chunk_keys(pb, key));
    const int nch = (int) key.size();
    std::vector<int> cseg((size_t) nch, -1);
    for (int c = 0; c < nch; c++)
        if (key[c] != INT32_MAX)
            cseg[c] = (int) (std::upper_bound(G.h_q.begin(), G.h_q.end(), key[c]) - G.h_q.begin()) - 1;
    const ToeplitzPlan &T = pb.tp;
    std::vector<int> bseg((size_t) std::max(T.nb, 1), -1);
    for (int b = 0; b < T.nb; b++) {
        const int64_t c = T.h_start[b] / kChunk;
        if (c < nch) bseg[b] = cseg[c];
    }
    for (int **p : {&G.d_chunk_seg, &G.d_block_seg})
        if (*p) {
            pb.arena.release(*p);
            *p = nullptr;
        }
    MGPU_RET(pb.arena.alloc(&G.d_chunk_seg, (size_t) nch));
    MGPU_RET(pb.arena.alloc(&G.d_block_seg, bseg.size()));
    MGPU_HIP(hipMemcpyAsync(G.d_chunk_seg, cseg.data(), sizeof(int) * nch, hipMemcpyHostToDevice, s));
    MGPU_HIP(hipMemcpyAsync(G.d_block_seg, bseg.data(), sizeof(int) * bseg.size(), hipMemcpyHostToDevice, s));
    MGPU_HIP(hipStreamSynchronize(s));
    G.h_chunk_seg = std::move(cseg);
    G.h_block_seg = std::move(bseg);
    G.npix = npix;
    return MGPU_OK;
}

int run_lists_update(Problem &pb, const std::vector<int> &state) {
    const Segments &G = pb.sg;
    if (G.h_chunk_seg.empty()) return MGPU_OK;
    hipStream_t s = pb.stream;
    auto running = [&](int sg) { return sg < 0 || state[sg] == 0; };
    // the lists are built on the host and copied while the stream is idle
    // (between blocks of iterations); the host vectors live until the sync below
    std::vector<std::vector<int>> keep_perm;
    std::vector<std::vector<TileDesc>> keep_tiles;
    Pointing &P = pb.pt;
    if (!P.h_chunk_order.empty()) {
        std::vector<int> ord;
        ord.reserve(P.h_chunk_order.size());
        for (int c : P.h_chunk_order)
            if (running(G.h_chunk_seg[c])) ord.push_back(c);
        if (!P.d_perm_run) MGPU_RET(pb.arena.alloc(&P.d_perm_run, P.h_chunk_order.size()));
        keep_perm.push_back(deal_to_xcds(ord));
        const std::vector<int> &perm = keep_perm.back();
        if (!perm.empty())
            MGPU_HIP(hipMemcpyAsync(P.d_perm_run, perm.data(), sizeof(int) * perm.size(), hipMemcpyHostToDevice, s));
        P.nch_run = (int) perm.size();
    }
    auto tiles = [&](const std::vector<TileDesc> &all, TileDesc *&d, int &n) -> int {
        if (all.empty()) return MGPU_OK;
        std::vector<TileDesc> run;
        run.reserve(all.size());
        for (const TileDesc &t : all)
            if (running(G.h_block_seg[t.block])) run.push_back(t);
        if (!d) MGPU_RET(pb.arena.alloc(&d, all.size()));
        keep_tiles.push_back(deal_to_xcds(run));
        const std::vector<TileDesc> &o = keep_tiles.back();
        if (!o.empty()) MGPU_HIP(hipMemcpyAsync(d, o.data(), sizeof(TileDesc) * o.size(), hipMemcpyHostToDevice, s));
        n = (int) o.size();
        return MGPU_OK;
    };
    for (TileClass &C : pb.tp.cls) {
        MGPU_RET(tiles(C.h_tiles, C.d_tiles_run, C.ntiles_run));
        MGPU_RET(tiles(C.h_tiles32, C.d_tiles32_run, C.ntiles32_run));
    }
    MGPU_HIP(hipStreamSynchronize(s));
    return MGPU_OK;
}

void run_lists_clear(Problem &pb) {
    pb.pt.nch_run = -1;
    for (TileClass &C : pb.tp.cls) {
        C.ntiles_run = -1;
        C.ntiles32_run = -1;
    }
}

int ensure_work(Problem &pb) {
    const int64_t m = pb.pt.m;
    const int n0 = 3 * pb.pt.npix0;
    if (!pb.d_tod) {
        MGPU_RET(pb.arena.alloc(&pb.d_tod, (size_t) m));
        MGPU_RET(pb.arena.alloc(&pb.d_tod2, (size_t) m));
    }
    if (!pb.d_map[0])
        for (int i = 0; i < 8; i++) MGPU_RET(pb.arena.alloc(&pb.d_map[i], (size_t) n0));
    // partials: 3 per map workgroup (at most one extra workgroup per segment)
    const int64_t need_red = 3 * ((int64_t) grid_for(pb.pt.npix0, kMB) + pb.sg.nseg + 1) + 64;
    if (pb.d_red && pb.nred < need_red) {
        pb.arena.release(pb.d_red);
        pb.d_red = nullptr;
    }
    if (!pb.d_red) {
        pb.nred = need_red;
        MGPU_RET(pb.arena.alloc(&pb.d_red, (size_t) pb.nred));
    }
    const int64_t nscal = (int64_t) S_NSCAL * pb.sg.nseg;
    if (pb.d_scal && pb.nscal < nscal) {
        pb.arena.release(pb.d_scal);
        pb.d_scal = nullptr;
    }
    if (!pb.d_scal) {
        pb.nscal = nscal;
        MGPU_RET(pb.arena.alloc(&pb.d_scal, (size_t) nscal));
    }
    if (!pb.d_flag) MGPU_RET(pb.arena.alloc(&pb.d_flag, (size_t) F_NFLAGS));
    if (!pb.ev0) MGPU_HIP(hipEventCreate(&pb.ev0));
    if (!pb.ev1) MGPU_HIP(hipEventCreate(&pb.ev1));
    if (pb.opt.precision == MGPU_MIXED && !pb.d_todf) {
        MGPU_RET(pb.arena.alloc(&pb.d_todf, (size_t) m));
        MGPU_RET(pb.arena.alloc(&pb.d_tod2f, (size_t) m));
        MGPU_RET(pb.arena.alloc(&pb.d_mapf, (size_t) n0));
    }
    return MGPU_OK;
}

// A v = P^T N^-1 P v in FP64 (map -> map), fused <v, Av> partials optional
int apply_A64(Problem &pb, const double *d_v, double *d_out, double *d_dot) {
    pb.prof.begin(MGPU_KT_P, pb.stream);
    MGPU_RET(launch_matvec(pb, d_v, pb.d_tod, true));
    pb.prof.end(pb.stream);
    if (pb.tm.live)  // apply_weights and the template part of PCG_GLS_templates (pcg_true.c:426-458)
        MGPU_RET(templates_noise(pb, pb.d_tod, pb.d_tod2));
    else
        MGPU_RET(toeplitz_apply(pb, pb.d_tod, pb.d_tod2, false));
    pb.prof.begin(MGPU_KT_PT, pb.stream);
    MGPU_RET(launch_trmatvec(pb, pb.d_tod2, d_out, d_dot ? d_v : nullptr, d_dot, nullptr));
    pb.prof.end(pb.stream);
    return MGPU_OK;
}

static int apply_A32(Problem &pb, const double *d_v, double *d_out, double *d_dot, const int *d_flag) {
    const int n = 3 * pb.pt.npix;
    pb.prof.begin(MGPU_KT_P, pb.stream);
    k_map_to_f32<<<grid_for(n, 256), 256, 0, pb.stream>>>(d_v, pb.d_mapf, n, d_flag);
    MGPU_LAUNCHED();
    MGPU_RET(launch_matvec_f32(pb, pb.d_mapf, pb.d_todf, true));
    pb.prof.end(pb.stream);
    MGPU_RET(toeplitz_apply_f32(pb, pb.d_todf, pb.d_tod2f, false));
    pb.prof.begin(MGPU_KT_PT, pb.stream);
    MGPU_RET(launch_trmatvec_f32(pb, pb.d_tod2f, d_out, d_dot ? d_v : nullptr, d_dot, nullptr));
    pb.prof.end(pb.stream);
    return MGPU_OK;
}

// owns a captured graph of one block of iterations
struct GraphBlock {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~GraphBlock() {
        if (exec) (void) hipGraphExecDestroy(exec);
        if (graph) (void) hipGraphDestroy(graph);
    }
};

int pcg_run(Problem &pb, const double *d_signal, const double *d_noise, double tol, int maxiter, int stop_after,
            double *h_x, double *h_hist, mgpu_report *rep, std::vector<SegReport> *segs,
            const std::vector<double> *seg_tol, const std::vector<int> *seg_maxiter, const SegDone *seg_done) {
    MGPU_ARG(pb.pc.built, "pcg before precond_build");
    MGPU_ARG(maxiter >= 1, "maxiter=%d must be >= 1", maxiter);
    MGPU_RET(segments_ensure(pb));
    MGPU_RET(ensure_work(pb));
    const bool mixed = pb.opt.precision == MGPU_MIXED;
    MGPU_ARG(!(mixed && pb.tm.on), "templates are solved in FP64 only");
    if (mixed) MGPU_RET(pointing_make_fp32(pb));
    // two-level preconditioner: Z, A Z and pinv(E) (build_precond, precond.c:1342-1398), timed apart
    const bool two = pb.opt.precond != 0;
    if (two) {
        const bool keep = pb.prof.on;
        pb.prof.on = false;
        const int rc = twolvl_build(pb, d_signal, d_noise);
        pb.prof.on = keep;
        MGPU_RET(rc);
    }
    hipStream_t s = pb.stream;
    const Segments &G = pb.sg;
    const int nseg = G.nseg, nwg = G.nwg;
    const int npix = pb.pt.npix, n = 3 * npix;
    const int64_t m = pb.pt.m;
    MGPU_ARG(nwg > 0, "no pixel left to solve for");
    double *x = pb.d_map[0], *h = pb.d_map[1], *gA = pb.d_map[2], *gB = pb.d_map[3], *qv = pb.d_map[4],
           *Cg = pb.d_map[5], *bvec = pb.d_map[6], *Ax = pb.d_map[7];
    double *part = pb.d_red;
    double *S = pb.d_scal;
    int *flag = pb.d_flag;
    int *state = G.d_state;
    const int *wseg = G.d_wg_seg, *wq0 = G.d_wg_q0, *sq = G.d_q, *wg0 = G.d_wg0;
    double *d_hist = nullptr;
    if (h_hist) {
        MGPU_RET(pb.arena.alloc(&d_hist, (size_t) maxiter * nseg));
        MGPU_HIP(hipMemsetAsync(d_hist, 0, sizeof(double) * (size_t) maxiter * nseg, s));
    }

    // scalars of every segment: tol, stop_after, maxiter
    MGPU_ARG(!seg_tol || (int) seg_tol->size() == nseg, "%zu tolerances for %d segments", seg_tol->size(), nseg);
    MGPU_ARG(!seg_maxiter || (int) seg_maxiter->size() == nseg, "%zu iteration caps for %d segments",
             seg_maxiter->size(), nseg);
    std::vector<double> init((size_t) S_NSCAL * nseg, 0.0);
    for (int r = 0; r < nseg; r++) {
        const int mr = seg_maxiter ? (*seg_maxiter)[r] : maxiter;
        MGPU_ARG(mr >= 1 && mr <= maxiter, "segment %d: maxiter %d outside [1, %d]", r, mr, maxiter);
        init[(size_t) r * S_NSCAL + S_TOL] = seg_tol ? (*seg_tol)[r] : tol;
        init[(size_t) r * S_NSCAL + S_STOPK] = stop_after > 0 ? (double) stop_after : 0.0;
        init[(size_t) r * S_NSCAL + S_MAXK] = (double) mr;
    }
    MGPU_HIP(hipMemcpyAsync(S, init.data(), sizeof(double) * init.size(), hipMemcpyHostToDevice, s));
    MGPU_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    MGPU_HIP(hipMemsetAsync(flag, 0, sizeof(int) * F_NFLAGS, s));
    MGPU_HIP(hipEventRecord(pb.ev0, s));
    std::unique_ptr<TimerRange> phase(new TimerRange("PCG initial residual"));  // pcg_true.c:75-98

    // g0 = P^T N^-1 (s + n) (FP64 in both modes: it is b, kept for refreshes)
    // the template part of N^-1 (PCG_GLS_templates) from here on; the two-level
    // space above is built on the plain N^-1, as build_precond does in MTmap
    struct TemplatesLive {
        Problem &pb;
        ~TemplatesLive() { pb.tm.live = false; }
    } tm_live{pb};
    pb.tm.live = pb.tm.on;
    k_rhs_tod<<<grid_for(m, 256), 256, 0, s>>>(d_signal, d_noise, pb.d_tod, m);
    MGPU_LAUNCHED();
    if (pb.tm.live)  // pcg_true.c:318-371
        MGPU_RET(templates_noise(pb, pb.d_tod, pb.d_tod2));
    else
        MGPU_RET(toeplitz_apply(pb, pb.d_tod, pb.d_tod2, true));
    MGPU_RET(launch_trmatvec(pb, pb.d_tod2, gA, nullptr, nullptr, nullptr));
    MGPU_HIP(hipMemcpyAsync(bvec, gA, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (two) {
        MGPU_RET(twolvl_apply(pb, gA, nullptr, Cg, h, part, true, nullptr, nullptr));
    } else {
        k_init_dots<<<nwg, kMB, 0, s>>>(pb.pc.d_inv, gA, Cg, h, wseg, wq0, sq, part);
        MGPU_LAUNCHED();
    }
    k_init_scalars<<<nseg, kRB, 0, s>>>(part, nwg, wg0, S, d_hist, maxiter, state, nseg, flag);
    MGPU_LAUNCHED();

    // from here on P, N^-1 and P^T skip the work of stopped segments (every
    // segment is computed again after the loop, for the true residual)
    struct LiveState {
        Problem &pb;
        ~LiveState() {
            pb.live_state = nullptr;
            run_lists_clear(pb);
        }
    } live{pb};
    pb.live_state = state;
    run_lists_clear(pb);

    // solves of the request loop keep the full work lists and full refreshes
    const bool serving = serving_thread();
    const int poll = pb.opt.poll_every > 0 ? pb.opt.poll_every : 8;
    // refresh period of mixed precision; adaptive: no period, segments ask (S_PEND)
    const bool adaptive = mixed && pb.opt.refresh_every == MGPU_REFRESH_ADAPTIVE;
    const int refresh = !mixed || adaptive ? 0 : (pb.opt.refresh_every > 0 ? pb.opt.refresh_every : 10);
    // Iterations are queued in blocks between host polls.  A graph block must
    // repeat exactly: an even length (g/gp swap parity) and, in mixed
    // precision, a multiple of the refresh period; eager blocks have the same length.
    // The block length is the same with and without graphs (polls, and so the
    // iterations that refresh pending segments, fall on the same numbers)
    const bool use_graph = pb.opt.graph != 0 && !pb.prof.on;
    const int period = refresh ? (refresh % 2 ? 2 * refresh : refresh) : 2;
    const int block = std::max(1, (poll + period - 1) / period) * period;
    int launched = 0;  // iterations queued (k = 1 .. launched)
    double *g = gA, *gp = gB;
    GraphBlock gb;

    // refresh_pending: the segments whose FP32-updated residual met the stop test
    // (S_PEND) refresh their residual in FP64 in this iteration whatever its
    // number (every live segment does when the number is a multiple of refresh)
    auto enqueue_iteration = [&](bool refresh_pending) -> int {
        ++launched;
        pb.prof.tag = launched;
        const bool all = refresh && launched % refresh == 0;
        const bool do_refresh = all || (mixed && refresh_pending);
        const int kind = !mixed ? IT_FP64 : all ? IT_REFRESH_ALL : do_refresh ? IT_REFRESH_PENDING : IT_MIXED;
        std::swap(g, gp);  // pcg_true.c:140-142
        // q = A h with fused <h, q>
        if (mixed)
            MGPU_RET(apply_A32(pb, h, qv, part, flag));
        else
            MGPU_RET(apply_A64(pb, h, qv, part));
        pb.prof.begin(MGPU_KT_MAP, s);
        k_coeff<<<nseg, kRB, 0, s>>>(part, wg0, S, state, flag);
        MGPU_LAUNCHED();
        if (two) {
            k_update_xg<<<nwg, kMB, 0, s>>>(x, h, g, gp, qv, wseg, wq0, sq, S, state, flag);
            MGPU_LAUNCHED();
            MGPU_RET(twolvl_apply(pb, g, gp, Cg, nullptr, part, false, state, flag));
        } else {
            k_update<<<nwg, kMB, 0, s>>>(x, h, g, gp, qv, pb.pc.d_inv, Cg, wseg, wq0, sq, S, state, part, flag);
            MGPU_LAUNCHED();
        }
        pb.prof.end(s);
        if (do_refresh) {
            // g = b - A x in FP64, then Cg = M g and the three dots on the refreshed residual
            pb.prof.begin(MGPU_KT_REFRESH, s);
            const bool keep = pb.prof.on;
            pb.prof.on = false;  // the FP64 operator inside is accounted as a refresh
            if (kind == IT_REFRESH_PENDING && nseg > 1 && !serving) {
                // only the pending segments need A x: the others' chunks and tiles return at once
                k_refresh_mask<<<grid_for(nseg, 64), 64, 0, s>>>(state, S, nseg, G.d_rstate);
                MGPU_LAUNCHED();
                pb.live_state = G.d_rstate;
                const int rc = apply_A64(pb, x, Ax, nullptr);
                pb.live_state = state;
                MGPU_RET(rc);
            } else {
                MGPU_RET(apply_A64(pb, x, Ax, nullptr));
            }
            k_refresh<<<nwg, kMB, 0, s>>>(bvec, Ax, g, wseg, wq0, sq, state, flag, flag + F_NREF, S, kind);
            MGPU_LAUNCHED();
            if (two) {
                // segments not refreshed get the same Cg and partials again (same g, gp)
                MGPU_RET(twolvl_apply(pb, g, gp, Cg, nullptr, part, false, state, flag));
            } else {
                k_refresh_three<<<nwg, kMB, 0, s>>>(pb.pc.d_inv, g, gp, Cg, wseg, wq0, sq, state, part, flag, S,
                                                     kind);
                MGPU_LAUNCHED();
            }
            pb.prof.on = keep;
            pb.prof.end(s);
        }
        pb.prof.begin(MGPU_KT_MAP, s);
        k_scalars<<<nseg, kRB, 0, s>>>(part, nwg, wg0, S, d_hist, maxiter, state, nseg, flag, kind, adaptive ? 1 : 0);
        MGPU_LAUNCHED();
        k_hupdate<<<nwg, kMB, 0, s>>>(h, Cg, wseg, wq0, sq, S, state, flag);
        MGPU_LAUNCHED();
        pb.prof.end(s);
        return MGPU_OK;
    };

    phase.reset();  // ranges nest: the initial-residual range ends before the solve range starts
    phase.reset(new TimerRange("PCG solve time"));  // pcg_true.c:132-204
    auto seg_iterations = [&](int r, const std::vector<double> &hs, const std::vector<int> &hstate) {
        // reference: k at break, or K when the loop ran out (pcg_true.c:224)
        const int kr = (int) hs[(size_t) r * S_NSCAL + S_K];
        return hstate[r] == 1 || (stop_after > 0 && kr >= stop_after) ? kr
                                                                       : (seg_maxiter ? (*seg_maxiter)[r] : maxiter);
    };
    auto seg_report = [&](int r, const std::vector<double> &hs, const std::vector<int> &hstate) {
        SegReport o;
        o.iterations = seg_iterations(r, hs, hstate);
        o.converged = hstate[r] == 1;
        o.res0 = hs[(size_t) r * S_NSCAL + S_RES0];
        o.res = hs[(size_t) r * S_NSCAL + S_RES];
        o.q0 = G.h_q[r];
        o.q1 = G.h_q[r + 1];
        return o;
    };
    // segments already handed to seg_done
    std::vector<char> reported((size_t) nseg, 0);
    auto report_stopped = [&]() -> int {
        std::vector<double> hs((size_t) S_NSCAL * nseg);
        std::vector<int> hstate((size_t) nseg);
        MGPU_HIP(hipMemcpyAsync(hs.data(), S, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, s));
        MGPU_HIP(hipMemcpyAsync(hstate.data(), state, sizeof(int) * nseg, hipMemcpyDeviceToHost, s));
        MGPU_HIP(hipStreamSynchronize(s));
        for (int r = 0; r < nseg; r++)
            if (hstate[r] && !reported[r]) {
                reported[r] = 1;
                MGPU_RET((*seg_done)(r, seg_report(r, hs, hstate), x));
            }
        return MGPU_OK;
    };

    bool done = maxiter <= 1;  // maxiter == 1: no loop iteration (for k = 1; k < K)
    bool pending = false;      // mixed precision: some segment awaits its FP64 refresh
    int ndone_seen = 0;        // stopped segments at the last report
    int ndone_lists = 0;       // stopped segments when the work lists were last rebuilt
    while (!done) {
        if (pending) {
            // an eager block whose first iteration refreshes the pending segments;
            // blocks keep their length, so polls (and refreshes) fall on the same
            // iteration numbers whatever the other segments of a batch do
            for (int it = 0; it < block && launched + 1 < maxiter; it++) MGPU_RET(enqueue_iteration(it == 0));
        } else if (gb.exec) {
            MGPU_HIP(hipGraphLaunch(gb.exec, s));
            launched += block;
        } else if (use_graph && launched > 0) {
            // the first block ran eagerly (kernel attributes); capture the second
            MGPU_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = MGPU_OK;
            for (int it = 0; it < block && rc == MGPU_OK; it++) rc = enqueue_iteration(false);
            hipGraph_t graph = nullptr;
            const hipError_t ec = hipStreamEndCapture(s, &graph);
            gb.graph = graph;
            MGPU_RET(rc);
            MGPU_HIP(ec);
            MGPU_HIP(hipGraphInstantiate(&gb.exec, gb.graph, nullptr, nullptr, 0));
            MGPU_HIP(hipGraphLaunch(gb.exec, s));
        } else {
            for (int it = 0; it < block && launched + 1 < maxiter; it++) MGPU_RET(enqueue_iteration(false));
        }
        int hf[F_NFLAGS];
        MGPU_HIP(hipMemcpyAsync(hf, flag, sizeof(hf), hipMemcpyDeviceToHost, s));
        MGPU_HIP(hipStreamSynchronize(s));
        pb.prof.collect(hf[F_KMAX]);
        pending = mixed && hf[F_NPEND] > 0 && !hf[F_STOP];
        if (hf[F_STOP] || launched + 1 >= maxiter) done = true;
        if (seg_done && nseg > 1 && hf[F_NDONE] > ndone_seen && !done) {
            ndone_seen = hf[F_NDONE];
            MGPU_RET(report_stopped());
        }
        if (nseg > 1 && !serving && hf[F_NDONE] > ndone_lists && !done) {
            // segments stopped: later blocks launch only the running segments'
            // chunks and tiles (a captured block is captured again with them)
            ndone_lists = hf[F_NDONE];
            std::vector<int> hstate((size_t) nseg);
            MGPU_HIP(hipMemcpyAsync(hstate.data(), state, sizeof(int) * nseg, hipMemcpyDeviceToHost, s));
            MGPU_HIP(hipStreamSynchronize(s));
            MGPU_RET(run_lists_update(pb, hstate));
            if (gb.exec) {
                (void) hipGraphExecDestroy(gb.exec);
                gb.exec = nullptr;
            }
            if (gb.graph) {
                (void) hipGraphDestroy(gb.graph);
                gb.graph = nullptr;
            }
        }
    }
    pb.prof.tag = 0;
    pb.live_state = nullptr;
    run_lists_clear(pb);
    phase.reset();
    MGPU_HIP(hipEventRecord(pb.ev1, s));
    // results
    int hf[F_NFLAGS] = {0};
    std::vector<double> hs((size_t) S_NSCAL * nseg);
    std::vector<int> hstate((size_t) nseg);
    MGPU_HIP(hipMemcpyAsync(hf, flag, sizeof(hf), hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipMemcpyAsync(hs.data(), S, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipMemcpyAsync(hstate.data(), state, sizeof(int) * nseg, hipMemcpyDeviceToHost, s));
    if (h_x) MGPU_HIP(hipMemcpyAsync(h_x, x, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipStreamSynchronize(s));
    if (h_hist) MGPU_HIP(hipMemcpy(h_hist, d_hist, sizeof(double) * (size_t) maxiter * nseg, hipMemcpyDeviceToHost));
    float ms = 0.f;
    MGPU_HIP(hipEventElapsedTime(&ms, pb.ev0, pb.ev1));
    const int refreshes = hf[F_NREF];

    std::vector<double> true_rel((size_t) nseg, -1.0);
    if (mixed && (rep || segs)) {
        // true residual of the returned x in FP64: ||b - A x|| / ||b|| (not timed)
        const bool keep = pb.prof.on;
        pb.prof.on = false;
        const int rc = apply_A64(pb, x, Ax, nullptr);
        pb.prof.on = keep;
        MGPU_RET(rc);
        k_refresh<<<nwg, kMB, 0, s>>>(bvec, Ax, qv, wseg, wq0, sq, nullptr, pb.d_flag + F_NEVER, nullptr, nullptr,
                                      IT_REFRESH_ALL);
        MGPU_LAUNCHED();
        k_two_norms<<<nwg, kMB, 0, s>>>(qv, bvec, wseg, wq0, sq, part);
        MGPU_LAUNCHED();
        k_seg_norms<<<nseg, kRB, 0, s>>>(part, nwg, wg0, S);
        MGPU_LAUNCHED();
        MGPU_HIP(hipMemcpyAsync(hs.data(), S, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, s));
        MGPU_HIP(hipStreamSynchronize(s));
        for (int r = 0; r < nseg; r++)
            true_rel[r] = std::sqrt(hs[(size_t) r * S_NSCAL + S_TRUE_R2]) /
                          std::sqrt(hs[(size_t) r * S_NSCAL + S_TRUE_B2]);
    }
    if (rep) {
        rep->converged = hstate[0] == 1;
        rep->iterations = seg_iterations(0, hs, hstate);
        rep->res0 = hs[S_RES0];
        rep->res = hs[S_RES];
        rep->n_map = n;
        rep->trash_pix = pb.pt.trash;
        rep->lcount = 3 * npix + 3 * pb.pt.trash;
        rep->n_cut = pb.pc.n_cut;
        rep->t_solve_ms = ms;
        rep->t_setup_ms = pb.t_create_ms + pb.t_precond_ms;
        rep->refreshes = refreshes;
        rep->tile_variant = pb.tp.variant(mixed ? 1 : 0);
        rep->true_res_rel = true_rel[0];
        rep->t_2lvl_ms = two ? pb.tl.build_ms : 0.0;
        rep->dev_peak_bytes = (long long) pb.arena.peak;
    }
    if (segs) {
        segs->assign((size_t) nseg, SegReport{});
        for (int r = 0; r < nseg; r++) {
            SegReport &o = (*segs)[r];
            o = seg_report(r, hs, hstate);
            o.true_res_rel = true_rel[r];
        }
    }
    // segments that stopped at the last poll (or never stopped)
    if (seg_done)
        for (int r = 0; r < nseg; r++)
            if (!reported[r]) {
                reported[r] = 1;
                SegReport o = seg_report(r, hs, hstate);
                o.true_res_rel = true_rel[r];
                MGPU_RET((*seg_done)(r, o, x));
            }
    if (d_hist) pb.arena.release(d_hist);
    return MGPU_OK;
}

}  // namespace mgpu
