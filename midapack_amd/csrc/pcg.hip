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
// set, so the host queues `poll_every` iterations at a time and reads the flag
// once per batch.  Dot products are fused into the producing kernels as
// per-workgroup partials and reduced by a single workgroup in a fixed order:
// results are deterministic and independent of what else runs on the GPU.
//
// MGPU_MIXED: P, N^-1 and P^T run in FP32 on FP32 copies of the pointing
// weights and TOD; map vectors, dot products and the 3x3 blocks stay FP64,
// and every `refresh_every` iterations g is recomputed in FP64 as b - A x.
#include "mgpu_internal.h"

#include <cmath>
#include <cstring>

namespace mgpu {

// scalar slots in d_scal
enum {
    S_G2PIX = 0,
    S_RES,
    S_RES0,
    S_TOL2,
    S_RO,
    S_GAMMA,
    S_COEFF,
    S_POLAK,
    S_G2PIXP,
    S_K,        // iteration counter (as double)
    S_TOL,
    S_STOPK,    // stop_after (0 = off)
    S_MAXK,
    S_NSCAL
};

constexpr int kMB = 256;   // map-domain workgroup size
constexpr int kRedMax = 1 << 16;

__device__ __forceinline__ bool stopped(const int *flag) { return *(volatile const int *) flag != 0; }

__device__ __forceinline__ double block_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int) (blockDim.x >> 6); i++) s += red[i];
    return s;
}

// fixed-order reduction of nparts partials (single workgroup of 256)
__device__ __forceinline__ double reduce_parts(const double *__restrict__ p, int nparts, double *red) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) s += p[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    return red[0];
}

// tod = s + n (the first residual with x0 = 0, pcg_true.c:78-80), fused partials
__global__ __launch_bounds__(256) void k_rhs_tod(const double *__restrict__ s, const double *__restrict__ n,
                                                 double *__restrict__ tod, int64_t m) {
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) tod[t] = (s[t] + n[t]) - 0.0;
}

__global__ __launch_bounds__(256) void k_to_f32_tod(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float) a[i];
}

__global__ __launch_bounds__(256) void k_map_to_f32(const double *__restrict__ a, float *__restrict__ b, int n,
                                                    const int *__restrict__ flag) {
    if (flag && stopped(flag)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float) a[i];
}

// init: Cg = M g, h = Cg, partials of <Cg,g> and <g,g>
__global__ __launch_bounds__(kMB) void k_init_dots(const double *__restrict__ inv, const double *__restrict__ g,
                                                   double *__restrict__ Cg, double *__restrict__ h, int npix,
                                                   double *__restrict__ part) {
    __shared__ double red[kMB / 64];
    const int q = blockIdx.x * kMB + threadIdx.x;
    double a = 0.0, c = 0.0;
    if (q < npix) {
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

__global__ __launch_bounds__(256) void k_init_scalars(const double *__restrict__ part, int nparts,
                                                      double *__restrict__ S, double *__restrict__ hist,
                                                      int *__restrict__ flag) {
    __shared__ double red[256];
    const double g2pix = reduce_parts(part, nparts, red);
    __syncthreads();
    const double res = reduce_parts(part + nparts, nparts, red);
    if (threadIdx.x == 0) {
        S[S_G2PIX] = g2pix;
        S[S_RES] = res;
        S[S_RES0] = res;
        S[S_TOL2] = S[S_TOL] * S[S_TOL] * res;
        S[S_K] = 0.0;
        if (hist) hist[0] = sqrt(res) / sqrt(res);
        flag[0] = 0;
        flag[1] = 0;  // converged
        flag[2] = 0;  // iterations done
    }
}

// ro = g2pix / coeff
__global__ __launch_bounds__(256) void k_coeff(const double *__restrict__ part, int nparts, double *__restrict__ S,
                                               const int *__restrict__ flag) {
    if (stopped(flag)) return;
    __shared__ double red[256];
    const double coeff = reduce_parts(part, nparts, red);
    if (threadIdx.x == 0) {
        S[S_COEFF] = coeff;
        S[S_RO] = S[S_G2PIX] / coeff;
    }
}

// x += ro h; g = gp - ro q; Cg = M g; partials <Cg,g>, <Cg,g-gp>, <g,g>
__global__ __launch_bounds__(kMB) void k_update(double *__restrict__ x, const double *__restrict__ h,
                                                double *__restrict__ g, const double *__restrict__ gp,
                                                const double *__restrict__ qv, const double *__restrict__ inv,
                                                double *__restrict__ Cg, int npix, const double *__restrict__ S,
                                                double *__restrict__ part, const int *__restrict__ flag) {
    if (stopped(flag)) return;
    __shared__ double red[kMB / 64];
    const int q = blockIdx.x * kMB + threadIdx.x;
    double a = 0.0, b = 0.0, c = 0.0;
    if (q < npix) {
        const double ro = S[S_RO];
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

// dots -> g2pix, polak, res; history; stop test; gamma
__global__ __launch_bounds__(256) void k_scalars(const double *__restrict__ part, int nparts, double *__restrict__ S,
                                                 double *__restrict__ hist, int *__restrict__ flag) {
    if (stopped(flag)) return;
    __shared__ double red[256];
    const double g2pix = reduce_parts(part, nparts, red);
    __syncthreads();
    const double polak = reduce_parts(part + nparts, nparts, red);
    __syncthreads();
    const double res = reduce_parts(part + 2 * nparts, nparts, red);
    if (threadIdx.x == 0) {
        const double k = S[S_K] + 1.0;
        S[S_K] = k;
        S[S_G2PIXP] = S[S_G2PIX];
        S[S_G2PIX] = g2pix;
        S[S_POLAK] = polak;
        S[S_RES] = res;
        const int ki = (int) k;
        if (hist && ki < (int) S[S_MAXK]) hist[ki] = sqrt(res) / sqrt(S[S_RES0]);
        flag[2] = ki;
        if (res <= S[S_TOL2]) {
            flag[1] = 1;
            flag[0] = 1;
        } else if ((S[S_STOPK] > 0.0 && k >= S[S_STOPK]) || k + 1.0 >= S[S_MAXK]) {
            flag[0] = 1;
        }
        S[S_GAMMA] = polak / S[S_G2PIXP];
    }
}

// h = h gamma + Cg
__global__ __launch_bounds__(kMB) void k_hupdate(double *__restrict__ h, const double *__restrict__ Cg, int n,
                                                 const double *__restrict__ S, const int *__restrict__ flag) {
    if (stopped(flag)) return;
    const int i = blockIdx.x * kMB + threadIdx.x;
    if (i < n) h[i] = h[i] * S[S_GAMMA] + Cg[i];
}

// FP64 refresh of the residual: g = b - Ax (mixed precision)
__global__ __launch_bounds__(kMB) void k_refresh(const double *__restrict__ b, const double *__restrict__ Ax,
                                                 double *__restrict__ g, int n, const int *__restrict__ flag) {
    if (stopped(flag)) return;
    const int i = blockIdx.x * kMB + threadIdx.x;
    if (i < n) g[i] = b[i] - Ax[i];
}

__global__ __launch_bounds__(256) void k_dot_plain(const double *__restrict__ a, const double *__restrict__ b, int n,
                                                   double *__restrict__ part) {
    __shared__ double red[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double v = i < n ? a[i] * b[i] : 0.0;
    v = block_sum(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ __launch_bounds__(256) void k_sum_parts(const double *__restrict__ part, int nparts,
                                                   double *__restrict__ out) {
    __shared__ double red[256];
    const double s = reduce_parts(part, nparts, red);
    if (threadIdx.x == 0) *out = s;
}

// refreshed residual: Cg = M g and the three dots <Cg,g>, <Cg,g-gp>, <g,g>
__global__ __launch_bounds__(kMB) void k_refresh_three(const double *__restrict__ inv, const double *__restrict__ g,
                                                       const double *__restrict__ gp, double *__restrict__ Cg,
                                                       int npix, double *__restrict__ part,
                                                       const int *__restrict__ flag) {
    if (stopped(flag)) return;
    __shared__ double red[kMB / 64];
    const int q = blockIdx.x * kMB + threadIdx.x;
    double a = 0.0, b = 0.0, c = 0.0;
    if (q < npix) {
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

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
int ensure_work(Problem &pb) {
    const int64_t m = pb.pt.m;
    const int n0 = 3 * pb.pt.npix0;
    if (!pb.d_tod) {
        MGPU_RET(pb.arena.alloc(&pb.d_tod, (size_t) m));
        MGPU_RET(pb.arena.alloc(&pb.d_tod2, (size_t) m));
    }
    if (!pb.d_map[0])
        for (int i = 0; i < 8; i++) MGPU_RET(pb.arena.alloc(&pb.d_map[i], (size_t) n0));
    if (!pb.d_red) {
        pb.nred = 3 * ((int64_t) grid_for(pb.pt.npix0, kMB) + 1) + 4 * (int64_t) grid_for(n0, 256) + 64;
        MGPU_RET(pb.arena.alloc(&pb.d_red, (size_t) pb.nred));
        MGPU_RET(pb.arena.alloc(&pb.d_scal, (size_t) 32));
        MGPU_RET(pb.arena.alloc(&pb.d_flag, (size_t) 8));
        MGPU_HIP(hipHostMalloc((void **) &pb.h_pinned, 64 * sizeof(double), hipHostMallocDefault));
        MGPU_HIP(hipEventCreate(&pb.ev0));
        MGPU_HIP(hipEventCreate(&pb.ev1));
    }
    if (pb.opt.precision == MGPU_MIXED && !pb.d_todf) {
        MGPU_RET(pb.arena.alloc(&pb.d_todf, (size_t) m));
        MGPU_RET(pb.arena.alloc(&pb.d_tod2f, (size_t) m));
        MGPU_RET(pb.arena.alloc(&pb.d_mapf, (size_t) n0));
    }
    return MGPU_OK;
}

// A v = P^T N^-1 P v in FP64 (map -> map), fused <v, Av> partials optional
static int apply_A64(Problem &pb, const double *d_v, double *d_out, double *d_dot) {
    pb.prof.begin(MGPU_KT_P, pb.stream);
    MGPU_RET(launch_matvec(pb, d_v, pb.d_tod, true));
    pb.prof.end(pb.stream);
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

int pcg_run(Problem &pb, const double *d_signal, const double *d_noise, double tol, int maxiter, int stop_after,
            double *h_x, double *h_hist, mgpu_report *rep) {
    MGPU_ARG(pb.pc.built, "pcg before precond_build");
    MGPU_ARG(maxiter >= 1, "maxiter=%d must be >= 1", maxiter);
    MGPU_RET(ensure_work(pb));
    const bool mixed = pb.opt.precision == MGPU_MIXED;
    if (mixed) MGPU_RET(pointing_make_fp32(pb));
    hipStream_t s = pb.stream;
    const int npix = pb.pt.npix, n = 3 * npix;
    const int64_t m = pb.pt.m;
    const unsigned gmap = grid_for(npix, kMB);
    double *x = pb.d_map[0], *h = pb.d_map[1], *gA = pb.d_map[2], *gB = pb.d_map[3], *qv = pb.d_map[4],
           *Cg = pb.d_map[5], *bvec = pb.d_map[6], *Ax = pb.d_map[7];
    double *part = pb.d_red;
    double *S = pb.d_scal;
    int *flag = pb.d_flag;
    double *d_hist = nullptr;
    if (h_hist) MGPU_RET(pb.arena.alloc(&d_hist, (size_t) maxiter));

    // scalars: tol, stop_after, maxiter
    double init[S_NSCAL] = {0};
    init[S_TOL] = tol;
    init[S_STOPK] = stop_after > 0 ? (double) stop_after : 0.0;
    init[S_MAXK] = (double) maxiter;
    MGPU_HIP(hipMemcpyAsync(S, init, sizeof(init), hipMemcpyHostToDevice, s));
    MGPU_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    MGPU_HIP(hipMemsetAsync(flag, 0, sizeof(int) * 8, s));
    MGPU_HIP(hipEventRecord(pb.ev0, s));

    // g0 = P^T N^-1 (s + n) (FP64 in both modes: it is b, kept for refreshes)
    k_rhs_tod<<<grid_for(m, 256), 256, 0, s>>>(d_signal, d_noise, pb.d_tod, m);
    MGPU_LAUNCHED();
    MGPU_RET(toeplitz_apply(pb, pb.d_tod, pb.d_tod2, true));
    MGPU_RET(launch_trmatvec(pb, pb.d_tod2, gA, nullptr, nullptr, nullptr));
    MGPU_HIP(hipMemcpyAsync(bvec, gA, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    k_init_dots<<<gmap, kMB, 0, s>>>(pb.pc.d_inv, gA, Cg, h, npix, part);
    MGPU_LAUNCHED();
    k_init_scalars<<<1, 256, 0, s>>>(part, (int) gmap, S, d_hist, flag);
    MGPU_LAUNCHED();

    const int poll = pb.opt.poll_every > 0 ? pb.opt.poll_every : 8;
    const int refresh = mixed ? (pb.opt.refresh_every > 0 ? pb.opt.refresh_every : 10) : 0;
    int launched = 0;  // iterations queued (k = 1 .. launched)
    int refreshes = 0;
    double *g = gA, *gp = gB;
    int nparts_pt = 0;
    bool done = false;
    // maxiter == 1: no loop iteration (for k = 1; k < K)
    if (maxiter <= 1) done = true;
    while (!done) {
        for (int it = 0; it < poll && launched + 1 < maxiter; it++) {
            ++launched;
            pb.prof.tag = launched;
            std::swap(g, gp);  // pcg_true.c:140-142
            // q = A h with fused <h, q>
            if (mixed)
                MGPU_RET(apply_A32(pb, h, qv, part, flag));
            else
                MGPU_RET(apply_A64(pb, h, qv, part));
            nparts_pt = trmatvec_nparts(pb);
            pb.prof.begin(MGPU_KT_MAP, s);
            k_coeff<<<1, 256, 0, s>>>(part, nparts_pt, S, flag);
            MGPU_LAUNCHED();
            k_update<<<gmap, kMB, 0, s>>>(x, h, g, gp, qv, pb.pc.d_inv, Cg, npix, S, part, flag);
            MGPU_LAUNCHED();
            pb.prof.end(s);
            if (refresh && launched % refresh == 0) {
                // g = b - A x in FP64, then Cg = M g and g2pix on the refreshed residual
                pb.prof.begin(MGPU_KT_REFRESH, s);
                const bool keep = pb.prof.on;
                pb.prof.on = false;  // the FP64 operator inside is accounted as a refresh
                MGPU_RET(apply_A64(pb, x, Ax, nullptr));
                k_refresh<<<grid_for(n, kMB), kMB, 0, s>>>(bvec, Ax, g, n, flag);
                MGPU_LAUNCHED();
                // the three dots of the stop test and of gamma on the refreshed g
                k_refresh_three<<<gmap, kMB, 0, s>>>(pb.pc.d_inv, g, gp, Cg, npix, part, flag);
                MGPU_LAUNCHED();
                pb.prof.on = keep;
                pb.prof.end(s);
                refreshes++;
            }
            pb.prof.begin(MGPU_KT_MAP, s);
            k_scalars<<<1, 256, 0, s>>>(part, (int) gmap, S, d_hist, flag);
            MGPU_LAUNCHED();
            k_hupdate<<<grid_for(n, kMB), kMB, 0, s>>>(h, Cg, n, S, flag);
            MGPU_LAUNCHED();
            pb.prof.end(s);
        }
        int hf[4];
        MGPU_HIP(hipMemcpyAsync(hf, flag, sizeof(int) * 3, hipMemcpyDeviceToHost, s));
        MGPU_HIP(hipStreamSynchronize(s));
        pb.prof.collect(hf[2]);
        if (hf[0] || launched + 1 >= maxiter) done = true;
    }
    pb.prof.tag = 0;
    MGPU_HIP(hipEventRecord(pb.ev1, s));
    // results
    int hf[4] = {0, 0, 0, 0};
    double hs[S_NSCAL];
    MGPU_HIP(hipMemcpyAsync(hf, flag, sizeof(int) * 3, hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipMemcpyAsync(hs, S, sizeof(hs), hipMemcpyDeviceToHost, s));
    if (h_x) MGPU_HIP(hipMemcpyAsync(h_x, x, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipStreamSynchronize(s));
    const int kdone = hf[2];
    if (h_hist) {
        const int nh = std::min(maxiter, kdone + 1);
        MGPU_HIP(hipMemcpy(h_hist, d_hist, sizeof(double) * nh, hipMemcpyDeviceToHost));
    }
    float ms = 0.f;
    MGPU_HIP(hipEventElapsedTime(&ms, pb.ev0, pb.ev1));
    if (rep) {
        rep->converged = hf[1];
        // reference: k at break, or K when the loop ran out (pcg_true.c:224)
        rep->iterations = hf[1] || (stop_after > 0 && kdone >= stop_after) ? kdone : maxiter;
        rep->res0 = hs[S_RES0];
        rep->res = hs[S_RES];
        rep->n_map = n;
        rep->trash_pix = pb.pt.trash;
        rep->lcount = 3 * npix + 3 * pb.pt.trash;
        rep->n_cut = pb.pc.n_cut;
        rep->t_solve_ms = ms;
        rep->refreshes = refreshes;
        rep->true_res_rel = -1.0;
    }
    if (mixed && rep) {
        // true residual of the returned x in FP64: ||b - A x|| / ||b|| (not timed)
        const bool keep = pb.prof.on;
        pb.prof.on = false;
        const int rc = apply_A64(pb, x, Ax, nullptr);
        pb.prof.on = keep;
        MGPU_RET(rc);
        k_refresh<<<grid_for(n, kMB), kMB, 0, s>>>(bvec, Ax, qv, n, pb.d_flag + 4);
        MGPU_LAUNCHED();
        const unsigned gd = grid_for(n, 256);
        k_dot_plain<<<gd, 256, 0, s>>>(qv, qv, n, part);
        k_sum_parts<<<1, 256, 0, s>>>(part, (int) gd, S + 20);
        k_dot_plain<<<gd, 256, 0, s>>>(bvec, bvec, n, part);
        k_sum_parts<<<1, 256, 0, s>>>(part, (int) gd, S + 21);
        MGPU_LAUNCHED();
        double two[2];
        MGPU_HIP(hipMemcpyAsync(two, S + 20, sizeof(two), hipMemcpyDeviceToHost, s));
        MGPU_HIP(hipStreamSynchronize(s));
        rep->true_res_rel = std::sqrt(two[0]) / std::sqrt(two[1]);
    }
    if (d_hist) pb.arena.release(d_hist);
    return MGPU_OK;
}

}  // namespace mgpu
