// twolvl.hip -- two-level preconditioners of PCG_GLS_true (precond = 1, 2).
//
// Reference: build_precond (precond.c:1310-1398), build_Z (798-953),
// build_AZ (1006-1030), build_Em1 (956-1003) with inverse_svd_bis (777-796),
// Lanczos_eig (1073-1306), and the two-level branch of apply_precond
// (1412-1422) with build_Qtx (1032-1055) and mul_ZQtx (1057-1071):
//
//   Cg = BJ (g - AZ Qtx) + Z Qtx,   Qtx = Em1 Z^T g,   Em1 = pinv(E),
//   E[i][j] = <Z_j, A Z_i>,         A = P^T N^-1 P.
//
// precond = 1 (a priori): Z from the per-chunk hit counts of the pointing
// (build_Z on one rank), A Z by Zn operator products.  precond = 2
// (a posteriori): Zn Lanczos steps on A from P^T N^-1 (signal + noise), the
// Ritz vectors of the Zn x Zn tridiagonal and their images.
//
// Device layout: Z and A Z are column-major (column j contiguous, stride
// 3 npix), so Z^T g is Zn coalesced dot products and Z Qtx / AZ Qtx are
// per-pixel combinations fused with the 3x3 block-Jacobi product and the PCG
// dots.  Per PCG iteration the preconditioner costs three launches: partial
// dots of Z^T g (chunks x Zn workgroups), one workgroup that sums them in a
// fixed order and multiplies by Em1, and the fused apply, which reads Z and
// A Z once.  All sums run in a fixed order, so results are deterministic.
// The Zn x Zn work (pinv of E, eigenvectors of the Lanczos tridiagonal) runs
// on the host: Zn is at most kMaxZn.
#include "pcg_device.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace mgpu {

constexpr int kZB = 256;        // threads of the dot workgroups
constexpr int kZChunk = 16384;  // map entries per dot workgroup

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// hits per (sample chunk, local pixel slot) for build_Z: chunk c = samples
// [c*per, (c+1)*per); slot = q + trash for valid samples, 0 for trash samples
// (the trash pixel's local indices 0..2 when trash != 0)
__global__ __launch_bounds__(256) void k_chunk_hits(const int *__restrict__ q, int64_t per, int64_t total, int trash,
                                                    int L, int *__restrict__ cnt) {
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int64_t c = t / per;
    const int qv = q[t];
    const int slot = qv < 0 ? 0 : qv + trash;
    atomicAdd(&cnt[c * L + slot], 1);
}

// column j of the a priori Z at valid pixel q (slot s = q + trash): the
// reference reads count[j*group + 3s] (precond.c:874; the count of local
// index i in chunk g is count[g*lcount + i], and every local index of a pixel
// holds that pixel's hit count) divided by the total count of 3s; Q and U
// entries are zero.
__global__ __launch_bounds__(256) void k_zprior(const int *__restrict__ cnt, int L, int group, int trash, int npix,
                                                int64_t ld, double *__restrict__ Z) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npix) return;
    const int j = blockIdx.y;
    const int slot = q + trash;
    int tc = 0;
    for (int c = 0; c < group; c++) tc += cnt[(int64_t) c * L + slot];
    const int64_t lc = 3 * (int64_t) L;
    const int64_t f = (int64_t) j * group + 3 * (int64_t) slot;
    const int64_t c = f / lc;
    const int idx = (int) ((f % lc) / 3);
    const int num = c < group ? cnt[c * L + idx] : 0;
    double *z = Z + (int64_t) j * ld + 3 * (int64_t) q;
    z[0] = (double) ((double) num / (double) tc);
    z[1] = 0.0;
    z[2] = 0.0;
}

// partial dots a.b over chunks of kZChunk entries
__global__ __launch_bounds__(kZB) void k_dot_part(const double *__restrict__ a, const double *__restrict__ b, int n,
                                                  double *__restrict__ part) {
    __shared__ double red[kZB / 64];
    const int k0 = blockIdx.x * kZChunk, k1 = min(n, k0 + kZChunk);
    double s = 0.0;
    for (int k = k0 + (int) threadIdx.x; k < k1; k += kZB) s += a[k] * b[k] * 1.0;
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// partial dots <Z_j, AZ_i> for every (i, j): blockIdx.x = i * zn + j, blockIdx.y = chunk
__global__ __launch_bounds__(kZB) void k_gram_part(const double *__restrict__ Z, const double *__restrict__ AZ,
                                                   int64_t ld, int n, int zn, int nchunk,
                                                   double *__restrict__ part) {
    __shared__ double red[kZB / 64];
    const int i = blockIdx.x / zn, j = blockIdx.x % zn;
    const double *z = Z + (int64_t) j * ld, *az = AZ + (int64_t) i * ld;
    const int k0 = blockIdx.y * kZChunk, k1 = min(n, k0 + kZChunk);
    double s = 0.0;
    for (int k = k0 + (int) threadIdx.x; k < k1; k += kZB) s += z[k] * az[k] * 1.0;
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[(int64_t) blockIdx.x * nchunk + blockIdx.y] = s;
}

// tod = s + n - P 0 (the start vector of Lanczos_eig, precond.c:1136-1138)
__global__ __launch_bounds__(256) void k_lz_rhs(const double *__restrict__ s, const double *__restrict__ nz,
                                                double *__restrict__ tod, int64_t m) {
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) tod[t] = s[t] + nz[t] - 0.0;
}

// v = w / beta; V[:, col] = v
__global__ __launch_bounds__(256) void k_lz_normalize(const double *__restrict__ w, double beta, double *__restrict__ v,
                                                      double *__restrict__ vcol, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = w[i] / beta;
    v[i] = x;
    vcol[i] = x;
}

// AmulV[:, col] = Av; w = Av - alpha v - beta vold; vold = v (beta = 0 and no
// vold for the first product, precond.c:1165-1168)
__global__ __launch_bounds__(256) void k_lz_step(const double *__restrict__ Av, double alpha, double beta,
                                                 const double *__restrict__ v, double *__restrict__ vold,
                                                 double *__restrict__ w, double *__restrict__ avcol, int n,
                                                 int first) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = Av[i];
    avcol[i] = a;
    if (first) {
        w[i] = a - (alpha * v[i]);
    } else {
        w[i] = a - (alpha * v[i]) - (beta * vold[i]);
        vold[i] = v[i];
    }
}

// out[:, j] = sum_k in[:, k] ev[k][j] (k ascending), the Ritz vectors of
// precond.c:1254-1282 with the eigenvectors in the columns of ev (K x K)
__global__ __launch_bounds__(256) void k_combine(const double *__restrict__ in, int64_t ld, int n,
                                                 const double *__restrict__ ev, int K, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = blockIdx.y;
    double s = 0.0;
    for (int k = 0; k < K; k++) s += in[(int64_t) k * ld + i] * ev[k * K + j];
    out[(int64_t) j * ld + i] = s;
}

// Z^T g partials: block (c, j) sums Z_j[k] g[k] over chunk c (build_Qtx,
// precond.c:1043-1046, pixpond = 1 on one rank)
__global__ __launch_bounds__(kZB) void k_zt_part(const double *__restrict__ Z, int64_t ld, int n,
                                                 const double *__restrict__ g, double *__restrict__ part, int nchunk,
                                                 const int *__restrict__ state, const int *__restrict__ flag) {
    if (flag && stopped(flag)) return;
    if (state && state[0]) return;
    __shared__ double red[kZB / 64];
    const int j = blockIdx.y;
    const double *z = Z + (int64_t) j * ld;
    const int k0 = blockIdx.x * kZChunk, k1 = min(n, k0 + kZChunk);
    double s = 0.0;
    for (int k = k0 + (int) threadIdx.x; k < k1; k += kZB) s += z[k] * g[k] * 1.0;
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[(int64_t) j * nchunk + blockIdx.x] = s;
}

// w = Z^T g summed over chunks in order, Qtx = Em1 w (precond.c:1050-1054)
__global__ __launch_bounds__(kZB) void k_zt_fin(const double *__restrict__ part, int nchunk, int zn,
                                                const double *__restrict__ Em1, double *__restrict__ qtx,
                                                const int *__restrict__ state, const int *__restrict__ flag) {
    if (flag && stopped(flag)) return;
    if (state && state[0]) return;
    __shared__ double w[kMaxZn];
    const int t = threadIdx.x;
    if (t < zn) {
        double s = 0.0;
        for (int c = 0; c < nchunk; c++) s += part[(int64_t) t * nchunk + c];
        w[t] = s;
    }
    __syncthreads();
    if (t < zn) {
        double qv = 0.0;
        for (int j = 0; j < zn; j++) qv += Em1[t * zn + j] * w[j];
        qtx[t] = qv;
    }
}

// Cg = BJ (g - AZ Qtx) + Z Qtx for the pixels of the workgroup, then the PCG
// dot partials: INIT: h = Cg, [<Cg,g> | <g,g>]; else [<Cg,g> | <Cg,g-gp> | <g,g>]
template <bool INIT>
__global__ __launch_bounds__(kMB) void k_twolvl_apply(
        const double *__restrict__ Z, const double *__restrict__ AZ, int64_t ld, int zn,
        const double *__restrict__ qtx, const double *__restrict__ inv, const double *__restrict__ g,
        const double *__restrict__ gp, double *__restrict__ Cg, double *__restrict__ h,
        const int *__restrict__ wg_seg, const int *__restrict__ wg_q0, const int *__restrict__ seg_q,
        const int *__restrict__ state, double *__restrict__ part, const int *__restrict__ flag) {
    if (flag && stopped(flag)) return;
    const SegPix sp = seg_pixel(wg_seg, wg_q0, seg_q);
    if (state && state[sp.r]) return;  // uniform over the workgroup
    __shared__ double red[kMB / 64];
    __shared__ double qs[kMaxZn];
    for (int j = threadIdx.x; j < zn; j += kMB) qs[j] = qtx[j];
    __syncthreads();
    const int q = sp.q;
    double a = 0.0, b = 0.0, c = 0.0;
    if (q < sp.qend) {
        double qg[3], r[3], gv[3];
        for (int e = 0; e < 3; e++) {
            const int64_t i = 3 * (int64_t) q + e;
            double zq = 0.0, azq = 0.0;  // mul_ZQtx: vec[k] += Z[j][k] Qtx[j], j ascending
            for (int j = 0; j < zn; j++) {
                zq += Z[(int64_t) j * ld + i] * qs[j];
                azq += AZ[(int64_t) j * ld + i] * qs[j];
            }
            gv[e] = g[i];
            qg[e] = zq;
            r[e] = gv[e] - azq;
        }
        const double *M = inv + 9 * (int64_t) q;
        for (int e = 0; e < 3; e++) {
            double y = 0.0;
            y += M[3 * e] * r[0];
            y += M[3 * e + 1] * r[1];
            y += M[3 * e + 2] * r[2];
            y += qg[e];
            Cg[3 * q + e] = y;
            if (INIT) h[3 * q + e] = y;
            a += y * gv[e] * 1.0;
            if (!INIT) b += y * (gv[e] - gp[3 * q + e]) * 1.0;
            c += gv[e] * gv[e] * 1.0;
        }
    }
    a = block_sum(a, red);
    if (!INIT) b = block_sum(b, red);
    c = block_sum(c, red);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        if (INIT) {
            part[gridDim.x + blockIdx.x] = c;
        } else {
            part[gridDim.x + blockIdx.x] = b;
            part[2 * gridDim.x + blockIdx.x] = c;
        }
    }
}

// ---------------------------------------------------------------------------
// host: small dense algebra
// ---------------------------------------------------------------------------

// pinv of the row-major n x n matrix a with singular values <= rcond * s_max
// dropped: the minimum-norm solution of min ||a X - I|| that inverse_svd_bis
// obtains from LAPACKE_dgelss (precond.c:777-796, rcond = 1e-15,
// precond.c:25).  One-sided Jacobi SVD (Hestenes 1958): rotate column pairs
// of W = a until orthogonal, a = W V^T, s_j = ||W_j||; pinv = sum_j V_j W_j^T / s_j^2.
static void pinv_jacobi(int n, const std::vector<double> &a, double rcond, std::vector<double> &out) {
    std::vector<double> W(a), V((size_t) n * n, 0.0), s2((size_t) n);
    for (int i = 0; i < n; i++) V[(size_t) i * n + i] = 1.0;
    for (int sweep = 0; sweep < 80; sweep++) {
        bool rotated = false;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double al = 0, be = 0, ga = 0;
                for (int i = 0; i < n; i++) {
                    const double wp = W[(size_t) i * n + p], wq = W[(size_t) i * n + q];
                    al += wp * wp;
                    be += wq * wq;
                    ga += wp * wq;
                }
                if (ga == 0.0 || std::fabs(ga) <= 1e-15 * std::sqrt(al * be)) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
                for (int i = 0; i < n; i++) {
                    double &wp = W[(size_t) i * n + p], &wq = W[(size_t) i * n + q];
                    const double x = wp, y = wq;
                    wp = c * x - sn * y;
                    wq = sn * x + c * y;
                    double &vp = V[(size_t) i * n + p], &vq = V[(size_t) i * n + q];
                    const double u = vp, v = vq;
                    vp = c * u - sn * v;
                    vq = sn * u + c * v;
                }
            }
        if (!rotated) break;
    }
    double smax = 0.0;
    for (int j = 0; j < n; j++) {
        double t = 0.0;
        for (int i = 0; i < n; i++) t += W[(size_t) i * n + j] * W[(size_t) i * n + j];
        s2[j] = t;
        smax = std::max(smax, std::sqrt(t));
    }
    out.assign((size_t) n * n, 0.0);
    for (int j = 0; j < n; j++) {
        if (!(std::sqrt(s2[j]) > rcond * smax)) continue;
        for (int i = 0; i < n; i++)
            for (int k = 0; k < n; k++) out[(size_t) i * n + k] += V[(size_t) i * n + j] * W[(size_t) k * n + j] / s2[j];
    }
}

// eigenvectors of the symmetric row-major n x n matrix a by cyclic Jacobi
// rotations (Golub & Van Loan 8.5), as columns of ev ordered by ascending
// eigenvalue -- what LAPACKE_dsyev(ROW_MAJOR, 'V') leaves in its matrix
// (precond.c:1240)
static void symeig_jacobi(int n, const std::vector<double> &a, std::vector<double> &ev) {
    std::vector<double> A(a), V((size_t) n * n, 0.0);
    for (int i = 0; i < n; i++) V[(size_t) i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                const double x = A[(size_t) i * n + j];
                tot += x * x;
                if (i != j) off += x * x;
            }
        if (off == 0.0 || off <= 1e-32 * tot) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A[(size_t) p * n + q];
                if (apq == 0.0) continue;
                const double theta = (A[(size_t) q * n + q] - A[(size_t) p * n + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    double &x = A[(size_t) k * n + p], &y = A[(size_t) k * n + q];
                    const double u = x, v = y;
                    x = c * u - s * v;
                    y = s * u + c * v;
                }
                for (int k = 0; k < n; k++) {
                    double &x = A[(size_t) p * n + k], &y = A[(size_t) q * n + k];
                    const double u = x, v = y;
                    x = c * u - s * v;
                    y = s * u + c * v;
                }
                for (int k = 0; k < n; k++) {
                    double &x = V[(size_t) k * n + p], &y = V[(size_t) k * n + q];
                    const double u = x, v = y;
                    x = c * u - s * v;
                    y = s * u + c * v;
                }
            }
    }
    std::vector<int> ord((size_t) n);
    for (int i = 0; i < n; i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int x, int y) { return A[(size_t) x * n + x] < A[(size_t) y * n + y]; });
    ev.assign((size_t) n * n, 0.0);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) ev[(size_t) i * n + j] = V[(size_t) i * n + ord[j]];
}

// ---------------------------------------------------------------------------
// host: build and apply
// ---------------------------------------------------------------------------

void twolvl_release(Problem &pb) {
    TwoLevel &T = pb.tl;
    for (double **p : {&T.d_Z, &T.d_AZ, &T.d_Em1, &T.d_qtx, &T.d_part})
        if (*p) {
            pb.arena.release(*p);
            *p = nullptr;
        }
    T.zn = 0;
    T.ld = 0;
    T.nchunk = 0;
}

// sequential sum of device partials on the host (deterministic)
static int host_sum(Problem &pb, const double *d_part, int np, double *out) {
    std::vector<double> h((size_t) np);
    MGPU_HIP(hipMemcpyAsync(h.data(), d_part, sizeof(double) * np, hipMemcpyDeviceToHost, pb.stream));
    MGPU_HIP(hipStreamSynchronize(pb.stream));
    double s = 0.0;
    for (int i = 0; i < np; i++) s += h[i];
    *out = s;
    return MGPU_OK;
}

// Lanczos_eig (precond.c:1073-1306) from x = 0: V and AmulV get K+1 columns;
// the Ritz combination of the first K fills Z and A Z.  Kept as written: the
// first loop step subtracts no beta * v_prev (vold starts at 0).
static int lanczos(Problem &pb, const double *d_signal, const double *d_noise, int K, double *d_V, double *d_AV) {
    TwoLevel &T = pb.tl;
    hipStream_t s = pb.stream;
    const int n = 3 * pb.pt.npix;
    const int64_t m = pb.pt.m, ld = T.ld;
    const double eps = 1.0e-15;  // precond.c:25
    double *d_w, *d_v, *d_vold, *d_Av, *d_dp;
    MGPU_RET(pb.arena.alloc(&d_w, (size_t) n));
    MGPU_RET(pb.arena.alloc(&d_v, (size_t) n));
    MGPU_RET(pb.arena.alloc(&d_vold, (size_t) n));
    MGPU_RET(pb.arena.alloc(&d_Av, (size_t) n));
    MGPU_RET(pb.arena.alloc(&d_dp, (size_t) T.nchunk));
    MGPU_HIP(hipMemsetAsync(d_v, 0, sizeof(double) * n, s));
    MGPU_HIP(hipMemsetAsync(d_vold, 0, sizeof(double) * n, s));
    auto dot = [&](const double *a, const double *b, double *r) -> int {
        k_dot_part<<<T.nchunk, kZB, 0, s>>>(a, b, n, d_dp);
        MGPU_LAUNCHED();
        return host_sum(pb, d_dp, T.nchunk, r);
    };
    const unsigned gn = grid_for(n, 256);
    // w = P^T N^-1 (b + noise - P x), x = 0
    k_lz_rhs<<<grid_for(m, 256), 256, 0, s>>>(d_signal, d_noise, pb.d_tod, m);
    MGPU_LAUNCHED();
    MGPU_RET(toeplitz_apply(pb, pb.d_tod, pb.d_tod2, true));
    MGPU_RET(launch_trmatvec(pb, pb.d_tod2, d_w, nullptr, nullptr, nullptr));
    const int K1 = K + 1;
    std::vector<double> Tm((size_t) K1 * K1, 0.0);
    double r, beta, alpha;
    MGPU_RET(dot(d_w, d_w, &r));
    beta = std::sqrt(r);
    if (beta > eps) {
        k_lz_normalize<<<gn, 256, 0, s>>>(d_w, beta, d_v, d_V, n);
        MGPU_LAUNCHED();
    }
    MGPU_RET(apply_A64(pb, d_v, d_Av, nullptr));
    MGPU_RET(dot(d_v, d_Av, &alpha));
    k_lz_step<<<gn, 256, 0, s>>>(d_Av, alpha, 0.0, d_v, d_vold, d_w, d_AV, n, 1);
    MGPU_LAUNCHED();
    for (int i = 0; i < K; i++) {
        MGPU_RET(dot(d_w, d_w, &r));
        beta = std::sqrt(r);
        if (beta > eps) {
            k_lz_normalize<<<gn, 256, 0, s>>>(d_w, beta, d_v, d_V + (int64_t) (i + 1) * ld, n);
            MGPU_LAUNCHED();
        }
        Tm[(size_t) i * K1 + i] = alpha;
        Tm[(size_t) i * K1 + i + 1] = beta;
        Tm[(size_t) (i + 1) * K1 + i] = beta;
        MGPU_RET(apply_A64(pb, d_v, d_Av, nullptr));
        MGPU_RET(dot(d_v, d_Av, &alpha));
        k_lz_step<<<gn, 256, 0, s>>>(d_Av, alpha, beta, d_v, d_vold, d_w, d_AV + (int64_t) (i + 1) * ld, n, 0);
        MGPU_LAUNCHED();
    }
    // T restricted to K x K, its eigenvectors, Ritz vectors Z = V ev and A Z = AmulV ev
    std::vector<double> Tk((size_t) K * K), ev;
    for (int i = 0; i < K; i++)
        for (int j = 0; j < K; j++) Tk[(size_t) i * K + j] = Tm[(size_t) i * K1 + j];
    symeig_jacobi(K, Tk, ev);
    double *d_ev;
    MGPU_RET(pb.arena.alloc(&d_ev, (size_t) K * K));
    MGPU_HIP(hipMemcpyAsync(d_ev, ev.data(), sizeof(double) * K * K, hipMemcpyHostToDevice, s));
    k_combine<<<dim3(gn, K), 256, 0, s>>>(d_V, ld, n, d_ev, K, T.d_Z);
    MGPU_LAUNCHED();
    k_combine<<<dim3(gn, K), 256, 0, s>>>(d_AV, ld, n, d_ev, K, T.d_AZ);
    MGPU_LAUNCHED();
    MGPU_HIP(hipStreamSynchronize(s));
    for (double *p : {d_w, d_v, d_vold, d_Av, d_dp, d_ev}) pb.arena.release(p);
    return MGPU_OK;
}

int twolvl_build(Problem &pb, const double *d_signal, const double *d_noise) {
    const int precond = pb.opt.precond;
    const int npix = pb.pt.npix, n = 3 * npix;
    const int zn = pb.opt.z_2lvl == 0 ? 1 : pb.opt.z_2lvl;  // Z_2lvl = 0: communicator size (pcg_true.c:50)
    MGPU_ARG(precond == 1 || precond == 2, "precond=%d: 0 block Jacobi, 1 two-level a priori, 2 a posteriori",
             precond);
    MGPU_ARG(pb.opt.z_2lvl >= 0 && zn <= kMaxZn && zn <= n, "z_2lvl=%d must be in [0, min(%d, n_map=%d)]",
             pb.opt.z_2lvl, kMaxZn, n);
    MGPU_ARG(pb.sg.nseg == 1, "two-level preconditioners solve one problem at a time");
    twolvl_release(pb);
    // the products of build_AZ and Lanczos_eig call stbmmProd (precond.c:1017-1022,
    // 1151-1153, 1205-1207): with lambda 1 they also remove the block means
    struct StbmmOffsets {
        Problem &pb;
        ~StbmmOffsets() { pb.stbmm_offsets = false; }
    } stbmm{pb};
    pb.stbmm_offsets = true;
    TwoLevel &T = pb.tl;
    hipStream_t s = pb.stream;
    MGPU_HIP(hipEventRecord(pb.ev0, s));
    T.zn = zn;
    T.ld = n;
    T.nchunk = (int) ((n + kZChunk - 1) / kZChunk);
    const int64_t ld = T.ld;
    MGPU_RET(pb.arena.alloc(&T.d_Em1, (size_t) zn * zn));
    MGPU_RET(pb.arena.alloc(&T.d_qtx, (size_t) zn));
    MGPU_RET(pb.arena.alloc(&T.d_part, (size_t) zn * T.nchunk));
    MGPU_RET(pb.arena.alloc(&T.d_Z, (size_t) zn * ld));
    MGPU_RET(pb.arena.alloc(&T.d_AZ, (size_t) zn * ld));
    if (precond == 1) {
        // build_Z (precond.c:798-953) on one rank: group = Zn, chunk = m / group samples
        const int L = npix + pb.pt.trash;
        const int64_t per = pb.pt.m / zn, total = per * zn;
        MGPU_ARG(per > 0, "z_2lvl=%d exceeds the number of samples", zn);
        int *d_cnt;
        MGPU_RET(pb.arena.alloc(&d_cnt, (size_t) zn * L));
        MGPU_HIP(hipMemsetAsync(d_cnt, 0, sizeof(int) * (size_t) zn * L, s));
        k_chunk_hits<<<grid_for(total, 256), 256, 0, s>>>(pb.pt.d_q, per, total, pb.pt.trash, L, d_cnt);
        MGPU_LAUNCHED();
        k_zprior<<<dim3(grid_for(npix, 256), zn), 256, 0, s>>>(d_cnt, L, zn, pb.pt.trash, npix, ld, T.d_Z);
        MGPU_LAUNCHED();
        // build_AZ (precond.c:1006-1030)
        for (int j = 0; j < zn; j++) MGPU_RET(apply_A64(pb, T.d_Z + j * ld, T.d_AZ + j * ld, nullptr));
        MGPU_HIP(hipStreamSynchronize(s));
        pb.arena.release(d_cnt);
    } else {
        double *d_V, *d_AV;
        MGPU_RET(pb.arena.alloc(&d_V, (size_t) (zn + 1) * ld));
        MGPU_RET(pb.arena.alloc(&d_AV, (size_t) (zn + 1) * ld));
        MGPU_HIP(hipMemsetAsync(d_V, 0, sizeof(double) * (size_t) (zn + 1) * ld, s));
        const int rc = lanczos(pb, d_signal, d_noise, zn, d_V, d_AV);
        MGPU_HIP(hipStreamSynchronize(s));
        pb.arena.release(d_V);
        pb.arena.release(d_AV);
        MGPU_RET(rc);
    }
    // build_Em1 (precond.c:956-1003): E[i*Zn + j] = <Z_j, AZ_i>, Em1 = pinv(E)
    double *d_gp;
    const int nprt = zn * zn * T.nchunk;
    MGPU_RET(pb.arena.alloc(&d_gp, (size_t) nprt));
    k_gram_part<<<dim3(zn * zn, T.nchunk), kZB, 0, s>>>(T.d_Z, T.d_AZ, ld, n, zn, T.nchunk, d_gp);
    MGPU_LAUNCHED();
    std::vector<double> hp((size_t) nprt), E((size_t) zn * zn), Em1;
    MGPU_HIP(hipMemcpyAsync(hp.data(), d_gp, sizeof(double) * nprt, hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipStreamSynchronize(s));
    pb.arena.release(d_gp);
    for (int y = 0; y < zn * zn; y++) {
        double e = 0.0;
        for (int c = 0; c < T.nchunk; c++) e += hp[(size_t) y * T.nchunk + c];
        E[y] = e;
    }
    pinv_jacobi(zn, E, 1.0e-15, Em1);
    MGPU_HIP(hipMemcpyAsync(T.d_Em1, Em1.data(), sizeof(double) * zn * zn, hipMemcpyHostToDevice, s));
    MGPU_HIP(hipEventRecord(pb.ev1, s));
    MGPU_HIP(hipEventSynchronize(pb.ev1));
    float ms = 0.f;
    MGPU_HIP(hipEventElapsedTime(&ms, pb.ev0, pb.ev1));
    T.build_ms = ms;
    return MGPU_OK;
}

int twolvl_apply(Problem &pb, const double *d_g, const double *d_gp, double *d_Cg, double *d_h, double *d_part,
                 bool init, const int *d_state, const int *d_flag) {
    const TwoLevel &T = pb.tl;
    const Segments &G = pb.sg;
    hipStream_t s = pb.stream;
    const int n = 3 * pb.pt.npix;
    MGPU_ARG(T.zn > 0 && T.ld == n, "two-level preconditioner not built for this map");
    k_zt_part<<<dim3(T.nchunk, T.zn), kZB, 0, s>>>(T.d_Z, T.ld, n, d_g, T.d_part, T.nchunk, d_state, d_flag);
    MGPU_LAUNCHED();
    k_zt_fin<<<1, kZB, 0, s>>>(T.d_part, T.nchunk, T.zn, T.d_Em1, T.d_qtx, d_state, d_flag);
    MGPU_LAUNCHED();
    if (init)
        k_twolvl_apply<true><<<G.nwg, kMB, 0, s>>>(T.d_Z, T.d_AZ, T.ld, T.zn, T.d_qtx, pb.pc.d_inv, d_g, d_gp, d_Cg,
                                                   d_h, G.d_wg_seg, G.d_wg_q0, G.d_q, d_state, d_part, d_flag);
    else
        k_twolvl_apply<false><<<G.nwg, kMB, 0, s>>>(T.d_Z, T.d_AZ, T.ld, T.zn, T.d_qtx, pb.pc.d_inv, d_g, d_gp,
                                                    d_Cg, d_h, G.d_wg_seg, G.d_wg_q0, G.d_q, d_state, d_part,
                                                    d_flag);
    MGPU_LAUNCHED();
    return MGPU_OK;
}

}  // namespace mgpu
