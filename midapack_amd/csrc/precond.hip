// precond.hip -- block-Jacobi preconditioner with the degenerate-pixel cut (gfx950).
//
// Reference semantics (one rank, precond=0):
//   getlocalW      precond.c:42-161: B_q[j][l] += (w_j w_l) T_b(t)[0] over the
//                  samples t of pixel q in increasing t (samples after the last
//                  block get w_j w_l); lhits_q = number of samples.
//   precondblockjacobilike  precond.c:456-702, per pixel:
//                  anorm = dlange('1'), dgetrf, rcond = dgecon('1');
//                  rcond > 0.1 -> BJ_inv = dgetri, else the pixel is cut: its
//                  samples point to the trash pixel, their signal is zeroed,
//                  and the pointing matrix is rebuilt (MatInit again).
//   apply_precond  precond.c:1403-1409: Cg = BJ_inv g (MatVecProd, nnz = 3).
//
// The LAPACK routines are restated for n = 3 in the operation order of the
// reference LAPACK/BLAS (dgetrf2 recursion with DLASWP/DTRSM/DGEMM, DGECON with
// DLACN2 and the DLATRS->DTRSV path, DGETRI with DTRTI2 and the unblocked
// column sweep).  With contraction disabled every value is the IEEE result of
// that sequence, so the kernel is bit-exact with the CPU restatement.
//
// One lane per pixel; the pixel's samples are read from the sliced-ELL layout
// (coalesced), the block index of each sample advances monotonically.
#include <hipcub/hipcub.hpp>

#include "mgpu_internal.h"

namespace mgpu {

constexpr int kSliceP = 64;

#define A3(i, j) a[(i) * 3 + (j)]

__device__ __forceinline__ double dlange1_3(const double *a) {
    double value = 0.0;
    for (int j = 0; j < 3; j++) {
        double sum = 0.0;
        for (int i = 0; i < 3; i++) sum += fabs(A3(i, j));
        if (value < sum || isnan(sum)) value = sum;
    }
    return value;
}

// DGETRF2 recursion unrolled for n = 3 (n1 = 1, then the 2x2 with n1 = 1)
__device__ __forceinline__ int dgetrf_3(double *a, int *ipiv) {
    const double sfmin = 2.2250738585072014e-308;
    int info = 0;
    // -- panel: column 0, rows 0..2
    {
        int ip = 0;
        double dmax = fabs(A3(0, 0));
        if (fabs(A3(1, 0)) > dmax) {
            ip = 1;
            dmax = fabs(A3(1, 0));
        }
        if (fabs(A3(2, 0)) > dmax) ip = 2;
        ipiv[0] = ip;
        if (A3(ip, 0) != 0.0) {
            if (ip != 0) {
                double t = A3(0, 0);
                A3(0, 0) = A3(ip, 0);
                A3(ip, 0) = t;
            }
            if (fabs(A3(0, 0)) >= sfmin) {
                const double r = 1.0 / A3(0, 0);
                A3(1, 0) = r * A3(1, 0);
                A3(2, 0) = r * A3(2, 0);
            } else {
                A3(1, 0) = A3(1, 0) / A3(0, 0);
                A3(2, 0) = A3(2, 0) / A3(0, 0);
            }
        } else {
            info = 1;
        }
    }
    // -- DLASWP on columns 1..2 for pivot 0
    if (ipiv[0] != 0)
        for (int j = 1; j < 3; j++) {
            double t = A3(0, j);
            A3(0, j) = A3(ipiv[0], j);
            A3(ipiv[0], j) = t;
        }
    // -- DTRSM with a 1x1 unit triangle: nothing
    // -- DGEMM: A22 <- A22 - A21 A12
    for (int j = 1; j < 3; j++) {
        const double temp = -1.0 * A3(0, j);
        for (int i = 1; i < 3; i++) A3(i, j) = A3(i, j) + temp * A3(i, 0);
    }
    // -- 2x2 recursion on rows/cols 1..2
    int info2 = 0;
    {
        int ip = 1;
        if (fabs(A3(2, 1)) > fabs(A3(1, 1))) ip = 2;
        ipiv[1] = ip;
        if (A3(ip, 1) != 0.0) {
            if (ip != 1) {
                double t = A3(1, 1);
                A3(1, 1) = A3(ip, 1);
                A3(ip, 1) = t;
            }
            if (fabs(A3(1, 1)) >= sfmin) {
                const double r = 1.0 / A3(1, 1);
                A3(2, 1) = r * A3(2, 1);
            } else {
                A3(2, 1) = A3(2, 1) / A3(1, 1);
            }
        } else {
            info2 = 1;
        }
        if (ipiv[1] != 1) {
            double t = A3(1, 2);
            A3(1, 2) = A3(ipiv[1], 2);
            A3(ipiv[1], 2) = t;
        }
        const double temp = -1.0 * A3(1, 2);
        A3(2, 2) = A3(2, 2) + temp * A3(2, 1);
        ipiv[2] = 2;
        if (info2 == 0 && A3(2, 2) == 0.0) info2 = 2;
    }
    if (info == 0 && info2 > 0) info = info2 + 1;
    // -- DLASWP on column 0 for pivots 1..2
    for (int k = 1; k < 3; k++) {
        const int p = ipiv[k];
        if (p != k) {
            double t = A3(k, 0);
            A3(k, 0) = A3(p, 0);
            A3(p, 0) = t;
        }
    }
    return info;
}

__device__ __forceinline__ int idamax3(const double *x) {
    int imx = 0;
    double dmax = fabs(x[0]);
    if (fabs(x[1]) > dmax) {
        imx = 1;
        dmax = fabs(x[1]);
    }
    if (fabs(x[2]) > dmax) imx = 2;
    return imx;
}

__device__ __forceinline__ void trsv_lower_unit_n3(const double *a, double *x) {
    for (int j = 0; j < 3; j++)
        if (x[j] != 0.0) {
            const double temp = x[j];
            for (int i = j + 1; i < 3; i++) x[i] = x[i] - temp * A3(i, j);
        }
}
__device__ __forceinline__ void trsv_upper_nonunit_n3(const double *a, double *x) {
    for (int j = 2; j >= 0; j--)
        if (x[j] != 0.0) {
            x[j] = x[j] / A3(j, j);
            const double temp = x[j];
            for (int i = j - 1; i >= 0; i--) x[i] = x[i] - temp * A3(i, j);
        }
}
__device__ __forceinline__ void trsv_upper_nonunit_t3(const double *a, double *x) {
    for (int j = 0; j < 3; j++) {
        double temp = x[j];
        for (int i = 0; i < j; i++) temp = temp - A3(i, j) * x[i];
        temp = temp / A3(j, j);
        x[j] = temp;
    }
}
__device__ __forceinline__ void trsv_lower_unit_t3(const double *a, double *x) {
    for (int j = 2; j >= 0; j--) {
        double temp = x[j];
        for (int i = 2; i > j; i--) temp = temp - A3(i, j) * x[i];
        x[j] = temp;
    }
}

// DLATRS growth test: 1 when DLATRS takes its DTRSV path
__device__ __forceinline__ bool latrs_fast_ok3(const double *a, const double *x, bool upper, bool unit, bool trans) {
    const double smlnum = 2.2250738585072014e-308 / 1.1102230246251565e-16;
    double cnorm[3];
    for (int j = 0; j < 3; j++) {
        double s = 0.0;
        if (upper)
            for (int i = 0; i < j; i++) s += fabs(A3(i, j));
        else
            for (int i = j + 1; i < 3; i++) s += fabs(A3(i, j));
        cnorm[j] = s;
    }
    const double xmax = fabs(x[idamax3(x)]);
    double grow;
    if (!trans) {
        const int jfirst = upper ? 2 : 0, jlast = upper ? -1 : 3, jinc = upper ? -1 : 1;
        if (!unit) {
            grow = 1.0 / fmax(xmax, smlnum);
            double xbnd = grow;
            for (int j = jfirst; j != jlast; j += jinc) {
                if (grow <= smlnum) break;
                const double tjj = fabs(A3(j, j));
                xbnd = fmin(xbnd, fmin(1.0, tjj) * grow);
                if (tjj + cnorm[j] >= smlnum)
                    grow = grow * (tjj / (tjj + cnorm[j]));
                else
                    grow = 0.0;
            }
            grow = xbnd;
        } else {
            grow = fmin(1.0, 1.0 / fmax(xmax, smlnum));
            for (int j = jfirst; j != jlast; j += jinc) {
                if (grow <= smlnum) break;
                grow = grow * (1.0 / (1.0 + cnorm[j]));
            }
        }
    } else {
        const int jfirst = upper ? 0 : 2, jlast = upper ? 3 : -1, jinc = upper ? 1 : -1;
        if (!unit) {
            grow = 1.0 / fmax(xmax, smlnum);
            double xbnd = grow;
            for (int j = jfirst; j != jlast; j += jinc) {
                if (grow <= smlnum) break;
                const double xj = 1.0 + cnorm[j];
                grow = fmin(grow, xbnd / xj);
                const double tjj = fabs(A3(j, j));
                if (xj > tjj) xbnd = xbnd * (tjj / xj);
            }
            grow = fmin(grow, xbnd);
        } else {
            grow = fmin(1.0, 1.0 / fmax(xmax, smlnum));
            for (int j = jfirst; j != jlast; j += jinc) {
                if (grow <= smlnum) break;
                grow = grow / (1.0 + cnorm[j]);
            }
        }
    }
    return grow > smlnum;
}

// DGECON('1') for n = 3 with the DLACN2 state machine
__device__ double dgecon1_3(const double *lu, double anorm) {
    if (anorm == 0.0) return 0.0;
    if (isnan(anorm)) return anorm;
    if (isinf(anorm)) return 0.0;
    double x[3], v[3];
    int isgn[3];
    double est = 0.0, estold;
    int kase = 1, jump = 1, iter = 0, j = 0, jlast;
    for (int i = 0; i < 3; i++) x[i] = 1.0 / 3.0;
    for (int guard = 0; guard < 32; guard++) {
        if (kase == 1) {
            if (!latrs_fast_ok3(lu, x, false, true, false)) return 0.0;
            trsv_lower_unit_n3(lu, x);
            if (!latrs_fast_ok3(lu, x, true, false, false)) return 0.0;
            trsv_upper_nonunit_n3(lu, x);
        } else {
            if (!latrs_fast_ok3(lu, x, true, false, true)) return 0.0;
            trsv_upper_nonunit_t3(lu, x);
            if (!latrs_fast_ok3(lu, x, false, true, true)) return 0.0;
            trsv_lower_unit_t3(lu, x);
        }
        bool restart50 = false, final120 = false;
        if (jump == 1) {
            est = 0.0;
            for (int i = 0; i < 3; i++) est += fabs(x[i]);
            for (int i = 0; i < 3; i++) {
                x[i] = (x[i] >= 0.0) ? 1.0 : -1.0;
                isgn[i] = (int) x[i];
            }
            kase = 2;
            jump = 2;
            continue;
        } else if (jump == 2) {
            j = idamax3(x);
            iter = 2;
            restart50 = true;
        } else if (jump == 3) {
            for (int i = 0; i < 3; i++) v[i] = x[i];
            estold = est;
            est = 0.0;
            for (int i = 0; i < 3; i++) est += fabs(v[i]);
            bool same = true;
            for (int i = 0; i < 3; i++) {
                const double xs = (x[i] >= 0.0) ? 1.0 : -1.0;
                if ((int) xs != isgn[i]) {
                    same = false;
                    break;
                }
            }
            if (same || est <= estold) {
                final120 = true;
            } else {
                for (int i = 0; i < 3; i++) {
                    x[i] = (x[i] >= 0.0) ? 1.0 : -1.0;
                    isgn[i] = (int) x[i];
                }
                kase = 2;
                jump = 4;
                continue;
            }
        } else if (jump == 4) {
            jlast = j;
            j = idamax3(x);
            if (x[jlast] != fabs(x[j]) && iter < 5) {
                iter++;
                restart50 = true;
            } else {
                final120 = true;
            }
        } else {  // jump == 5
            double s = 0.0;
            for (int i = 0; i < 3; i++) s += fabs(x[i]);
            const double temp = 2.0 * (s / 9.0);
            if (temp > est) est = temp;
            break;
        }
        if (restart50) {
            for (int i = 0; i < 3; i++) x[i] = 0.0;
            x[j] = 1.0;
            kase = 1;
            jump = 3;
        } else if (final120) {
            double altsgn = 1.0;
            for (int i = 0; i < 3; i++) {
                x[i] = altsgn * (1.0 + (double) i / 2.0);
                altsgn = -altsgn;
            }
            kase = 1;
            jump = 5;
        }
    }
    (void) v;
    if (est == 0.0) return 0.0;
    return (1.0 / est) / anorm;
}

// DGETRI for n = 3 (DTRTI2 + unblocked sweep + column interchanges)
__device__ __forceinline__ int dgetri_3(double *a, const int *ipiv) {
    for (int i = 0; i < 3; i++)
        if (A3(i, i) == 0.0) return i + 1;
    for (int j = 0; j < 3; j++) {
        A3(j, j) = 1.0 / A3(j, j);
        const double ajj = -A3(j, j);
        for (int jj = 0; jj < j; jj++)
            if (A3(jj, j) != 0.0) {
                const double temp = A3(jj, j);
                for (int i = 0; i < jj; i++) A3(i, j) = A3(i, j) + temp * A3(i, jj);
                A3(jj, j) = A3(jj, j) * A3(jj, jj);
            }
        for (int i = 0; i < j; i++) A3(i, j) = ajj * A3(i, j);
    }
    double work[3];
    for (int j = 2; j >= 0; j--) {
        for (int i = j + 1; i < 3; i++) {
            work[i] = A3(i, j);
            A3(i, j) = 0.0;
        }
        for (int jj = j + 1; jj < 3; jj++) {
            const double temp = -1.0 * work[jj];
            for (int i = 0; i < 3; i++) A3(i, j) = A3(i, j) + temp * A3(i, jj);
        }
    }
    for (int j = 1; j >= 0; j--) {
        const int jp = ipiv[j];
        if (jp != j)
            for (int i = 0; i < 3; i++) {
                const double t = A3(i, j);
                A3(i, j) = A3(i, jp);
                A3(i, jp) = t;
            }
    }
    return 0;
}
#undef A3

// one lane per pixel: accumulate B over the ELL hits, then rcond / inverse
template <bool W0>
__global__ __launch_bounds__(256) void k_blocks(const int64_t *__restrict__ base, const int *__restrict__ ct,
                                                const double *__restrict__ cw0, const double2 *__restrict__ cw12,
                                                int npix,
                                                const int64_t *__restrict__ start, int nb,
                                                const double *__restrict__ T0, double *__restrict__ blocks_out,
                                                double *__restrict__ inv, double *__restrict__ cond,
                                                int *__restrict__ hits, unsigned *__restrict__ kept) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npix) return;
    const int s = q / kSliceP, lane = q % kSliceP;
    const int64_t b0 = base[s], width = (base[s + 1] - b0) / kSliceP;
    double B[9];
    for (int i = 0; i < 9; i++) B[i] = 0.0;
    int nh = 0, blk = 0;
    const int64_t covered = start[nb];
    for (int64_t i = 0; i < width; i++) {
        const int64_t e = b0 + i * kSliceP + lane;
        const int t = ct[e];
        if (t < 0) break;
        while (blk < nb && start[blk + 1] <= t) blk++;
        const double2 wqu = cw12[e];
        const double w[3] = {W0 ? cw0[e] : 1.0, wqu.x, wqu.y};
        nh++;
        if (t < covered) {
            const double d = T0[blk];
            for (int j = 0; j < 3; j++)
                for (int l = 0; l < 3; l++) B[3 * j + l] += w[j] * w[l] * d;
        } else {
            for (int j = 0; j < 3; j++)
                for (int l = 0; l < 3; l++) B[3 * j + l] += w[j] * w[l];
        }
    }
    if (blocks_out)
        for (int i = 0; i < 9; i++) blocks_out[9 * (int64_t) q + i] = B[i];
    const double anorm = dlange1_3(B);
    int ipiv[3];
    dgetrf_3(B, ipiv);
    const double rcond = dgecon1_3(B, anorm);
    cond[q] = rcond;
    hits[q] = nh;
    const bool ok = rcond > 1e-1;
    kept[q] = ok ? 1u : 0u;
    if (ok) {
        dgetri_3(B, ipiv);
        for (int i = 0; i < 9; i++) inv[9 * (int64_t) q + i] = B[i];
    } else {
        for (int i = 0; i < 9; i++) inv[9 * (int64_t) q + i] = 0.0;
    }
}

// signal and weights of the samples of cut pixels are zeroed (precond.c:599-616)
__global__ void k_cut_samples(const int *__restrict__ q, int64_t m, const unsigned *__restrict__ kept,
                              double *__restrict__ signal) {
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) {
        const int qq = q[t];
        if (qq >= 0 && !kept[qq] && signal) signal[t] = 0.0;
    }
}

__global__ void k_compact_precond(const unsigned *__restrict__ kept, const unsigned *__restrict__ newidx, int npix,
                                  const double *__restrict__ inv_in, const double *__restrict__ cond_in,
                                  const int *__restrict__ hits_in, double *__restrict__ inv_out,
                                  double *__restrict__ cond_out, int *__restrict__ hits_out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npix || !kept[q]) return;
    const unsigned d = newidx[q];
    for (int i = 0; i < 9; i++) inv_out[9 * (int64_t) d + i] = inv_in[9 * (int64_t) q + i];
    cond_out[d] = cond_in[q];
    hits_out[d] = hits_in[q];
}

__global__ void k_kept_u(const unsigned *__restrict__ kept, int npix, unsigned *__restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < npix) out[q] = kept[q];
    if (q == npix) out[q] = 0u;
}

int precond_build(Problem &pb, double *d_signal) {
    Pointing &P = pb.pt;
    Precond &C = pb.pc;
    ToeplitzPlan &T = pb.tp;
    hipStream_t s = pb.stream;
    MGPU_ARG(!C.built, "the preconditioner is already built for this problem");
    const int npix = P.npix;
    double *d_inv, *d_cond;
    int *d_hits;
    unsigned *d_kept;
    MGPU_RET(pb.arena.alloc(&d_inv, (size_t) npix * 9));
    MGPU_RET(pb.arena.alloc(&d_cond, (size_t) npix));
    MGPU_RET(pb.arena.alloc(&d_hits, (size_t) npix));
    MGPU_RET(pb.arena.alloc(&d_kept, (size_t) npix + 1));
    if (P.unit_w0)
        k_blocks<false><<<grid_for(npix, 256), 256, 0, s>>>(P.d_ell_base, P.d_ct, nullptr, P.d_cw12, npix,
                                                            T.d_start, T.nb, T.d_T0, nullptr, d_inv, d_cond, d_hits,
                                                            d_kept);
    else
        k_blocks<true><<<grid_for(npix, 256), 256, 0, s>>>(P.d_ell_base, P.d_ct, P.d_cw0, P.d_cw12, npix,
                                                           T.d_start, T.nb, T.d_T0, nullptr, d_inv, d_cond, d_hits,
                                                           d_kept);
    MGPU_LAUNCHED();
    // count kept pixels
    unsigned *d_ku, *d_new;
    MGPU_RET(pb.arena.alloc(&d_ku, (size_t) npix + 1));
    MGPU_RET(pb.arena.alloc(&d_new, (size_t) npix + 1));
    k_kept_u<<<grid_for(npix + 1, 256), 256, 0, s>>>(d_kept, npix, d_ku);
    MGPU_LAUNCHED();
    size_t tb = 0;
    MGPU_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, d_ku, d_new, npix + 1, s));
    MGPU_RET(pb.cub_tmp.ensure(tb));
    MGPU_HIP(hipcub::DeviceScan::ExclusiveSum(pb.cub_tmp.p, tb, d_ku, d_new, npix + 1, s));
    unsigned nkept = 0;
    MGPU_HIP(hipMemcpyAsync(&nkept, d_new + npix, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipStreamSynchronize(s));
    C.n_cut = npix - (int) nkept;
    MGPU_ARG(nkept > 0, "every pixel was cut by the rcond <= 0.1 test (precond.c:560)");
    if (C.n_cut > 0) {
        k_cut_samples<<<grid_for(P.m, 256), 256, 0, s>>>(P.d_q, P.m, d_kept, d_signal);
        MGPU_LAUNCHED();
        double *d_inv2, *d_cond2;
        int *d_hits2;
        MGPU_RET(pb.arena.alloc(&d_inv2, (size_t) nkept * 9));
        MGPU_RET(pb.arena.alloc(&d_cond2, (size_t) nkept));
        MGPU_RET(pb.arena.alloc(&d_hits2, (size_t) nkept));
        k_compact_precond<<<grid_for(npix, 256), 256, 0, s>>>(d_kept, d_new, npix, d_inv, d_cond, d_hits, d_inv2,
                                                              d_cond2, d_hits2);
        MGPU_LAUNCHED();
        MGPU_RET(pointing_compact(pb, d_kept));
        P.trash = 1;
        pb.arena.release(d_inv);
        pb.arena.release(d_cond);
        pb.arena.release(d_hits);
        d_inv = d_inv2;
        d_cond = d_cond2;
        d_hits = d_hits2;
    }
    MGPU_HIP(hipStreamSynchronize(s));
    pb.arena.release(d_ku);
    pb.arena.release(d_new);
    pb.arena.release(d_kept);
    C.d_inv = d_inv;
    C.d_cond = d_cond;
    C.d_hits = d_hits;
    C.built = true;
    return MGPU_OK;
}

// Cg = BJ_inv g with the reference's summation order ((0 + a0 g0) + a1 g1) + a2 g2
__global__ void k_bj_apply(const double *__restrict__ inv, const double *__restrict__ g, double *__restrict__ Cg,
                           int npix) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npix) return;
    const double g0 = g[3 * q], g1 = g[3 * q + 1], g2 = g[3 * q + 2];
    const double *a = inv + 9 * (int64_t) q;
    for (int r = 0; r < 3; r++) {
        double y = 0.0;
        y += a[3 * r] * g0;
        y += a[3 * r + 1] * g1;
        y += a[3 * r + 2] * g2;
        Cg[3 * q + r] = y;
    }
}

int launch_precond_apply(Problem &pb, const double *d_g, double *d_Cg) {
    MGPU_ARG(pb.pc.built, "precond_apply before precond_build");
    k_bj_apply<<<grid_for(pb.pt.npix, 256), 256, 0, pb.stream>>>(pb.pc.d_inv, d_g, d_Cg, pb.pt.npix);
    MGPU_LAUNCHED();
    return MGPU_OK;
}

}  // namespace mgpu
