/*
 * oracle.h -- CPU restatement of the MAPPRAISER/MIDAPACK GLS map-making path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product path (libmappraiser_gpu.so) never
 * links, loads or calls it.
 *
 * Each function restates one reference routine on ONE rank (every MPI
 * reduction on the path is an identity there).  Citations are to files of the
 * reference tree (B3Dcmb/midapack).  Arithmetic is plain IEEE double with
 * floating-point contraction disabled (-ffp-contract=off), matching a build of
 * the reference without FMA.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   orc_local_shape, orc_matvec, orc_trmatvec  -- bit-exact against the reference
 *       mapmat.c/csort.c/cindex.c compiled from source (oracle/_ref) and against
 *       numpy unique/searchsorted brute force.
 *   orc_stbmm / orc_apply_weights -- FFT shift-and-overlap path equals the direct
 *       banded product (closed form, mean removal + zero-boundary band) to 1e-13;
 *       special cases lambda=1, T=delta, constant input.
 *   orc_fft -- against numpy.fft.
 *   orc_dgetrf / orc_dgecon1 / orc_dgetri / orc_dlange1 -- against scipy's LAPACK
 *       (dgetrf/dgecon/dgetri/dlange) on random and degenerate 3x3 blocks, and the
 *       inverse against the adjugate closed form.
 *   orc_getlocalW -- closed form sum_t w w^T T0.
 *   orc_gls_solve -- tiny systems against a dense LU solve of the GLS normal
 *       equations (with mean-removal N^-1), and x_k against a dense-operator PCG.
 *   orc_pinv / orc_symeig -- against numpy's LAPACK pinv, lstsq and eigh
 *       (full-rank, rank-deficient, badly scaled), Penrose conditions.
 *   orc_tmpl_* / orc_gls_solve_tmpl (templates.c, tests/test_templates_oracle.py) --
 *       T, T^T and the kernels against a dense numpy T from the definitions, one
 *       kernel entry against its ordered sum, InvKernel against the Penrose
 *       conditions and numpy's pinv/rank, the filter against the dense projector,
 *       the template PCG against the dense normal equations and template invariance.
 *   orc_gls_solve2 (two-level, tests/test_twolvl_oracle.py) -- build_Z column 0
 *       against brute-force hit fractions and the other columns through the
 *       reference's own index; A Z against the pinned products; the Lanczos Z
 *       inside the Krylov space of A from P^T N^-1 d (and not inside a smaller
 *       one); Em1 against numpy's pinv of Z^T-products; the deflated PCG
 *       reaches the block-Jacobi solution.
 */
#ifndef MAPPRAISER_ORACLE_H
#define MAPPRAISER_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_EINVAL (-1)
#define ORC_ENOMEM (-2)

/* ---- index tables (mapmat.c:496-521, csort.c:111-135, cindex.c:34-41,102-115) ---- */
int orc_counting_sort(int *indices, int64_t count);
int orc_dichotomy(int nT, const int *T, int e);
int orc_local_shape(int64_t count, int *indices, int *lindices, int *lcount, int *trash_pix);

/* ---- pointing products (mapmat.c:704-729, 819-851) ---- */
void orc_matvec(int m, int nnz, const int *indices, const double *values, int trash_pix,
                const double *x, double *y);
void orc_trmatvec(int m, int nnz, const int *indices, const double *values, int trash_pix,
                  int lcount, const double *y, double *x);

/* ---- pixel -> samples linked list (mappraiser.c:110-145) ---- */
void orc_pixel_lists(int m, int nnz, const int *indices, int trash_pix, int npix_valid,
                     int *id_last_pix, int *ll);

/* ---- FFT used by the Toeplitz restatement (stands for FFTW's r2c/c2r) ---- */
/* in/out interleaved complex (re,im), length n, sign -1 forward / +1 backward,
 * unnormalised. Any n >= 1. */
int orc_fft(int n, const double *in, double *out, int sign);

/* ---- Toeplitz (toeplitz.c:168-253, 441-470, 569-700, 731-865; toeplitz_block.c:115-636;
 *      toeplitz_wizard.c:80-111; pcg_true.c:578-600) ---- */
int orc_define_blocksize(int n, int lambda);
/* How the Toeplitz tiles are computed (all other steps are unchanged):
 *   ORC_NINV_REFERENCE  FFT shift-and-overlap with the reference's in-place
 *                       complex multiply (toeplitz.c:587-588), whose imaginary
 *                       line reuses the updated real part: the rounding noise
 *                       of Im FFT(T_circ) enters multiplied by Re FFT(T_circ),
 *                       an error of order eps*|T|^2 (about 6e-10 relative on
 *                       the reference's own test data, T0 ~ 1e7).  Default.
 *   ORC_NINV_REAL_SPECTRUM  same tiles, multiplied by the real spectrum.
 *   ORC_NINV_DIRECT     direct banded product (the closed form).
 * Returns the previous mode. */
#define ORC_NINV_REFERENCE 0
#define ORC_NINV_REAL_SPECTRUM 1
#define ORC_NINV_DIRECT 2
int orc_set_ninv_mode(int mode);
/* Per-block offset handling of N^-1 (C = I - 1 1^T/n per block):
 *   ORC_OFFSET_REFERENCE  T C, the reference (toeplitz_wizard.c:86-95).  Default.
 *   ORC_OFFSET_SYMMETRIC  C T C, the projection applied again after the product.
 *   ORC_OFFSET_NONE       T.
 * Returns the previous mode. */
#define ORC_OFFSET_REFERENCE 0
#define ORC_OFFSET_SYMMETRIC 1
#define ORC_OFFSET_NONE 2
int orc_set_offset_mode(int mode);
/* one block, in place: banded symmetric Toeplitz product with zero boundaries.
 * use_fft=1: shift-and-overlap with the reference's tile rule; 0: direct O(n lambda). */
int orc_stmm_block(double *V, int n, const double *T, int lambda, int use_fft);
/* N^-1 on a local TOD vector: per-block mean removal then the Toeplitz product. */
int orc_stbmm(int nb, const int *block_sizes, int lambda, const double *invtt, int64_t m,
              double *V, int use_fft);
/* apply_weights: lambda==1 -> diagonal scale by T0 (no mean removal), else orc_stbmm. */
int orc_apply_weights(int nb, const int *block_sizes, int lambda, const double *invtt,
                      int64_t m, double *V, int use_fft);

/* ---- LAPACK restatement (reference LAPACK algorithms, row-major n x n, small n) ---- */
double orc_dlange1(int n, const double *a);
int orc_dgetrf(int n, double *a, int *ipiv);          /* returns LAPACK info */
double orc_dgecon1(int n, const double *lu, double anorm);
int orc_dgetri(int n, double *a, const int *ipiv);    /* returns LAPACK info */

/* ---- preconditioner (precond.c:42-161, 456-702) ---- */
void orc_getlocalW(int m, int nnz, const int *indices, const double *values, int trash_pix,
                   int lcount, int nb, const int *block_sizes, int lambda, const double *invtt,
                   double *vpixBlock, int *lhits);

typedef struct {
    int    iterations;   /* value of k when the loop ended (K if unconverged) */
    int    converged;
    double res0;         /* ||g0||^2 */
    double res;          /* final ||g||^2 */
    int    n_map;        /* map size after the pixel cut */
    int    lcount;       /* final local index count (including trash entries) */
    int    trash_pix;    /* final trash flag */
    int    n_cut;        /* number of pixels removed by the rcond cut */
} orc_report;

/* Whole one-rank drop-in: MLmap setup (mappraiser.c:64-195) + PCG_GLS_true
 * (pcg_true.c:20-241) with the block-Jacobi preconditioner (precond=0).
 * pix (global on input, local on output) and pixweights/signal are modified in
 * place exactly as the reference modifies them.  Output buffers:
 *   x_out[m*nnz cap], lindices_out[m*nnz cap], cond_out/lhits_out[m cap],
 *   res_hist[maxiter cap] (rel. residual sqrt(res)/sqrt(res0) per k, k=0..).
 * stop_after>0 stops after that many loop iterations regardless of tol (for
 * iterate-by-iterate parity tests). */
int orc_gls_solve(int m, int nnz, int *pix, double *pixweights, double *signal,
                  const double *noise, int nb, const int *block_sizes, int lambda,
                  const double *invtt, double tol, int maxiter, int stop_after,
                  double *x_out, int *lindices_out, double *cond_out, int *lhits_out,
                  double *res_hist, orc_report *rep);

/* orc_gls_solve with the preconditioner choice of PCG_GLS_true (pcg_true.c:52):
 * precond 0 = block Jacobi, 1 = two-level a priori (build_Z, precond.c:798-953),
 * 2 = two-level a posteriori (Lanczos_eig, precond.c:1073-1306); Z_2lvl = number
 * of deflation vectors (0 -> 1, the one-rank communicator size).  When not
 * NULL, z_out / az_out receive the Zn columns of Z and A Z (column j at j*n_map)
 * and em1_out the Zn x Zn row-major pseudo-inverse of E (precond.c:956-1003). */
int orc_gls_solve2(int m, int nnz, int *pix, double *pixweights, double *signal,
                   const double *noise, int nb, const int *block_sizes, int lambda,
                   const double *invtt, double tol, int maxiter, int stop_after, int precond, int Z_2lvl,
                   double *x_out, int *lindices_out, double *cond_out, int *lhits_out,
                   double *res_hist, double *z_out, double *az_out, double *em1_out, orc_report *rep);

/* ---- template marginalisation (MTmap, mappraiser.c:387-700; PCG_GLS_templates,
 *      pcg_true.c:243-576; algebra/src/templates/templates.c) -- templates.c ---- */
typedef struct {
    int npoly;          /* polynomial orders 0..npoly-1 on every sweep */
    int ground;         /* 1: one scan-synchronous class of n_sss_bins azimuth bins */
    int n_sss_bins;
    int nhwp;           /* HWP harmonics 1..nhwp (cos and sin) per time bin of delta_t */
    double delta_t, sampling_freq;
    int nces, ndet;     /* blocks c*ndet + j (CES-major), each as long as its CES */
    const int *block_sizes;     /* [nces*ndet] */
    const int *nsweeps;         /* [nces] */
    const int *sweeptstamps;    /* CES c: nsweeps[c]+1 stamps, concatenated */
    const int *az_binned;       /* per CES sample, concatenated (ground) */
    const double *hwp_cos;      /* cos / sin of the HWP angle per CES sample, concatenated (nhwp) */
    const double *hwp_sin;
    const double *w;            /* [nces*ndet] white-noise weight per block (invtt, lambda 1) */
} orc_tmpl;
/* bin_az (templates.c:736-757); ORC_EINVAL for a sample below az_min */
int orc_bin_az(int n, const double *az, double az_min, double az_max, int nbins, int *out);
/* cos and sin of the HWP angle (MTmap, mappraiser.c:586-597) */
void orc_hwp_mod(int n, const double *angle, double *c, double *s);
/* templates per block (nt_block may be NULL); returns their total or < 0 */
int orc_tmpl_sizes(const orc_tmpl *t, int *nt_block);
int orc_tmpl_tvec(const orc_tmpl *t, const double *tau, double *out);    /* TVecProd */
int orc_tmpl_trtvec(const orc_tmpl *t, const double *d, double *out);    /* TrTVecProd */
int orc_tmpl_kernel(const orc_tmpl *t, double *B);   /* BuildKernel, blocks nt_b^2 concatenated */
int orc_tmpl_invkernel(const orc_tmpl *t, const double *B, double *Binv, int *ranks); /* InvKernel */
/* tod (already weighted by W) -= W T Binv T^T tod */
int orc_tmpl_filter(const orc_tmpl *t, const double *Binv, double *tod);
/* MTmap + PCG_GLS_templates on one rank; invtt[nb] (lambda 1); ranks_out[nb]
 * (may be NULL) receives the kernel ranks InvKernel reports. */
int orc_gls_solve_tmpl(int m, int nnz, int *pix, double *pixweights, double *signal, const double *noise,
                       const int *block_sizes, const double *invtt, double tol, int maxiter, int stop_after,
                       int precond, int Z_2lvl, const orc_tmpl *tm, double *x_out, int *lindices_out,
                       double *cond_out, int *lhits_out, double *res_hist, int *ranks_out, orc_report *rep);

/* threads of the all-cores timing build (liboracle_omp.so): sets n (> 0) and
 * returns the count in use; 1 in the sequential build */
int orc_set_threads(int n);

/* pinv(a) with singular values <= rcond * s_max dropped (inverse_svd_bis,
 * precond.c:777-796), one-sided Jacobi SVD; row-major n x n. */
int orc_pinv(int n, const double *a, double rcond, double *out);
/* the same, with the number of kept singular values in *rank (may be NULL) */
int orc_pinv_rank(int n, const double *a, double rcond, double *out, int *rank);
/* symmetric eigen-decomposition (LAPACKE_dsyev row-major 'V', precond.c:1240):
 * ascending eigenvalues w, eigenvector j in column j of row-major v. */
int orc_symeig(int n, const double *a, double *w, double *v);

#ifdef __cplusplus
}
#endif
#endif
