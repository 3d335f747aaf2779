/*
 * mappraiser_gpu.h -- C-ABI of libmappraiser_gpu.so, the MI355X (gfx950) GLS
 * map-making hot path of MAPPRAISER/MIDAPACK for one rank's data on one GPU.
 *
 * The library replaces everything to the right of the PCG_GLS_true call in
 * MLmap (mappraiser/src/mappraiser.c:210-211): the pointing products, the
 * block-Toeplitz inverse-noise product, the block-Jacobi preconditioner with
 * its degenerate-pixel cut, and the PCG iteration.  All entry points
 *   - return an int status (MGPU_OK = 0) and never call exit();
 *   - copy from / into caller buffers and never keep pointers to them;
 *   - are stream-ordered on the handle's own HIP stream: calls on different
 *     handles may run concurrently; calls on one handle must not overlap.
 * Plain C types only: int / int32 indices, double data, as the reference's
 * ctypes wrapper passes them (mappraiser_wrapper/__init__.py:23-28, 58-84).
 *
 * Pointing layout: m samples x nnz=3 entries, pix[3t+k] = 3p+k (global pixel
 * p, mappraiser.py:1214-1216); flagged samples carry -3,-2,-1
 * (utils.py:148-150).  Toeplitz blocks are contiguous and block b uses
 * invtt[b*lambda .. b*lambda+lambda) (create_toeplitz.c:27-47).
 */
#ifndef MAPPRAISER_GPU_H
#define MAPPRAISER_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define MGPU_OK 0
#define MGPU_EINVAL 1  /* invalid or unsupported input (message in last_error) */
#define MGPU_EHIP 2    /* HIP runtime error */
#define MGPU_ENOMEM 3  /* device allocation failed */
#define MGPU_ESTATE 4  /* call out of order (e.g. precond_apply before precond_build) */

/* ---- precision modes ---- */
#define MGPU_FP64 0    /* every kernel in FP64 (parity mode) */
#define MGPU_MIXED 1   /* FP32 Toeplitz FFT + P/Pt, FP64 dots, 3x3 blocks and periodic FP64 residual */

typedef struct mgpu_problem mgpu_problem;

typedef struct mgpu_options {
    int precision;      /* MGPU_FP64 or MGPU_MIXED */
    int refresh_every;  /* MGPU_MIXED: recompute the FP64 residual every k iterations (default 10);
                           MGPU_REFRESH_ADAPTIVE: when a segment's FP32-updated residual norm has
                           fallen 10x below its last FP64 value or 64 iterations after it, at
                           the start of the next poll block */
    int tile_log2;      /* real FFT tile length 2^k of the Toeplitz product; 0 = autotuned choice */
    int poll_every;     /* PCG: iterations launched between host convergence polls (default 8) */
    int profile;        /* 1: time every kernel class with HIP events on the handle's stream */
    int offset_mode;    /* per-block offset handling of N^-1 (MGPU_OFFSET_*), default REFERENCE */
    int tile_variant;   /* Toeplitz tile kernel variant of the precision of the PCG iterations
                           (FP32 under MGPU_MIXED): 0 = the fixed default of the tile length
                           (results depend on the inputs only), k > 0 forces variant k,
                           MGPU_TILE_AUTOTUNE times every variant on the first product (once
                           per process, device, precision and tile length) and keeps the
                           fastest; variants round differently, so autotuned FP64 results
                           may differ between processes in the last bits.  FP64 products of
                           a mixed solve always use the FP64 default.  See
                           mappraiser_gpu_tile_variants */
    int graph;          /* 1: capture each block of PCG iterations once into a HIP graph and
                           replay it (same kernels, bit-identical results; ignored when
                           profile = 1) */
    int precond;        /* PCG_GLS_true's precond (pcg_true.c:52, build_precond precond.c:1310-1398):
                           0 block Jacobi, 1 two-level a priori (build_Z), 2 two-level
                           a posteriori (Lanczos_eig); two-level solves one problem per
                           call (not the batch / server entry points); with lambda = 1 its
                           products remove the block means as stbmmProd does */
    int z_2lvl;         /* Z_2lvl: deflation vectors of the two-level preconditioners
                           (0 -> 1, the one-rank communicator size, pcg_true.c:50; at most 256) */
} mgpu_options;

/* Per-block offset handling in N^-1 (block b: C = I - 1 1^T / n_b, T = Toeplitz band):
 *   MGPU_OFFSET_REFERENCE  N^-1 = T C, the reference's stbmmProd (toeplitz_wizard.c:86-95):
 *                          mean removal before the product.  Not symmetric: on workloads
 *                          with steep 1/f noise PCG can stagnate or break down.
 *   MGPU_OFFSET_SYMMETRIC  N^-1 = C T C: the same offset marginalisation as a symmetric
 *                          projection; equals T C except within lambda of block edges.
 *   MGPU_OFFSET_NONE       N^-1 = T (no offset removal).
 * lambda == 1 (white noise) never removes offsets, as in apply_weights (pcg_true.c:583-593). */
#define MGPU_TILE_AUTOTUNE (-1)  /* mgpu_options.tile_variant: opt into timing-based variant choice */
#define MGPU_REFRESH_ADAPTIVE (-1)  /* mgpu_options.refresh_every: residual refreshes on demand */

#define MGPU_OFFSET_REFERENCE 0
#define MGPU_OFFSET_SYMMETRIC 1
#define MGPU_OFFSET_NONE 2

/* kernel classes reported by mappraiser_gpu_kernel_times */
#define MGPU_KT_P 0        /* P (MatVecProd) with the fused per-block sums */
#define MGPU_KT_MEAN 1     /* per-block mean of N^-1 */
#define MGPU_KT_TILES 2    /* Toeplitz FFT tiles of N^-1 */
#define MGPU_KT_PT 3       /* P^T (TrMatVecProd) with the fused dot */
#define MGPU_KT_MAP 4      /* PCG map-domain updates, dots and preconditioner */
#define MGPU_KT_REFRESH 5  /* FP64 residual recomputation (MGPU_MIXED) */
#define MGPU_KT_TMPL 6     /* template marginalisation (T^T, Binv, T) */
#define MGPU_NKT 7

typedef struct mgpu_report {
    int    iterations;  /* value of the loop counter k when PCG stopped (maxiter if not converged) */
    int    converged;   /* 1 if ||g_k||^2 <= tol^2 ||g_0||^2 (pcg_true.c:200) */
    double res0;        /* ||g_0||^2 */
    double res;         /* final ||g_k||^2 */
    int    n_map;       /* map size after the pixel cut (3 x kept pixels) */
    int    lcount;      /* local index count incl. trash entries (Mat.lcount after the rebuild) */
    int    trash_pix;   /* Mat.trash_pix after the rebuild */
    int    n_cut;       /* pixels removed by the rcond <= 0.1 cut (precond.c:560) */
    double t_setup_ms;  /* device time: pointing build + Toeplitz plan (problem constructor) and
                           the preconditioner build with its cut (uploads excluded) */
    double t_solve_ms;  /* device time: PCG iterations */
    double true_res_rel;/* ||b - A x||/||b - A x0|| recomputed in FP64 at exit (MGPU_MIXED), else -1 */
    int    refreshes;   /* FP64 residual recomputations performed (MGPU_MIXED) */
    int    tile_variant;  /* tile kernel variant used by the solve's Toeplitz product (0: none) */
    double t_2lvl_ms;   /* device time building the two-level space Z, A Z, pinv(E) (precond 1, 2) */
    long long dev_peak_bytes; /* device memory the solve held at its high-water mark (inputs staged on
                                 the device, pointing tables, Toeplitz plan, preconditioner, work
                                 vectors); for a batch item, that of the whole batch it ran in */
} mgpu_report;

/* ---- library ---- */
/* Select the HIP device for subsequent handles.  No reference counterpart
 * (MPI_Init plays this role for the CPU path). */
int mappraiser_gpu_init(int device);
int mappraiser_gpu_finalize(void);
/* Message of the last failing call on this thread ("" if none). */
const char *mappraiser_gpu_last_error(void);
void mappraiser_gpu_options_default(mgpu_options *opt);

/* ---- problem (replaces MatInit + defineBlocks_avg/defineTpltz_avg,
 *      mappraiser.c:93-195, mapmat.c:54-73, create_toeplitz.c:10-47) ---- */
/* Host inputs: pix int32[m*nnz] global indices, pixweights f64[m*nnz],
 * local_blocks_sizes int32[nb], invtt f64[nb*lambda].  opt may be NULL. */
int mappraiser_gpu_problem_create(mgpu_problem **out, int m, int nnz, const int *pix,
                                  const double *pixweights, int nb, const int *local_blocks_sizes,
                                  int lambda, const double *invtt, const mgpu_options *opt);
/* Same inputs already resident in device memory (d_pix, d_pixweights,
 * d_invtt are device pointers; local_blocks_sizes is a host array). */
int mappraiser_gpu_problem_create_device(mgpu_problem **out, int m, int nnz, const int *d_pix,
                                         const double *d_pixweights, int nb,
                                         const int *local_blocks_sizes, int lambda,
                                         const double *d_invtt, const mgpu_options *opt);
int mappraiser_gpu_problem_destroy(mgpu_problem *pb);

/* MatLocalShape results (mapmat.c:496-521): lcount, trash flag, sorted unique
 * indices (lindices, cap lcount) and the local indices (cap m*nnz).  Either
 * array may be NULL.  After precond_build these describe the rebuilt matrix
 * (precond.c:659-670). */
int mappraiser_gpu_local_shape(mgpu_problem *pb, int *lcount, int *trash_pix, int *lindices,
                               int *indices);

/* ---- operators on host buffers (test and integration entry points) ---- */
/* y[m] = P x (MatVecProd, mapmat.c:704-729); x has lcount - 3*trash_pix entries. */
int mappraiser_gpu_matvec(mgpu_problem *pb, const double *x, double *y);
/* x = P^T y (TrMatVecProd, mapmat.c:819-859; the MPI reduction is the identity). */
int mappraiser_gpu_trmatvec(mgpu_problem *pb, const double *y, double *x);
/* tod <- N^-1 tod in place (apply_weights, pcg_true.c:578-600 -> stbmmProd,
 * toeplitz_wizard.c:80-111). */
int mappraiser_gpu_apply_weights(mgpu_problem *pb, double *tod);
/* d_out <- N^-1 d_in on device vectors of m doubles (FP64), the product of
 * mappraiser_gpu_apply_weights without any allocation: no reference
 * counterpart; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  The product's scratch (per-chunk
 * partial sums, block means, output-mean slices) lives in the library's
 * static device array of 2^20 doubles (8 MiB per GPU, held until the process
 * ends): problems needing more (nchunks + 10 nb doubles, nchunks = ceil(m / 1024))
 * are rejected with MGPU_EINVAL.  Launched on `stream` (a hipStream_t; NULL:
 * the handle's stream) and not synchronised.  Calls that overlap on one GPU
 * share that array and race.  A handle set to MGPU_TILE_AUTOTUNE that has not
 * yet chosen its tile variant uses the fixed default. */
int mappraiser_gpu_apply_weights_noalloc(mgpu_problem *pb, const double *d_in, double *d_out, void *stream);

/* Block-Jacobi preconditioner with the degenerate-pixel cut
 * (precondblockjacobilike, precond.c:456-702; build_precond precond=0,
 * precond.c:1310-1340).  signal[m] is modified in place like the reference
 * (cut samples set to 0).  cond/lhits receive n_map/3 entries (cap: pixels
 * before the cut); n_map receives the map size after the cut.  Any output
 * pointer may be NULL. */
int mappraiser_gpu_precond_build(mgpu_problem *pb, double *signal, double *cond, int *lhits,
                                 int *n_map);
/* Cg = BJ_inv g (apply_precond, BJ branch, precond.c:1403-1409). */
int mappraiser_gpu_precond_apply(mgpu_problem *pb, const double *g, double *Cg);

/* PCG_GLS_true iteration (pcg_true.c:20-241) after precond_build: x0 = 0,
 * rhs = P^T N^-1 (signal + noise), Polak-Ribiere PCG.  Writes x[n_map] and,
 * when res_hist is not NULL, sqrt(res_k)/sqrt(res_0) for k = 0.. (cap maxiter),
 * like the pcg_residuals_<ref>.dat file (pcg_true.c:116-125,195).
 * stop_after > 0 stops after that many iterations regardless of tol. */
int mappraiser_gpu_pcg(mgpu_problem *pb, const double *signal, const double *noise, double tol,
                       int maxiter, int stop_after, double *x, double *res_hist, mgpu_report *rep);

/* Device-resident variant of precond_build + pcg for callers that keep TOD in
 * HBM (signal and noise are device pointers of m doubles; d_signal is
 * modified in place on cut samples).  x/res_hist/cond/lhits are host. */
int mappraiser_gpu_solve_device(mgpu_problem *pb, double *d_signal, const double *d_noise,
                                double tol, int maxiter, double *x, double *res_hist,
                                double *cond, int *lhits, mgpu_report *rep);

/* Kernel timing of a handle created with opt.profile = 1: accumulated device
 * milliseconds and launch counts per kernel class (MGPU_KT_*), measured with
 * HIP events on the handle's stream; reset != 0 clears them afterwards.  With
 * pb == NULL: the totals over every drop-in solve (mappraiser_gpu_pcg_gls,
 * MLmap, MTmap) and batch (mappraiser_gpu_pcg_batch) run with opt.profile = 1
 * in this process.  No reference
 * counterpart (the reference prints MPI_Wtime splits,
 * pcg_true.c:27,98-99,178-179). */
int mappraiser_gpu_kernel_times(mgpu_problem *pb, double *ms, long long *counts, int reset);

/* Tile kernel variants of the Toeplitz product for a precision (MGPU_FP64 /
 * MGPU_MIXED) and real tile length 2^tile_log2: *count = number of variants
 * (0 if the length has no tile kernel); if desc != NULL, a text description of
 * variant `variant` (1-based) is written (at most desc_len bytes, terminated).
 * No reference counterpart (the reference's FFT is FFTW's, toeplitz.c:418-423). */
int mappraiser_gpu_tile_variants(int precision, int tile_log2, int variant, int *count, char *desc, int desc_len);

/* ---- noise-model builder ----
 * compute_invtt / noise2invtt (TOAST_interface/utils.py:435-601) on the
 * device: for every Toeplitz block of the noise timestream, a one-segment
 * Welch periodogram (Hann window, linear detrend), a bounded fit of the
 * log-PSD model a + log10(1 + ((f + fmin)/fknee)^alpha) (the reference's
 * p0 and bounds, utils.py:530-538), the inverse PSD of the fit, its inverse
 * FFT over the block length, and the apodisation window's first lambda
 * points (utils.py:562-577).  noise: [sum of sizes]; invtt: [nb*lambda]; both
 * host memory, or device memory when on_device != 0.  fit_params: host
 * [nb*4] (a, alpha, fknee, fmin) or NULL.  The reference fits with scipy's
 * trust-region-reflective least squares, this build with a bounded
 * Levenberg-Marquardt iteration: the minimiser agrees to the fit tolerance. */
#define MGPU_APOD_CHEBWIN 0   /* ifftshift(chebwin(2 lambda, 150 dB)) (the reference default) */
#define MGPU_APOD_GAUSSIAN 1  /* general Gaussian, p = 1, sigma = lambda / 3 */
int mappraiser_gpu_noise_invtt(const double *noise, int nb, const int *local_blocks_sizes, int lambda, double fsamp,
                               int apod_window, int on_device, double *invtt, double *fit_params);

/* ---- batched solves ----
 * One item is one independent GLS problem (a Monte Carlo noise realization, a
 * detector subset, a serving request) with its own pointing, signal, noise,
 * Toeplitz blocks (their sizes must sum to m) and lambda.  All items are
 * solved together as one segmented problem, whatever their bands and sizes:
 * every kernel of a PCG iteration is launched once for all of them (the
 * Toeplitz tiles once per (lambda, tile length) class, each block keeping the
 * tile length its item would use alone), each item keeps its own PCG scalars
 * and stop test, and converged items are frozen while the others go on.
 * Each item's map, iteration count and residuals are bit-identical to its
 * solo solve with the same options (with MGPU_TILE_AUTOTUNE: when both runs
 * pick the same tile variants).  Items with diagonal noise (lambda = 1), and
 * items whose band is too wide for a tile, form problems of their own, solved
 * after the others.  No reference counterpart: the reference solves one
 * problem per MLmap call (mappraiser.c:34-232). */
typedef struct mgpu_batch_item {
    /* inputs: host memory, or device memory when inputs_on_device != 0 */
    int m;                          /* samples */
    const int *pix;                 /* [3m] global indices 3p+k (0 <= p < grid_npix), flagged -3,-2,-1 */
    const double *pixweights;       /* [3m] */
    const double *signal;           /* [m] (read only: the pixel cut acts on an internal copy) */
    const double *noise;            /* [m] */
    int nb;                         /* Toeplitz blocks */
    const int *local_blocks_sizes;  /* [nb], host memory, summing to m */
    int lambda;
    const double *invtt;            /* [nb*lambda] */
    /* outputs (host memory; NULL pointers are skipped) */
    double *x;                      /* [x_cap] map of the kept pixels, (I,Q,U) per pixel, pixel order */
    int *pixels;                    /* [x_cap/3] global pixel p of each map entry */
    double *cond;                   /* [x_cap/3] rcond of each kept pixel */
    int *lhits;                     /* [x_cap/3] hits of each kept pixel */
    int x_cap;                      /* capacity of x in doubles (>= 3 * pixels kept) */
    mgpu_report rep;                /* per-item report (t_setup_ms/t_solve_ms: the item's group) */
} mgpu_batch_item;

int mappraiser_gpu_pcg_batch(int nitems, mgpu_batch_item *items, int grid_npix, int inputs_on_device, double tol,
                             int maxiter, const mgpu_options *opt);

/* ---- serving: request loop with HBM-budget admission and continuous batching ----
 * Requests are batch items (host arrays, owned by the caller until
 * server_wait returns for their ticket).  A dispatcher admits the oldest
 * queued request as soon as a slot is free and its estimated device footprint
 * fits what the running solves leave of the HBM budget, together with every
 * later queued request that fits, up to max_batch, whatever their sizes,
 * bands, grids, tolerances and iteration caps; each admitted group runs as
 * one batched solve (as mappraiser_gpu_pcg_batch, each request with its own
 * grid, tol and maxiter) on a stream and host thread of its own (opt->graph
 * is ignored).  One group runs on the device at a time: max_slots must be 1
 * (other values are refused with MGPU_EINVAL; with several concurrent groups
 * the loop stopped answering under sustained load, cause not found).  A request is answered
 * as soon as its own map has converged, while the rest of its group iterates
 * on (its report then carries no solve time and, in MGPU_MIXED, no FP64 true
 * residual).  Results are bit-identical to the requests' solo solves.
 * At creation the server reserves 80% of the free device memory (after
 * returning the library's cached blocks) as one pool that all its solves
 * allocate from, so serving makes no device allocation or free through the
 * HIP runtime; create it while no other solve of the process runs on the
 * device.  The admission budget is that pool, or hbm_budget_gb when smaller
 * (<= 0: the pool).  Footprints are
 * estimated per sample and per grid pixel and, once groups have finished,
 * raised by 1.05 x the largest ratio of measured to estimated bytes seen when
 * that exceeds 1.  A group whose device
 * allocations fail while other solves run, or that merged several requests,
 * returns to the queue head; its requests then run one at a time on an idle
 * device (a failure there is reported).  A merged group stays within 2^30
 * samples and 2^28 grid pixels.  No reference counterpart. */
typedef struct mgpu_server mgpu_server;

int mappraiser_gpu_server_create(mgpu_server **out, double hbm_budget_gb, int max_slots, int max_batch,
                                 const mgpu_options *opt);
int mappraiser_gpu_server_submit(mgpu_server *s, mgpu_batch_item *item, int grid_npix, double tol, int maxiter,
                                 long long *ticket);
/* blocks until the request is solved; latency = done - submitted, queue = admitted - submitted */
int mappraiser_gpu_server_wait(mgpu_server *s, long long ticket, double *latency_ms, double *queue_ms);
/* admitted: admissions (a requeued request counts again); groups: batched solves started;
 * peak_gb: largest budget charge; requeued: requests sent back after an allocation failure */
int mappraiser_gpu_server_stats(mgpu_server *s, long long *admitted, long long *groups, double *peak_gb,
                                int *peak_running, long long *requeued);
/* finishes every queued and running request, then frees the server */
int mappraiser_gpu_server_destroy(mgpu_server *s);

/* Drop-in for PCG_GLS_true(outpath, ref, &A, &Nm1, x, signal, noise, cond,
 * lhits, tol, maxiter, precond=0, Z_2lvl) at mappraiser.c:211 for one rank,
 * taking MLmap's raw arguments.  Side effects mirror the reference:
 *   pix        -> local indices of the rebuilt matrix (A.indices, precond.c:659-670)
 *   pixweights -> zeroed on cut samples (precond.c:601,613)
 *   signal     -> zeroed on cut samples (precond.c:605,610)
 *   x, cond, lhits -> first n_map, n_map/3, n_map/3 entries (cap: pixels before the cut)
 *   lindices   -> rebuilt A.lindices (cap m*3), lcount/trash_pix in rep.
 * Writes <outpath>/pcg_residuals_<ref>.dat like pcg_true.c:116-125 when outpath
 * and ref are not NULL. */
int mappraiser_gpu_pcg_gls(const char *outpath, const char *ref, int m, int nnz, int *pix,
                           double *pixweights, double *signal, const double *noise, int nb,
                           const int *local_blocks_sizes, int lambda, const double *invtt,
                           double tol, int maxiter, double *x, double *cond, int *lhits,
                           int *lindices, const mgpu_options *opt, mgpu_report *rep);

/* ---- MLmap and its outputs ---- */
/* data type codes of mappraiser_gpu_write_map (cfitsio's TDOUBLE / TINT values,
 * so the reference's write_map calls carry over unchanged) */
#define MGPU_TDOUBLE 82
#define MGPU_TINT 31

/* x2map_pol (mappraiser.c:836-853): scatter the local map x[xsize] (I,Q,U per
 * pixel, global indices lstid = 3p+k) into full maps mapI/mapQ/mapU, Cond and
 * hits (12 nside^2 entries each).  lstid[i]/3 must index those maps. */
int mappraiser_gpu_x2map_pol(double *mapI, double *mapQ, double *mapU, double *Cond, int *hits,
                             const double *x, const int *lstid, const double *cond, const int *lhits,
                             int xsize);

/* write_map (iofiles.c:19-53): a HEALPix map of 12 nside^2 values (type
 * MGPU_TDOUBLE or MGPU_TINT) as a FITS binary table with one 8-byte float
 * column "SIGNAL", ORDERING NESTED if nest else RING, COORDSYS from coordsys[0].
 * Written without cfitsio; byte-identical to the reference's files apart from
 * the DATE value.  An existing file is overwritten. */
int mappraiser_gpu_write_map(const void *signal, int type, long nside, const char *filename, char nest,
                             const char *coordsys);

/* MLmap (mappraiser.c:34-380, declared in map.h:10-13) for ONE rank on one GPU:
 * the same arguments, so the reference's ctypes wrapper
 * (mappraiser_wrapper/__init__.py:58-84) binds it unchanged.  comm is not read
 * (call it from a one-rank communicator; MPICH's int and Open MPI's pointer
 * handles both arrive in this first integer argument); m = data_size_proc[0].
 * solver must be 0 (PCG); precond / Z_2lvl select block Jacobi or the
 * two-level preconditioners.  The solve is mappraiser_gpu_pcg_gls with default
 * options (FP64, reference offsets; the environment variables
 * MGPU_PRECISION=mixed and MGPU_OFFSET_MODE=symmetric|none override them),
 * with the same side effects on pix, pixweights and signal and the residual
 * file <outpath>/pcg_residuals_<ref>.dat; then x2map_pol and the five FITS
 * maps <outpath>/{mapI,mapQ,mapU,Cond,Hits}_<ref>.fits (old files removed
 * first, mappraiser.c:297-355).  void like the reference: failures are
 * printed to stderr and left in mappraiser_gpu_last_error (no exit). */
void MLmap(intptr_t comm, char *outpath, char *ref, int solver, int precond, int Z_2lvl, int pointing_commflag,
           double tol, int maxiter, int enlFac, int ortho_alg, int bs_red, int nside, void *data_size_proc,
           int nb_blocks_loc, void *local_blocks_sizes, int Nnz, void *pix, void *pixweights, void *signal,
           double *noise, int lambda, double *invtt);

/* ---- template marginalisation (MTmap, mappraiser.c:387-700; PCG_GLS_templates,
 *      pcg_true.c:243-576; algebra/src/templates/templates.c) ---- */
typedef struct mgpu_templates {
    int npoly;              /* Legendre orders 0..npoly-1 on every sweep (Polyinit) */
    int ground;             /* 1: a scan-synchronous class of n_sss_bins azimuth bins (SSSinit) */
    int n_sss_bins;
    int nhwp;               /* HWP harmonics 1..nhwp, cos and sin, per time bin of delta_t s (HWPSSinit) */
    double delta_t;
    double sampling_freq;
    int nces;               /* the problem's nb blocks are nces x ndet, CES-major, each as long as its CES */
    int *const *sweeptstamps;  /* per CES: nsweeps[c]+1 non-decreasing stamps, the last = the CES length */
    const int *nsweeps;
    double *const *az;      /* per CES: boresight azimuth per sample (ground) */
    const double *az_min, *az_max;
    double *const *hwp_angle;  /* per CES: HWP angle per sample (nhwp) */
} mgpu_templates;

/* Attach template classes to a problem created with lambda = 1 (MTmap's
 * white-noise model: invtt[b] is block b's weight, create_toeplitz.c:27-47 with
 * lambda_block_avg = 1).  Builds the tables of Tlist_init/bin_az
 * (templates.c:560-757), the kernels B_b = w_b T_b^T T_b (BuildKernel,
 * templates.c:191-436) and their pseudo-inverses (InvKernel, templates.c:449-465:
 * singular values <= 1e-12 s_max dropped).  Later pcg / solve_device calls
 * marginalise over the templates (PCG_GLS_templates) in FP64; MGPU_MIXED is
 * refused.  A second call replaces the classes. */
int mappraiser_gpu_templates_set(mgpu_problem *pb, const mgpu_templates *t);
/* Template count, per-block counts nt_block[nb] and kernel ranks[nb] (the
 * effective ranks InvKernel prints); any pointer may be NULL. */
int mappraiser_gpu_templates_info(mgpu_problem *pb, long long *ntot, int *nt_block, int *ranks);
#define MGPU_TMPL_T 0       /* out[m] = T in[ntot]           (TVecProd, templates.c:27-103) */
#define MGPU_TMPL_TT 1      /* out[ntot] = T^T in[m]         (TrTVecProd, templates.c:106-188) */
#define MGPU_TMPL_FILTER 2  /* out[m] = in - W T Binv T^T in (pcg_true.c:426-458; in already weighted by W) */
int mappraiser_gpu_templates_apply(mgpu_problem *pb, int op, const double *in, double *out);
/* The kernels and their pseudo-inverses as row-major nt_b x nt_b blocks,
 * concatenated in block order (sum of nt_b^2 entries each); either may be NULL. */
int mappraiser_gpu_templates_kernel(mgpu_problem *pb, double *B, double *Binv);

/* MTmap (mappraiser.c:387-700, map.h:15-18) with the reference's signature, so
 * its ctypes wrapper (mappraiser_wrapper/__init__.py:86-122) binds it
 * unchanged.  One rank as for MLmap: m = data_size_proc[0], the
 * nb_blocks_loc blocks are nces x (nb_blocks_loc / nces) CES-major and each
 * must be as long as its CES (sweeptstamps[c][nsweeps[c]]).  The noise model is
 * MTmap's: lambda 1 with invtt[b] per block.  The solve is PCG_GLS_templates
 * (mappraiser_gpu_templates_set on the problem, then the drop-in solve), with
 * the side effects, residual file and FITS maps of MLmap.  void like the
 * reference: failures are printed to stderr and left in
 * mappraiser_gpu_last_error (no exit). */
void MTmap(intptr_t comm, char *outpath, char *ref, int solver, int precond, int Z_2lvl, int pointing_commflag,
           int npoly, int nhwp, double delta_t, int ground, int n_sss_bins, double tol, int maxiter, int enlFac,
           int ortho_alg, int bs_red, int nside, int **sweeptstamps, int *nsweeps, double **az, double *az_min,
           double *az_max, double **hwp_angle, int nces, void *data_size_proc, int nb_blocks_loc,
           void *local_blocks_sizes, int Nnz, void *pix, void *pixweights, void *signal, double *noise,
           double sampling_freq, double *invtt);

#ifdef __cplusplus
}
#endif
#endif /* MAPPRAISER_GPU_H */
