/*
 * oracle.c -- CPU restatement of the MAPPRAISER/MIDAPACK one-rank GLS path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline).  The
 * product library never links or calls anything here.  See oracle.h for what
 * pins each function.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 */
#include "oracle.h"

/* loops that the all-cores timing build (-fopenmp) runs in parallel */
#ifdef _OPENMP
#include <omp.h>
#define ORC_PFOR _Pragma("omp parallel for schedule(static)")
#else
#define ORC_PFOR
#endif

int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void) n;
    return 1;
#endif
}

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double re, im;
} cplx;

/* ========================================================================
 * Index tables
 * ====================================================================== */

/* csort.c:111-135 -- dense bitmap over [min,max], returns the sorted unique
 * values in place and their count. */
int orc_counting_sort(int *indices, int64_t count) {
    if (count <= 0) return 0;
    int mn = indices[0], mx = indices[0];
    for (int64_t i = 1; i < count; i++) {
        if (indices[i] > mx)
            mx = indices[i];
        else if (indices[i] < mn)
            mn = indices[i];
    }
    int64_t span = (int64_t) mx - (int64_t) mn + 1;
    unsigned char *buf = (unsigned char *) calloc((size_t) span, 1);
    if (!buf) return ORC_ENOMEM;
    for (int64_t i = 0; i < count; i++) buf[(int64_t) indices[i] - mn] = 1;
    int j = 0;
    for (int64_t i = 0; i < span; i++)
        if (buf[i]) indices[j++] = (int) (mn + i);
    free(buf);
    return j;
}

/* cindex.c:102-115 -- binary search returning the pivot position. */
int orc_dichotomy(int nT, const int *T, int e) {
    int lo = 0, hi = nT - 1, piv = (hi - lo) / 2;
    while (e != T[piv] && hi > lo) {
        if (T[piv] < e)
            lo = piv + 1;
        else
            hi = piv;
        piv = lo + (hi - lo) / 2;
    }
    return piv;
}

/* mapmat.c:496-521 (MatLocalShape with ssort flag 3, csort.c:172-204) and
 * cindex.c:34-41 (sindex). indices are rewritten in place to local positions. */
int orc_local_shape(int64_t count, int *indices, int *lindices, int *lcount, int *trash_pix) {
    int *tmp = (int *) malloc((size_t) (count > 0 ? count : 1) * sizeof(int));
    if (!tmp) return ORC_ENOMEM;
    memcpy(tmp, indices, (size_t) count * sizeof(int));
    int n = orc_counting_sort(tmp, count);
    if (n < 0) {
        free(tmp);
        return n;
    }
    memcpy(lindices, tmp, (size_t) n * sizeof(int));
    free(tmp);
    for (int64_t i = 0; i < count; i++) indices[i] = orc_dichotomy(n, lindices, indices[i]);
    *lcount = n;
    if (n > 0 && lindices[0] < 0) *trash_pix = 1;
    return ORC_OK;
}

/* ========================================================================
 * Pointing products
 * ====================================================================== */

/* mapmat.c:704-729 -- y = P x; trash rows (first local index 0) give 0 and the
 * map offset is -nnz. */
void orc_matvec(int m, int nnz, const int *indices, const double *values, int trash_pix,
                const double *x, double *y) {
    for (int i = 0; i < m; i++) y[i] = 0.0;
    if (trash_pix) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (int64_t r = 0; r < m; r++) {
            const int64_t i = r * nnz;
            if (indices[i] != 0)
                for (int j = 0; j < nnz; j++) y[r] += values[i + j] * x[indices[i + j] - nnz];
        }
    } else {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (int64_t r = 0; r < m; r++) {
            const int64_t i = r * nnz;
            for (int j = 0; j < nnz; j++) y[r] += values[i + j] * x[indices[i + j]];
        }
    }
}

/* mapmat.c:819-851 -- x = P^T y (greedyreduce is an identity on one rank). */
void orc_trmatvec(int m, int nnz, const int *indices, const double *values, int trash_pix,
                  int lcount, const double *y, double *x) {
    if (trash_pix) {
        for (int i = 0; i < lcount - nnz; i++) x[i] = 0.0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (int64_t r = 0; r < m; r++) {
            const int64_t i = r * nnz;
            if (indices[i] != 0)
                for (int j = 0; j < nnz; j++) {
                    const double v = values[i + j] * y[r];
#ifdef _OPENMP
#pragma omp atomic
#endif
                    x[indices[i + j] - nnz] += v;
                }
        }
    } else {
        for (int i = 0; i < lcount; i++) x[i] = 0.0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (int64_t r = 0; r < m; r++) {
            const int64_t i = r * nnz;
            for (int j = 0; j < nnz; j++) {
                const double v = values[i + j] * y[r];
#ifdef _OPENMP
#pragma omp atomic
#endif
                x[indices[i + j]] += v;
            }
        }
    }
}

/* mappraiser.c:110-145 -- last sample of each valid pixel and the chain of
 * earlier samples. */
void orc_pixel_lists(int m, int nnz, const int *indices, int trash_pix, int npix_valid,
                     int *id_last_pix, int *ll) {
    for (int i = 0; i < m; i++) ll[i] = -1;
    for (int j = 0; j < npix_valid; j++) id_last_pix[j] = -1;
    for (int i = 0; i < m; i++) {
        if (!trash_pix || indices[(int64_t) i * nnz] != 0) {
            int v = indices[(int64_t) i * nnz] / nnz - trash_pix;
            if (id_last_pix[v] == -1) {
                id_last_pix[v] = i;
            } else {
                ll[i] = id_last_pix[v];
                id_last_pix[v] = i;
            }
        }
    }
}

/* ========================================================================
 * FFT (stands for FFTW's plans in toeplitz.c; unnormalised, any length)
 * ====================================================================== */

typedef struct {
    int   n;
    cplx *tw; /* exp(-2 pi i k / n), k < n */
} twiddle_cache;

#define TW_SLOTS 256
static twiddle_cache g_tw[TW_SLOTS];
static int g_tw_next = 0;

/* Tables stay allocated once made: with OpenMP (the all-cores timing build)
 * other threads may be reading them, so a full cache computes an uncached
 * table instead of evicting one (the caller then owns nothing: it leaks one
 * table per call, which no workload here reaches with its few FFT sizes). */
static const cplx *get_twiddles(int n) {
    const cplx *found = NULL;
#ifdef _OPENMP
#pragma omp critical(orc_twiddles)
#endif
    {
        for (int s = 0; s < TW_SLOTS && !found; s++)
            if (g_tw[s].n == n && g_tw[s].tw) found = g_tw[s].tw;
        if (!found) {
            cplx *t = (cplx *) malloc((size_t) n * sizeof(cplx));
            for (int k = 0; k < n; k++) {
                double a = -2.0 * M_PI * (double) k / (double) n;
                t[k].re = cos(a);
                t[k].im = sin(a);
            }
            if (g_tw_next < TW_SLOTS) {
                g_tw[g_tw_next].tw = t;
                g_tw[g_tw_next].n = n;
                g_tw_next++;
            }
            found = t;
        }
    }
    return found;
}

static inline cplx cmul(cplx a, cplx b) {
    cplx r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

/* twiddle exp(sign * 2 pi i e / N) from the forward table */
static inline cplx twid(const cplx *tw, int N, int64_t e, int sign) {
    cplx w = tw[e % N];
    if (sign > 0) w.im = -w.im;
    return w;
}

/* iterative radix-2 decimation in time, n a power of two, in place */
static void fft_pow2(int n, cplx *a, int sign) {
    const cplx *tw = get_twiddles(n);
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            cplx t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int j = 0; j < half; j++) {
                cplx w = twid(tw, n, (int64_t) j * step, sign);
                cplx u = a[i + j];
                cplx v = cmul(a[i + j + half], w);
                a[i + j].re = u.re + v.re;
                a[i + j].im = u.im + v.im;
                a[i + j + half].re = u.re - v.re;
                a[i + j + half].im = u.im - v.im;
            }
        }
    }
}

static int smallest_factor(int n) {
    if (n % 2 == 0) return 2;
    for (int p = 3; (int64_t) p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* recursive mixed radix: out[0..n) = DFT of in[0], in[s], in[2s], ... */
static void fft_rec(int n, const cplx *in, int64_t s, cplx *out, int sign, const cplx *tw, int N,
                    int64_t twstride) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = smallest_factor(n), mlen = n / p;
    for (int j = 0; j < p; j++) fft_rec(mlen, in + j * s, s * p, out + j * mlen, sign, tw, N, twstride * p);
    cplx *t = (cplx *) malloc((size_t) p * sizeof(cplx));
    cplx *r = (cplx *) malloc((size_t) p * sizeof(cplx));
    for (int k = 0; k < mlen; k++) {
        for (int j = 0; j < p; j++) t[j] = cmul(out[j * mlen + k], twid(tw, N, (int64_t) j * k * twstride, sign));
        for (int q = 0; q < p; q++) {
            cplx acc = {0.0, 0.0};
            for (int j = 0; j < p; j++) {
                cplx w = twid(tw, N, ((int64_t) j * q % p) * mlen * twstride, sign);
                cplx v = cmul(t[j], w);
                acc.re += v.re;
                acc.im += v.im;
            }
            r[q] = acc;
        }
        for (int q = 0; q < p; q++) out[k + q * mlen] = r[q];
    }
    free(t);
    free(r);
}

int orc_fft(int n, const double *in, double *out, int sign) {
    if (n < 1) return ORC_EINVAL;
    const cplx *src = (const cplx *) in;
    cplx *dst = (cplx *) out;
    if ((n & (n - 1)) == 0) {
        if (dst != src) memmove(dst, src, (size_t) n * sizeof(cplx));
        fft_pow2(n, dst, sign);
        return ORC_OK;
    }
    const cplx *tw = get_twiddles(n);
    cplx *tmp = (cplx *) malloc((size_t) n * sizeof(cplx));
    if (!tmp) return ORC_ENOMEM;
    fft_rec(n, src, 1, tmp, sign, tw, n, 1);
    memcpy(dst, tmp, (size_t) n * sizeof(cplx));
    free(tmp);
    return ORC_OK;
}

/* real -> half spectrum X[0..n/2] (FFTW r2c semantics), n even */
static void rfft(int n, const double *x, cplx *X, cplx *work) {
    int h = n / 2;
    for (int j = 0; j < h; j++) {
        work[j].re = x[2 * j];
        work[j].im = x[2 * j + 1];
    }
    orc_fft(h, (const double *) work, (double *) work, -1);
    const cplx *tw = get_twiddles(n);
    for (int k = 0; k <= h; k++) {
        cplx zk = work[k % h], zc = work[(h - k) % h];
        zc.im = -zc.im;
        cplx e = {0.5 * (zk.re + zc.re), 0.5 * (zk.im + zc.im)};
        cplx o = {0.5 * (zk.im - zc.im), -0.5 * (zk.re - zc.re)}; /* (zk - zc)/(2i) */
        cplx wo = cmul(tw[k % n], o);
        X[k].re = e.re + wo.re;
        X[k].im = e.im + wo.im;
    }
}

/* half spectrum -> real, unnormalised (FFTW c2r semantics: imaginary parts of
 * the DC and Nyquist bins are ignored), n even */
static void irfft(int n, const cplx *Xin, double *x, cplx *work) {
    int h = n / 2;
    const cplx *tw = get_twiddles(n);
    for (int k = 0; k < h; k++) {
        cplx xk = Xin[k], xc = Xin[h - k];
        if (k == 0) {
            xk.im = 0.0;
            xc.im = 0.0;
        }
        xc.im = -xc.im;
        cplx e = {xk.re + xc.re, xk.im + xc.im};
        cplx d = {xk.re - xc.re, xk.im - xc.im};
        cplx wi = tw[k];
        wi.im = -wi.im; /* exp(+2 pi i k/n) */
        cplx o = cmul(d, wi);
        work[k].re = e.re - o.im; /* e + i*o */
        work[k].im = e.im + o.re;
    }
    orc_fft(h, (const double *) work, (double *) work, +1);
    for (int j = 0; j < h; j++) {
        x[2 * j] = work[j].re;
        x[2 * j + 1] = work[j].im;
    }
}

/* generic (odd length) real transforms through the complex FFT */
static void rfft_any(int n, const double *x, cplx *X, cplx *work) {
    if (n % 2 == 0) {
        rfft(n, x, X, work);
        return;
    }
    for (int j = 0; j < n; j++) {
        work[j].re = x[j];
        work[j].im = 0.0;
    }
    orc_fft(n, (const double *) work, (double *) work, -1);
    for (int k = 0; k <= n / 2; k++) X[k] = work[k];
}

static void irfft_any(int n, const cplx *X, double *x, cplx *work) {
    if (n % 2 == 0) {
        irfft(n, X, x, work);
        return;
    }
    for (int k = 0; k <= n / 2; k++) work[k] = X[k];
    work[0].im = 0.0;
    for (int k = n / 2 + 1; k < n; k++) {
        work[k].re = X[n - k].re;
        work[k].im = -X[n - k].im;
    }
    orc_fft(n, (const double *) work, (double *) work, +1);
    for (int j = 0; j < n; j++) x[j] = work[j].re;
}

/* ========================================================================
 * Toeplitz
 * ====================================================================== */

/* toeplitz.c:200-206 (flag_bs 0, the auto strategy set by
 * toeplitz_params.c:88-96): bs = 2^ceil(log2(4 lambda)), or 4 lambda if that
 * exceeds n.  The expression is kept verbatim in form so that the rounding of
 * log()/log() is the same. */
int orc_define_blocksize(int n, int lambda) {
    int min_bs = 4 * lambda;
    int min_pow2 = (int) ceil(log(min_bs) / log(2));
    int bs = (int) pow(2, min_pow2);
    if (bs > n) bs = min_bs;
    return bs;
}

/* toeplitz.c:441-470 (circ_init_fftw): spectrum of the circulant embedding */
static void circ_spectrum(const double *T, int bs, int lambda, cplx *Tfft, cplx *work) {
    double *Tc = (double *) calloc((size_t) bs, sizeof(double));
    Tc[0] = T[0];
    for (int i = 1; i < lambda; i++) {
        Tc[i] = T[i];
        Tc[bs - i] = T[i];
    }
    rfft_any(bs, Tc, Tfft, work);
    free(Tc);
}

/* toeplitz.c:651-700 (scmm_basic) + 569-616 (scmm_direct), one column, nfft=1:
 * V_rfft = V_bloc; r2c; V_fft <- C_fft * V_fft (in place, the updated real part
 * is reused for the imaginary part as at lines 587-588); c2r; TV = V_rfft/bs. */
static int g_ninv_mode = ORC_NINV_REFERENCE;

int orc_set_ninv_mode(int mode) {
    int old = g_ninv_mode;
    if (mode >= ORC_NINV_REFERENCE && mode <= ORC_NINV_DIRECT) g_ninv_mode = mode;
    return old;
}

static void scmm(const double *Vb, double *TVb, int bs, const cplx *Cf, cplx *Vf, cplx *work) {
    int sizeT = bs / 2 + 1;
    rfft_any(bs, Vb, Vf, work);
    if (g_ninv_mode == ORC_NINV_REFERENCE) {
        for (int i = 0; i < sizeT; i++) {
            Vf[i].re = Cf[i].re * Vf[i].re - Cf[i].im * Vf[i].im;
            Vf[i].im = Cf[i].re * Vf[i].im + Cf[i].im * Vf[i].re;
        }
    } else {
        /* the spectrum of the real symmetric circulant embedding is real */
        for (int i = 0; i < sizeT; i++) {
            Vf[i].re = Cf[i].re * Vf[i].re;
            Vf[i].im = Cf[i].re * Vf[i].im;
        }
    }
    irfft_any(bs, Vf, TVb, work);
    const double norm = 1.0 / ((double) bs);
    for (int i = 0; i < bs; i++) TVb[i] = TVb[i] * norm;
}

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

/* direct banded product with zero boundaries: the closed form that every tile
 * of the shift-and-overlap algorithm reproduces (cf. toeplitz_nofft.c:104-136,
 * here computed on every row). */
static void stmm_direct(double *V, int n, const double *T, int lambda) {
    double *out = (double *) malloc((size_t) n * sizeof(double));
    for (int i = 0; i < n; i++) {
        int j0 = imax(i - (lambda - 1), 0), j1 = imin(i + lambda, n);
        double acc = 0.0;
        for (int j = j0; j < j1; j++) acc += T[abs(j - i)] * V[j];
        out[i] = acc;
    }
    memcpy(V, out, (size_t) n * sizeof(double));
    free(out);
}

/* toeplitz.c:888-932 (stmm_main, shortcut m_eff==1, nfft==1) and 731-865
 * (stmm_core, flag_offset=0): sliding window of tiles of length bs, each
 * keeping bs_eff = bs - 2(lambda-1) central outputs. */
int orc_stmm_block(double *V, int n, const double *T, int lambda, int use_fft) {
    const int dcm = lambda - 1;
    if (n < dcm) return 1; /* stmm_main:912-913: block left unchanged */
    if (!use_fft || g_ninv_mode == ORC_NINV_DIRECT) {
        stmm_direct(V, n, T, lambda);
        return ORC_OK;
    }
    const int bs = orc_define_blocksize(n, lambda);
    const int bs_eff = bs - 2 * dcm;
    const int nbloc = (int) ceil((1.0 * n) / bs_eff);
    double *Vb = (double *) calloc((size_t) bs, sizeof(double));
    double *TVb = (double *) calloc((size_t) bs, sizeof(double));
    cplx *Cf = (cplx *) malloc((size_t) (bs / 2 + 1) * sizeof(cplx));
    cplx *Vf = (cplx *) malloc((size_t) (bs / 2 + 1) * sizeof(cplx));
    cplx *work = (cplx *) malloc((size_t) bs * sizeof(cplx));
    if (!Vb || !TVb || !Cf || !Vf || !work) return ORC_ENOMEM;
    circ_spectrum(T, bs, lambda, Cf, work);

    int iV = 0, iTV = 0, cs;
    cs = imin(bs - dcm, n - iV);
    for (int i = 0; i < cs; i++) Vb[dcm + i] = V[iV + i];
    scmm(Vb, TVb, bs, Cf, Vf, work);
    iV = bs_eff - dcm;
    if (nbloc > 1) {
        cs = imin(bs, n - iV);
        if (cs != bs) memset(Vb, 0, (size_t) bs * sizeof(double));
        for (int i = 0; i < cs; i++) Vb[i] = V[iV + i];
    }
    cs = imin(bs_eff, n - iTV);
    for (int i = 0; i < cs; i++) V[iTV + i] = TVb[dcm + i];
    iTV += bs_eff;
    for (int k = 1; k < nbloc; k++) {
        scmm(Vb, TVb, bs, Cf, Vf, work);
        iV += bs_eff;
        if (k != nbloc - 1) {
            cs = imin(bs, n - iV);
            if (cs != bs) memset(Vb, 0, (size_t) bs * sizeof(double));
            for (int i = 0; i < cs; i++) Vb[i] = V[iV + i];
        }
        cs = imin(bs_eff, n - iTV);
        for (int i = 0; i < cs; i++) V[iTV + i] = TVb[dcm + i];
        iTV += bs_eff;
    }
    free(Vb);
    free(TVb);
    free(Cf);
    free(Vf);
    free(work);
    return ORC_OK;
}

/* toeplitz_wizard.c:80-111 (stbmmProd): per-block offset removal 86-95, then
 * mpi_stbmm (toeplitz_block.c:115-636), which on one rank with contiguous
 * blocks processes every block [idv, idv+n) independently.  Samples past the
 * last block are left untouched. */
static int g_offset_mode = ORC_OFFSET_REFERENCE;

int orc_set_offset_mode(int mode) {
    int old = g_offset_mode;
    if (mode >= ORC_OFFSET_REFERENCE && mode <= ORC_OFFSET_NONE) g_offset_mode = mode;
    return old;
}

int orc_stbmm(int nb, const int *block_sizes, int lambda, const double *invtt, int64_t m,
              double *V, int use_fft) {
    int64_t tot = 0;
    for (int b = 0; b < nb; b++) {
        if (block_sizes[b] <= 0) return ORC_EINVAL;
        tot += block_sizes[b];
    }
    if (tot > m) return ORC_EINVAL;
    /* blocks are independent: in the all-cores timing build they run in parallel */
    int64_t *b0 = (int64_t *) malloc((size_t) (nb + 1) * sizeof(int64_t));
    if (!b0) return ORC_ENOMEM;
    b0[0] = 0;
    for (int b = 0; b < nb; b++) b0[b + 1] = b0[b] + block_sizes[b];
    int err = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic) reduction(min : err)
#endif
    for (int b = 0; b < nb; b++) {
        const int64_t t_id = b0[b];
        const int n = block_sizes[b];
        if (g_offset_mode != ORC_OFFSET_NONE) {
            double offset = 0;
            for (int j = 0; j < n; j++) offset += V[t_id + j];
            offset /= n;
            for (int j = 0; j < n; j++) V[t_id + j] -= offset;
        }
        int rc = orc_stmm_block(V + t_id, n, invtt + (int64_t) b * lambda, lambda, use_fft);
        if (rc < 0) {
            err = rc < err ? rc : err;
            continue;
        }
        if (g_offset_mode == ORC_OFFSET_SYMMETRIC) {
            /* the projection C applied again after the product: C T C */
            double mo = 0;
            for (int j = 0; j < n; j++) mo += V[t_id + j];
            mo /= n;
            for (int j = 0; j < n; j++) V[t_id + j] -= mo;
        }
    }
    free(b0);
    return err < 0 ? err : ORC_OK;
}

/* pcg_true.c:578-600 */
int orc_apply_weights(int nb, const int *block_sizes, int lambda, const double *invtt,
                      int64_t m, double *V, int use_fft) {
    if (lambda == 1) {
        int64_t t_id = 0;
        for (int b = 0; b < nb; b++) {
            for (int j = 0; j < block_sizes[b]; j++) V[t_id + j] = invtt[b] * V[t_id + j];
            t_id += block_sizes[b];
        }
        return ORC_OK;
    }
    return orc_stbmm(nb, block_sizes, lambda, invtt, m, V, use_fft);
}

/* ========================================================================
 * LAPACK restatement (reference LAPACK / reference BLAS algorithms for a
 * small dense n x n matrix; row-major storage, element (i,j) at a[i*n+j]).
 * The LAPACKE row-major wrappers called at precond.c:530-566 transpose to
 * column-major and back, so they compute on the same logical matrix.
 * ====================================================================== */

/* DLANGE('1'): max column sum of |a|, rows summed in order. */
double orc_dlange1(int n, const double *a) {
    double value = 0.0;
    for (int j = 0; j < n; j++) {
        double sum = 0.0;
        for (int i = 0; i < n; i++) sum += fabs(a[i * n + j]);
        if (value < sum || isnan(sum)) value = sum;
    }
    return value;
}

static int idamax(int n, const double *x, int incx) {
    int imx = 0;
    double dmax = fabs(x[0]);
    for (int i = 1; i < n; i++)
        if (fabs(x[i * incx]) > dmax) {
            imx = i;
            dmax = fabs(x[i * incx]);
        }
    return imx;
}

/* DGETRF2 (recursive LU with partial pivoting) on the sub-matrix of rows
 * [r0, r0+mr) and columns [c0, c0+nc); ipiv holds 0-based absolute rows. */
static int dgetrf2_rec(int n, double *a, int r0, int c0, int mr, int nc, int *ipiv) {
#define A(i, j) a[(int64_t) (r0 + (i)) * n + (c0 + (j))]
    int info = 0;
    if (mr == 1) {
        ipiv[0] = r0;
        if (A(0, 0) == 0.0) info = 1;
        return info;
    }
    if (nc == 1) {
        const double sfmin = 2.2250738585072014e-308; /* DLAMCH('S') */
        int ip = 0;
        double dmax = fabs(A(0, 0));
        for (int i = 1; i < mr; i++)
            if (fabs(A(i, 0)) > dmax) {
                ip = i;
                dmax = fabs(A(i, 0));
            }
        ipiv[0] = r0 + ip;
        if (A(ip, 0) != 0.0) {
            if (ip != 0) {
                double t = A(0, 0);
                A(0, 0) = A(ip, 0);
                A(ip, 0) = t;
            }
            if (fabs(A(0, 0)) >= sfmin) {
                double r = 1.0 / A(0, 0);
                for (int i = 1; i < mr; i++) A(i, 0) = r * A(i, 0);
            } else {
                for (int i = 1; i < mr; i++) A(i, 0) = A(i, 0) / A(0, 0);
            }
        } else {
            info = 1;
        }
        return info;
    }
    int mn = mr < nc ? mr : nc;
    int n1 = mn / 2, n2 = nc - n1;
    int iinfo = dgetrf2_rec(n, a, r0, c0, mr, n1, ipiv);
    if (info == 0 && iinfo > 0) info = iinfo;
    /* DLASWP on A12/A22 for pivots 0..n1-1 */
    for (int k = 0; k < n1; k++) {
        int p = ipiv[k] - r0;
        if (p != k)
            for (int j = n1; j < nc; j++) {
                double t = A(k, j);
                A(k, j) = A(p, j);
                A(p, j) = t;
            }
    }
    /* DTRSM L,L,N,U: A12 <- L11^-1 A12 (column by column, reference BLAS order) */
    for (int j = n1; j < nc; j++)
        for (int k = 0; k < n1; k++)
            if (A(k, j) != 0.0)
                for (int i = k + 1; i < n1; i++) A(i, j) = A(i, j) - A(k, j) * A(i, k);
    /* DGEMM: A22 <- A22 - A21 A12 (reference BLAS: temp = alpha*B(l,j); C += temp*A(i,l)) */
    for (int j = n1; j < nc; j++)
        for (int l = 0; l < n1; l++) {
            double temp = -1.0 * A(l, j);
            for (int i = n1; i < mr; i++) A(i, j) = A(i, j) + temp * A(i, l);
        }
    iinfo = dgetrf2_rec(n, a, r0 + n1, c0 + n1, mr - n1, n2, ipiv + n1);
    if (info == 0 && iinfo > 0) info = iinfo + n1;
    /* DLASWP on A21 for pivots n1..mn-1 */
    for (int k = n1; k < mn; k++) {
        int p = ipiv[k] - r0;
        if (p != k)
            for (int j = 0; j < n1; j++) {
                double t = A(k, j);
                A(k, j) = A(p, j);
                A(p, j) = t;
            }
    }
    return info;
#undef A
}

int orc_dgetrf(int n, double *a, int *ipiv) { return dgetrf2_rec(n, a, 0, 0, n, n, ipiv); }

/* DTRSV variants on the packed LU (reference BLAS loop orders). */
static void trsv_lower_unit_n(int n, const double *a, double *x) {
    for (int j = 0; j < n; j++)
        if (x[j] != 0.0) {
            double temp = x[j];
            for (int i = j + 1; i < n; i++) x[i] = x[i] - temp * a[i * n + j];
        }
}
static void trsv_upper_nonunit_n(int n, const double *a, double *x) {
    for (int j = n - 1; j >= 0; j--)
        if (x[j] != 0.0) {
            x[j] = x[j] / a[j * n + j];
            double temp = x[j];
            for (int i = j - 1; i >= 0; i--) x[i] = x[i] - temp * a[i * n + j];
        }
}
static void trsv_upper_nonunit_t(int n, const double *a, double *x) {
    for (int j = 0; j < n; j++) {
        double temp = x[j];
        for (int i = 0; i < j; i++) temp = temp - a[i * n + j] * x[i];
        temp = temp / a[j * n + j];
        x[j] = temp;
    }
}
static void trsv_lower_unit_t(int n, const double *a, double *x) {
    for (int j = n - 1; j >= 0; j--) {
        double temp = x[j];
        for (int i = n - 1; i > j; i--) temp = temp - a[i * n + j] * x[i];
        x[j] = temp;
    }
}

/* DLATRS fast-path test: returns 1 when DLATRS would call DTRSV (no scaling,
 * SCALE = 1), 0 when it would take its scaled path.  Only the growth bound is
 * restated; the scaled path itself is reached only for blocks whose rcond is
 * far below the 0.1 cut (exactly singular U or growth below 1e-292), and there
 * the result is taken as rcond = 0. */
static int latrs_fast_ok(int n, const double *a, const double *x, int upper, int unit, int trans) {
    const double smlnum = 2.2250738585072014e-308 / 1.1102230246251565e-16; /* SAFMIN/EPS (DLAMCH 'P') */
    double cnorm[16];
    for (int j = 0; j < n; j++) {
        double s = 0.0;
        if (upper)
            for (int i = 0; i < j; i++) s += fabs(a[i * n + j]);
        else
            for (int i = j + 1; i < n; i++) s += fabs(a[i * n + j]);
        cnorm[j] = s;
    }
    double xmax = fabs(x[idamax(n, x, 1)]);
    double grow;
    if (!trans) {
        int jfirst = upper ? n - 1 : 0, jlast = upper ? -1 : n, jinc = upper ? -1 : 1;
        if (!unit) {
            grow = 1.0 / fmax(xmax, smlnum);
            double xbnd = grow;
            for (int j = jfirst; j != jlast; j += jinc) {
                if (grow <= smlnum) break;
                double tjj = fabs(a[j * n + j]);
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
        int jfirst = upper ? 0 : n - 1, jlast = upper ? n : -1, jinc = upper ? 1 : -1;
        if (!unit) {
            grow = 1.0 / fmax(xmax, smlnum);
            double xbnd = grow;
            for (int j = jfirst; j != jlast; j += jinc) {
                if (grow <= smlnum) break;
                double xj = 1.0 + cnorm[j];
                grow = fmin(grow, xbnd / xj);
                double tjj = fabs(a[j * n + j]);
                if (xj > tjj) xbnd = xbnd * (tjj / xj);
            }
            grow = fmin(grow, xbnd);
        } else {
            grow = fmin(1.0, 1.0 / fmax(xmax, smlnum));
            for (int j = jfirst; j != jlast; j += jinc) {
                if (grow <= smlnum) break;
                double xj = 1.0 + cnorm[j];
                grow = grow / xj;
            }
        }
    }
    return grow > smlnum; /* TSCAL = 1 for any finite block here */
}

/* DGECON('1') with DLACN2 (Higham's modification of Hager's estimator). */
double orc_dgecon1(int n, const double *lu, double anorm) {
    if (n == 0) return 1.0;
    if (anorm == 0.0) return 0.0;
    if (isnan(anorm)) return anorm;
    if (isinf(anorm)) return 0.0;
    double x[16], v[16];
    int isgn[16];
    double est = 0.0, estold;
    int kase, jump, j, jlast, iter;

    /* the estimator's reverse-communication loop, unrolled as a state machine */
    for (int i = 0; i < n; i++) x[i] = 1.0 / (double) n;
    kase = 1;
    jump = 1;
    iter = 0;
    j = 0;
    for (;;) {
        /* apply the operator requested by kase: 1 -> inv(U) inv(L), 2 -> inv(L^T) inv(U^T) */
        if (kase == 1) {
            if (!latrs_fast_ok(n, lu, x, 0, 1, 0)) return 0.0;
            trsv_lower_unit_n(n, lu, x);
            if (!latrs_fast_ok(n, lu, x, 1, 0, 0)) return 0.0;
            trsv_upper_nonunit_n(n, lu, x);
        } else {
            if (!latrs_fast_ok(n, lu, x, 1, 0, 1)) return 0.0;
            trsv_upper_nonunit_t(n, lu, x);
            if (!latrs_fast_ok(n, lu, x, 0, 1, 1)) return 0.0;
            trsv_lower_unit_t(n, lu, x);
        }
        /* DLACN2 entry points */
        if (jump == 1) {
            if (n == 1) {
                v[0] = x[0];
                est = fabs(v[0]);
                break;
            }
            est = 0.0;
            for (int i = 0; i < n; i++) est += fabs(x[i]);
            for (int i = 0; i < n; i++) {
                x[i] = (x[i] >= 0.0) ? 1.0 : -1.0;
                isgn[i] = (int) x[i];
            }
            kase = 2;
            jump = 2;
            continue;
        }
        if (jump == 2) {
            j = idamax(n, x, 1);
            iter = 2;
        L50:
            for (int i = 0; i < n; i++) x[i] = 0.0;
            x[j] = 1.0;
            kase = 1;
            jump = 3;
            continue;
        }
        if (jump == 3) {
            for (int i = 0; i < n; i++) v[i] = x[i];
            estold = est;
            est = 0.0;
            for (int i = 0; i < n; i++) est += fabs(v[i]);
            int same = 1;
            for (int i = 0; i < n; i++) {
                double xs = (x[i] >= 0.0) ? 1.0 : -1.0;
                if ((int) xs != isgn[i]) {
                    same = 0;
                    break;
                }
            }
            if (same || est <= estold) goto L120;
            for (int i = 0; i < n; i++) {
                x[i] = (x[i] >= 0.0) ? 1.0 : -1.0;
                isgn[i] = (int) x[i];
            }
            kase = 2;
            jump = 4;
            continue;
        }
        if (jump == 4) {
            jlast = j;
            j = idamax(n, x, 1);
            if (x[jlast] != fabs(x[j]) && iter < 5) {
                iter++;
                goto L50;
            }
        L120 : {
            double altsgn = 1.0;
            for (int i = 0; i < n; i++) {
                x[i] = altsgn * (1.0 + (double) i / (double) (n - 1));
                altsgn = -altsgn;
            }
        }
            kase = 1;
            jump = 5;
            continue;
        }
        if (jump == 5) {
            double s = 0.0;
            for (int i = 0; i < n; i++) s += fabs(x[i]);
            double temp = 2.0 * (s / (double) (3 * n));
            if (temp > est) est = temp;
            break;
        }
    }
    if (est == 0.0) return 0.0;
    return (1.0 / est) / anorm;
}

/* DGETRI with DTRTRI/DTRTI2 (upper, non-unit) and the unblocked column loop. */
int orc_dgetri(int n, double *a, const int *ipiv) {
#define A(i, j) a[(i) * n + (j)]
    for (int i = 0; i < n; i++)
        if (A(i, i) == 0.0) return i + 1;
    /* DTRTI2 upper non-unit */
    for (int j = 0; j < n; j++) {
        A(j, j) = 1.0 / A(j, j);
        double ajj = -A(j, j);
        /* DTRMV upper, no-trans, non-unit on column j, order j-1 */
        for (int jj = 0; jj < j; jj++)
            if (A(jj, j) != 0.0) {
                double temp = A(jj, j);
                for (int i = 0; i < jj; i++) A(i, j) = A(i, j) + temp * A(i, jj);
                A(jj, j) = A(jj, j) * A(jj, jj);
            }
        for (int i = 0; i < j; i++) A(i, j) = ajj * A(i, j);
    }
    double work[16];
    for (int j = n - 1; j >= 0; j--) {
        for (int i = j + 1; i < n; i++) {
            work[i] = A(i, j);
            A(i, j) = 0.0;
        }
        if (j < n - 1) {
            /* DGEMV N: A(:,j) <- A(:,j) - A(:,j+1:n) work(j+1:n) */
            for (int jj = j + 1; jj < n; jj++) {
                double temp = -1.0 * work[jj];
                for (int i = 0; i < n; i++) A(i, j) = A(i, j) + temp * A(i, jj);
            }
        }
    }
    for (int j = n - 2; j >= 0; j--) {
        int jp = ipiv[j];
        if (jp != j)
            for (int i = 0; i < n; i++) {
                double t = A(i, j);
                A(i, j) = A(i, jp);
                A(i, jp) = t;
            }
    }
    return 0;
#undef A
}

/* ========================================================================
 * Preconditioner
 * ====================================================================== */

/* precond.c:42-161 on one rank (idp = 0, vShft = 0, idv0 = 0): accumulate the
 * nnz x nnz blocks sum_t w_j w_l T_b(t)[0] and the hit counts.  Samples after
 * the last block get w_j w_l without T0 (lines 144-156). */
void orc_getlocalW(int m, int nnz, const int *indices, const double *values, int trash_pix,
                   int lcount, int nb, const int *block_sizes, int lambda, const double *invtt,
                   double *vpixBlock, int *lhits) {
    const int n_valid = lcount - trash_pix * nnz;
    for (int i = 0; i < nnz * n_valid; i++) vpixBlock[i] = 0.0;
    int64_t idv = 0;
    for (int k = 0; k < nb; k++) {
        int64_t istartn = (k + 1 < nb) ? idv + block_sizes[k] : m;
        int64_t istart = idv;
        int64_t il = block_sizes[k];
        if (k == nb - 1 && il > m - idv) il = m - idv;
        if (nb == 1 && il > m) il = m;
        const double diag = invtt[(int64_t) k * lambda];
        for (int64_t i = istart; i < istart + il; i++) {
            if (!trash_pix || indices[i * nnz] != 0) {
                int vpix_i = indices[i * nnz] / nnz - trash_pix;
                lhits[vpix_i] += 1;
                for (int j = 0; j < nnz; j++) {
                    int vpix_j = indices[i * nnz + j] - nnz * trash_pix;
                    for (int l = 0; l < nnz; l++)
                        vpixBlock[(int64_t) nnz * vpix_j + l] += values[i * nnz + j] * values[i * nnz + l] * diag;
                }
            }
        }
        for (int64_t i = istart + il; i < istartn; i++) {
            if (!trash_pix || indices[i * nnz] != 0) {
                int vpix_i = indices[i * nnz] / nnz - trash_pix;
                lhits[vpix_i] += 1;
                for (int j = 0; j < nnz; j++) {
                    int vpix_j = indices[i * nnz + j] - nnz * trash_pix;
                    for (int l = 0; l < nnz; l++)
                        vpixBlock[(int64_t) nnz * vpix_j + l] += values[i * nnz + j] * values[i * nnz + l];
                }
            }
        }
        idv += block_sizes[k];
    }
}

/* ========================================================================
 * Small dense kernels of the two-level preconditioners
 * ====================================================================== */

/* Pseudo-inverse of the row-major n x n matrix a: the minimum-norm solution X
 * of min ||a X - I|| with singular values s_j <= rcond * s_max treated as zero,
 * which is what inverse_svd_bis obtains from LAPACKE_dgelss against the
 * identity (precond.c:777-796, rcond = eps = 1e-15, precond.c:25).  The SVD is
 * the one-sided Jacobi method (Hestenes 1958; Demmel & Veselic, SIAM J. Matrix
 * Anal. Appl. 13 (1992)): rotate column pairs of W = a until they are
 * orthogonal, so that W = U S and a = U S V^T with V the accumulated rotations;
 * then pinv(a) = sum_j v_j w_j^T / s_j^2 over the kept j. */
int orc_pinv(int n, const double *a, double rcond, double *out) { return orc_pinv_rank(n, a, rcond, out, NULL); }

int orc_pinv_rank(int n, const double *a, double rcond, double *out, int *rank) {
    if (n <= 0) return ORC_EINVAL;
    double *W = (double *) malloc((size_t) n * n * sizeof(double));
    double *V = (double *) calloc((size_t) n * n, sizeof(double));
    if (!W || !V) {
        free(W);
        free(V);
        return ORC_ENOMEM;
    }
    memcpy(W, a, (size_t) n * n * sizeof(double));
    for (int i = 0; i < n; i++) V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 80; sweep++) {
        int rotated = 0;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double al = 0, be = 0, ga = 0;
                for (int i = 0; i < n; i++) {
                    al += W[i * n + p] * W[i * n + p];
                    be += W[i * n + q] * W[i * n + q];
                    ga += W[i * n + p] * W[i * n + q];
                }
                if (ga == 0.0 || fabs(ga) <= 1e-15 * sqrt(al * be)) continue;
                rotated = 1;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                for (int i = 0; i < n; i++) {
                    const double wp = W[i * n + p], wq = W[i * n + q];
                    W[i * n + p] = c * wp - sn * wq;
                    W[i * n + q] = sn * wp + c * wq;
                    const double vp = V[i * n + p], vq = V[i * n + q];
                    V[i * n + p] = c * vp - sn * vq;
                    V[i * n + q] = sn * vp + c * vq;
                }
            }
        if (!rotated) break;
    }
    double smax = 0;
    double *s2 = (double *) malloc((size_t) n * sizeof(double));
    for (int j = 0; j < n; j++) {
        double t = 0;
        for (int i = 0; i < n; i++) t += W[i * n + j] * W[i * n + j];
        s2[j] = t;
        if (sqrt(t) > smax) smax = sqrt(t);
    }
    for (int i = 0; i < n * n; i++) out[i] = 0.0;
    int kept = 0;
    for (int j = 0; j < n; j++) {
        if (!(sqrt(s2[j]) > rcond * smax)) continue;
        kept++;
        for (int i = 0; i < n; i++)
            for (int k = 0; k < n; k++) out[i * n + k] += V[i * n + j] * W[k * n + j] / s2[j];
    }
    free(s2);
    free(W);
    free(V);
    if (rank) *rank = kept;
    return ORC_OK;
}

/* Eigen-decomposition of the symmetric row-major n x n matrix a by cyclic
 * Jacobi rotations (Golub & Van Loan, Matrix Computations, 4th ed., 8.5):
 * eigenvalues ascending in w, eigenvector j in column j of the row-major v --
 * the layout LAPACKE_dsyev(LAPACK_ROW_MAJOR, 'V', ...) leaves in its matrix
 * argument (precond.c:1240). */
int orc_symeig(int n, const double *a, double *w, double *v) {
    if (n <= 0) return ORC_EINVAL;
    double *A = (double *) malloc((size_t) n * n * sizeof(double));
    double *Vt = (double *) calloc((size_t) n * n, sizeof(double));
    int *ord = (int *) malloc((size_t) n * sizeof(int));
    if (!A || !Vt || !ord) {
        free(A);
        free(Vt);
        free(ord);
        return ORC_ENOMEM;
    }
    memcpy(A, a, (size_t) n * n * sizeof(double));
    for (int i = 0; i < n; i++) Vt[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                tot += A[i * n + j] * A[i * n + j];
                if (i != j) off += A[i * n + j] * A[i * n + j];
            }
        if (off == 0.0 || off <= 1e-32 * tot) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A[p * n + q];
                if (apq == 0.0) continue;
                const double theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) { /* A <- A J (columns p, q) */
                    const double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) { /* A <- J^T A (rows p, q) */
                    const double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) { /* V <- V J */
                    const double vkp = Vt[k * n + p], vkq = Vt[k * n + q];
                    Vt[k * n + p] = c * vkp - s * vkq;
                    Vt[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++) ord[i] = i;
    for (int i = 1; i < n; i++) { /* insertion sort by eigenvalue */
        const int o = ord[i];
        int j = i - 1;
        while (j >= 0 && A[ord[j] * n + ord[j]] > A[o * n + o]) {
            ord[j + 1] = ord[j];
            j--;
        }
        ord[j + 1] = o;
    }
    for (int j = 0; j < n; j++) {
        w[j] = A[ord[j] * n + ord[j]];
        for (int i = 0; i < n; i++) v[i * n + j] = Vt[i * n + ord[j]];
    }
    free(A);
    free(Vt);
    free(ord);
    return ORC_OK;
}

/* ========================================================================
 * Two-level preconditioners (build_precond precond = 1, 2, precond.c:1342-1398)
 * ====================================================================== */

typedef struct {
    int Zn, n;
    double *Z, *AZ; /* column j at [j * n] */
    double *Em1;    /* Zn x Zn row-major */
    double *w, *Qtx, *Qg, *AQg;
} orc_twolvl;

typedef struct {
    int m, nnz, trash, lcount, nb, lambda;
    const int *pix, *block_sizes;
    const double *w, *invtt;
} orc_op;

/* A v = P^T stbmmProd(P v), the product of build_AZ (precond.c:1017-1022) and
 * Lanczos_eig (precond.c:1151-1153, 1205-1207): stbmmProd, not apply_weights,
 * so lambda == 1 also removes the offsets there. */
static int op_apply(const orc_op *o, const double *v, double *tod, double *out) {
    orc_matvec(o->m, o->nnz, o->pix, o->w, o->trash, v, tod);
    int rc = orc_stbmm(o->nb, o->block_sizes, o->lambda, o->invtt, o->m, tod, 1);
    if (rc < 0) return rc;
    orc_trmatvec(o->m, o->nnz, o->pix, o->w, o->trash, o->lcount, tod, out);
    return ORC_OK;
}

/* build_Z on one rank (precond.c:798-953): group = Zn (Zn > 1) or 1, chunk g =
 * samples [g*(m/group), (g+1)*(m/group)); count[g*lcount + i] = occurrences of
 * local index i in chunk g; column g at local index i (step nnz) =
 * count[g*group + i] / tcount[i] -- the reference's own index (precond.c:874),
 * kept as written -- then the trash entries are dropped (precond.c:943-945). */
static int build_Z(const orc_op *o, int Zn, double *Z, int n) {
    const int lcount = o->lcount, nnz = o->nnz;
    const int group = Zn > 1 ? Zn : 1;
    int *count = (int *) calloc((size_t) group * lcount, sizeof(int));
    int *tcount = (int *) calloc((size_t) (lcount > 0 ? lcount : 1), sizeof(int));
    double *col = (double *) malloc((size_t) (lcount > 0 ? lcount : 1) * sizeof(double));
    if (!count || !tcount || !col) {
        free(count);
        free(tcount);
        free(col);
        return ORC_ENOMEM;
    }
    const int64_t per = (int64_t) (o->m / group) * nnz;
    for (int g = 0; g < group; g++)
        for (int64_t i = 0; i < per; i++) count[(int64_t) g * lcount + o->pix[g * per + i]]++;
    for (int g = 0; g < group; g++)
        for (int i = 0; i < lcount; i++) tcount[i] += count[(int64_t) g * lcount + i];
    for (int g = 0; g < Zn; g++) {
        for (int i = 0; i < lcount; i++) col[i] = 0.0;
        for (int i = 0; i < lcount; i += nnz)
            col[i] = (double) ((double) count[(int64_t) g * group + i] / (double) tcount[i]);
        for (int i = 0; i < n; i++) Z[(int64_t) g * n + i] = col[i + nnz * o->trash];
    }
    free(count);
    free(tcount);
    free(col);
    return ORC_OK;
}

static double dot_seq(int n, const double *a, const double *b);

/* Lanczos_eig (precond.c:1073-1306) from x = 0: K steps on A from
 * w = P^T N^-1 (b + noise), the Ritz vectors of the K x K tridiagonal T and
 * their images under A.  Kept as written: the first step subtracts no
 * beta * v_prev term (vold starts at 0, precond.c:1214-1218). */
static int lanczos(const orc_op *o, const double *b, const double *noise, int K, int n, double *Z, double *AZ) {
    const double eps = 1.0e-15;
    const int m = o->m;
    const int K1 = K + 1;
    double *tod = (double *) malloc((size_t) (m > 0 ? m : 1) * sizeof(double));
    double *w = (double *) calloc((size_t) n, sizeof(double));
    double *v = (double *) calloc((size_t) n, sizeof(double));
    double *vold = (double *) calloc((size_t) n, sizeof(double));
    double *Av = (double *) calloc((size_t) n, sizeof(double));
    double *V = (double *) calloc((size_t) n * K1, sizeof(double));     /* column k at [k*n] */
    double *AmulV = (double *) calloc((size_t) n * K1, sizeof(double)); /* column k at [k*n] */
    double *T = (double *) calloc((size_t) K1 * K1, sizeof(double));
    double *Tk = (double *) calloc((size_t) K * K, sizeof(double));
    double *ev = (double *) calloc((size_t) K * K, sizeof(double));
    double *lam = (double *) calloc((size_t) K, sizeof(double));
    int rc = ORC_ENOMEM;
    if (!tod || !w || !v || !vold || !Av || !V || !AmulV || !T || !Tk || !ev || !lam) goto out;
    for (int i = 0; i < m; i++) tod[i] = b[i] + noise[i] - 0.0; /* P x with x = 0 */
    rc = orc_stbmm(o->nb, o->block_sizes, o->lambda, o->invtt, m, tod, 1);
    if (rc < 0) goto out;
    orc_trmatvec(m, o->nnz, o->pix, o->w, o->trash, o->lcount, tod, w);
    double beta = sqrt(dot_seq(n, w, w));
    if (beta > eps)
        for (int i = 0; i < n; i++) {
            v[i] = w[i] / beta;
            V[i] = v[i];
        }
    if ((rc = op_apply(o, v, tod, Av)) < 0) goto out;
    memcpy(w, Av, (size_t) n * sizeof(double));
    double alpha = dot_seq(n, v, Av);
    for (int i = 0; i < n; i++) {
        AmulV[i] = w[i];
        w[i] = w[i] - (alpha * v[i]);
    }
    for (int it = 0; it < K; it++) {
        beta = sqrt(dot_seq(n, w, w));
        if (beta > eps)
            for (int j = 0; j < n; j++) {
                v[j] = w[j] / beta;
                V[(int64_t) (it + 1) * n + j] = v[j];
            }
        T[it * K1 + it] = alpha;
        T[it * K1 + it + 1] = beta;
        T[(it + 1) * K1 + it] = beta;
        if ((rc = op_apply(o, v, tod, Av)) < 0) goto out;
        memcpy(w, Av, (size_t) n * sizeof(double));
        alpha = dot_seq(n, v, Av);
        for (int j = 0; j < n; j++) {
            AmulV[(int64_t) (it + 1) * n + j] = w[j];
            w[j] = w[j] - (alpha * v[j]) - (beta * vold[j]);
            vold[j] = v[j];
        }
    }
    for (int i = 0; i < K; i++)
        for (int j = 0; j < K; j++) Tk[i * K + j] = T[i * K1 + j];
    if ((rc = orc_symeig(K, Tk, lam, ev)) < 0) goto out;
    /* Ritz vectors: Z = V ev, AZ = AmulV ev over the first K columns (precond.c:1254-1282) */
    for (int j = 0; j < K; j++)
        for (int i = 0; i < n; i++) {
            double z = 0.0, az = 0.0;
            for (int k = 0; k < K; k++) {
                z += V[(int64_t) k * n + i] * ev[k * K + j];
                az += AmulV[(int64_t) k * n + i] * ev[k * K + j];
            }
            Z[(int64_t) j * n + i] = z;
            AZ[(int64_t) j * n + i] = az;
        }
    rc = ORC_OK;
out:
    free(tod);
    free(w);
    free(v);
    free(vold);
    free(Av);
    free(V);
    free(AmulV);
    free(T);
    free(Tk);
    free(ev);
    free(lam);
    return rc;
}

static void twolvl_free(orc_twolvl *t) {
    free(t->Z);
    free(t->AZ);
    free(t->Em1);
    free(t->w);
    free(t->Qtx);
    free(t->Qg);
    free(t->AQg);
    memset(t, 0, sizeof(*t));
}

/* build_precond for precond = 1 (build_Z + build_AZ) or 2 (Lanczos_eig), then
 * build_Em1 (precond.c:956-1003): E[i*Zn+j] = <Z_j, AZ_i>, Em1 = pinv(E). */
static int twolvl_build(orc_twolvl *t, const orc_op *o, int precond, int Zn, int n, const double *b,
                        const double *noise) {
    memset(t, 0, sizeof(*t));
    t->Zn = Zn;
    t->n = n;
    const size_t nz = (size_t) Zn * (n > 0 ? n : 1);
    t->Z = (double *) calloc(nz, sizeof(double));
    t->AZ = (double *) calloc(nz, sizeof(double));
    t->Em1 = (double *) calloc((size_t) Zn * Zn, sizeof(double));
    t->w = (double *) calloc((size_t) Zn, sizeof(double));
    t->Qtx = (double *) calloc((size_t) Zn, sizeof(double));
    t->Qg = (double *) calloc((size_t) (n > 0 ? n : 1), sizeof(double));
    t->AQg = (double *) calloc((size_t) (n > 0 ? n : 1), sizeof(double));
    double *tod = (double *) malloc((size_t) (o->m > 0 ? o->m : 1) * sizeof(double));
    double *E = (double *) calloc((size_t) Zn * Zn, sizeof(double));
    int rc = ORC_ENOMEM;
    if (!t->Z || !t->AZ || !t->Em1 || !t->w || !t->Qtx || !t->Qg || !t->AQg || !tod || !E) goto out;
    if (precond == 1) {
        if ((rc = build_Z(o, Zn, t->Z, n)) < 0) goto out;
        for (int i = 0; i < Zn; i++)
            if ((rc = op_apply(o, t->Z + (int64_t) i * n, tod, t->AZ + (int64_t) i * n)) < 0) goto out;
    } else {
        if ((rc = lanczos(o, b, noise, Zn, n, t->Z, t->AZ)) < 0) goto out;
    }
    for (int i = 0; i < Zn; i++)
        for (int j = 0; j < Zn; j++)
            for (int k = 0; k < n; k++)
                E[i * Zn + j] += (double) (t->Z[(int64_t) j * n + k] * t->AZ[(int64_t) i * n + k] * 1.0);
    rc = orc_pinv(Zn, E, 1.0e-15, t->Em1);
out:
    free(tod);
    free(E);
    if (rc < 0) twolvl_free(t);
    return rc;
}

/* apply_precond, two-level branch (precond.c:1412-1422 with build_Qtx 1032-1055
 * and mul_ZQtx 1057-1071): Cg = BJ (g - AZ Qtx) + Z Qtx, Qtx = Em1 Z^T g. */
static void twolvl_apply(orc_twolvl *t, int nnz, const double *inv, const double *g, double *Cg) {
    const int n = t->n, Zn = t->Zn;
    for (int j = 0; j < Zn; j++) {
        t->w[j] = 0.0;
        for (int k = 0; k < n; k++) t->w[j] += t->Z[(int64_t) j * n + k] * g[k] * 1.0;
    }
    for (int i = 0; i < Zn; i++) {
        t->Qtx[i] = 0.0;
        for (int j = 0; j < Zn; j++) t->Qtx[i] += t->Em1[i * Zn + j] * t->w[j];
    }
    for (int k = 0; k < n; k++) t->Qg[k] = 0.0;
    for (int j = 0; j < Zn; j++)
        for (int k = 0; k < n; k++) t->Qg[k] += t->Z[(int64_t) j * n + k] * t->Qtx[j];
    for (int k = 0; k < n; k++) t->AQg[k] = 0.0;
    for (int j = 0; j < Zn; j++)
        for (int k = 0; k < n; k++) t->AQg[k] += t->AZ[(int64_t) j * n + k] * t->Qtx[j];
    for (int i = 0; i < n; i++) t->AQg[i] = g[i] - t->AQg[i];
    for (int i = 0; i < n; i++) {
        const int q = i / nnz;
        double y = 0.0;
        for (int j = 0; j < nnz; j++) y += inv[(int64_t) i * nnz + j] * t->AQg[q * nnz + j];
        Cg[i] = y;
    }
    for (int i = 0; i < n; i++) Cg[i] += t->Qg[i];
}

/* ========================================================================
 * Whole solve: MLmap setup + build_precond + PCG_GLS_true
 * ====================================================================== */

static double dot_seq(int n, const double *a, const double *b) {
    double s = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : s) schedule(static)
#endif
    for (int i = 0; i < n; i++) s += a[i] * b[i] * 1.0; /* pixpond = 1 on one rank */
    return s;
}

/* BJ apply (precond.c:1403-1409 -> MatVecProd on BJ_inv with nnz entries per
 * row: row i of pixel q uses g[nnz q + j]). */
static void bj_apply(int n, int nnz, const double *inv, const double *g, double *Cg) {
    ORC_PFOR
    for (int i = 0; i < n; i++) {
        const int q = i / nnz;
        double y = 0.0;
        for (int j = 0; j < nnz; j++) y += inv[(int64_t) i * nnz + j] * g[q * nnz + j];
        Cg[i] = y;
    }
}

int orc_gls_solve(int m, int nnz, int *pix, double *pixweights, double *signal,
                  const double *noise, int nb, const int *block_sizes, int lambda,
                  const double *invtt, double tol, int maxiter, int stop_after,
                  double *x_out, int *lindices_out, double *cond_out, int *lhits_out,
                  double *res_hist, orc_report *rep) {
    return orc_gls_solve2(m, nnz, pix, pixweights, signal, noise, nb, block_sizes, lambda, invtt, tol, maxiter,
                          stop_after, 0, 0, x_out, lindices_out, cond_out, lhits_out, res_hist, NULL, NULL, NULL,
                          rep);
}

static int gls_core(int m, int nnz, int *pix, double *pixweights, double *signal, const double *noise, int nb,
                    const int *block_sizes, int lambda, const double *invtt, double tol, int maxiter,
                    int stop_after, int precond, int Z_2lvl, double *x_out, int *lindices_out, double *cond_out,
                    int *lhits_out, double *res_hist, double *z_out, double *az_out, double *em1_out,
                    orc_report *rep, const orc_tmpl *tm, const double *tm_binv);

int orc_gls_solve2(int m, int nnz, int *pix, double *pixweights, double *signal,
                   const double *noise, int nb, const int *block_sizes, int lambda,
                   const double *invtt, double tol, int maxiter, int stop_after, int precond, int Z_2lvl,
                   double *x_out, int *lindices_out, double *cond_out, int *lhits_out,
                   double *res_hist, double *z_out, double *az_out, double *em1_out, orc_report *rep) {
    return gls_core(m, nnz, pix, pixweights, signal, noise, nb, block_sizes, lambda, invtt, tol, maxiter, stop_after,
                    precond, Z_2lvl, x_out, lindices_out, cond_out, lhits_out, res_hist, z_out, az_out, em1_out, rep,
                    NULL, NULL);
}

/* MTmap on one rank (mappraiser.c:387-700): the white-noise model (lambda 1,
 * weight invtt[b] per block, create_toeplitz.c:27-47 with lambda_block_avg 1),
 * the template kernels B_b = w_b T_b^T T_b and their pseudo-inverses, then
 * PCG_GLS_templates (pcg_true.c:243-576): PCG_GLS_true with the template part
 * of the noise weighting subtracted after every apply_weights. */
int orc_gls_solve_tmpl(int m, int nnz, int *pix, double *pixweights, double *signal, const double *noise,
                       const int *block_sizes, const double *invtt, double tol, int maxiter, int stop_after,
                       int precond, int Z_2lvl, const orc_tmpl *tm, double *x_out, int *lindices_out,
                       double *cond_out, int *lhits_out, double *res_hist, int *ranks_out, orc_report *rep) {
    const int nb = tm->nces * tm->ndet;
    int nt_tot = 0;
    int64_t bsz = 0;
    int *nt = (int *) malloc((size_t) nb * sizeof(int));
    if (!nt) return ORC_ENOMEM;
    const int rs = orc_tmpl_sizes(tm, nt);
    if (rs < 0) {
        free(nt);
        return rs;
    }
    nt_tot = rs;
    for (int b = 0; b < nb; b++) bsz += (int64_t) nt[b] * nt[b];
    free(nt);
    (void) nt_tot;
    double *B = (double *) malloc((size_t) (bsz > 0 ? bsz : 1) * sizeof(double));
    double *Bi = (double *) malloc((size_t) (bsz > 0 ? bsz : 1) * sizeof(double));
    int rc = (!B || !Bi) ? ORC_ENOMEM : ORC_OK;
    if (!rc) rc = orc_tmpl_kernel(tm, B);
    if (!rc) rc = orc_tmpl_invkernel(tm, B, Bi, ranks_out);
    if (!rc)
        rc = gls_core(m, nnz, pix, pixweights, signal, noise, nb, block_sizes, 1, invtt, tol, maxiter, stop_after,
                      precond, Z_2lvl, x_out, lindices_out, cond_out, lhits_out, res_hist, NULL, NULL, NULL, rep, tm,
                      Bi);
    free(B);
    free(Bi);
    return rc;
}

static int gls_core(int m, int nnz, int *pix, double *pixweights, double *signal, const double *noise, int nb,
                    const int *block_sizes, int lambda, const double *invtt, double tol, int maxiter,
                    int stop_after, int precond, int Z_2lvl, double *x_out, int *lindices_out, double *cond_out,
                    int *lhits_out, double *res_hist, double *z_out, double *az_out, double *em1_out,
                    orc_report *rep, const orc_tmpl *tm, const double *tm_binv) {
    if (precond < 0 || precond > 2 || Z_2lvl < 0) return ORC_EINVAL;
    /* PCG_GLS_true: Z_2lvl == 0 means the communicator size (pcg_true.c:50) */
    const int Zn = Z_2lvl == 0 ? 1 : Z_2lvl;
    const int64_t mn = (int64_t) m * nnz;
    int lcount = 0, trash = 0, rc;
    int *lind = (int *) malloc((size_t) (mn > 0 ? mn : 1) * sizeof(int));
    if (!lind) return ORC_ENOMEM;
    memset(rep, 0, sizeof(*rep));

    /* MatInit (mappraiser.c:93-94) */
    rc = orc_local_shape(mn, pix, lind, &lcount, &trash);
    if (rc) return rc;
    int n_valid = lcount - (trash ? nnz : 0);
    int npix = n_valid / nnz;

    /* pixel -> time linked list (mappraiser.c:110-145) */
    int *id_last = (int *) malloc((size_t) (npix > 0 ? npix : 1) * sizeof(int));
    int *ll = (int *) malloc((size_t) (m > 0 ? m : 1) * sizeof(int));
    orc_pixel_lists(m, nnz, pix, trash, npix, id_last, ll);

    /* precondblockjacobilike (precond.c:456-702) */
    int *lhits = (int *) calloc((size_t) (npix > 0 ? npix : 1), sizeof(int));
    double *cond = (double *) calloc((size_t) (npix > 0 ? npix : 1), sizeof(double));
    double *W = (double *) malloc((size_t) (n_valid > 0 ? n_valid : 1) * nnz * sizeof(double));
    double *Winv = (double *) malloc((size_t) (n_valid > 0 ? n_valid : 1) * nnz * sizeof(double));
    orc_getlocalW(m, nnz, pix, pixweights, trash, lcount, nb, block_sizes, lambda, invtt, W, lhits);

    const int nnz2 = nnz * nnz;
    double xb[64];
    int ipiv[8];
    int n = n_valid;      /* shrinking effective size */
    int kept = 0, ncut = 0;
    for (int q = 0; q < npix; q++) {
        const double *blk = W + (int64_t) q * nnz2;
        memcpy(xb, blk, (size_t) nnz2 * sizeof(double));
        double anorm = orc_dlange1(nnz, xb);
        orc_dgetrf(nnz, xb, ipiv);
        double rcond = orc_dgecon1(nnz, xb, anorm);
        if (rcond > 1e-1) {
            orc_dgetri(nnz, xb, ipiv);
            memcpy(Winv + (int64_t) kept * nnz2, xb, (size_t) nnz2 * sizeof(double));
            memmove(W + (int64_t) kept * nnz2, blk, (size_t) nnz2 * sizeof(double));
            lhits[kept] = lhits[q];
            cond[kept] = rcond;
            kept++;
        } else {
            /* cut (precond.c:585-628); the pixel's own chain is followed
             * (q indexes id_last_pix directly, see Appendix quirk 5 of the
             * reference notes: the reference offsets it by the trash flag) */
            int j = id_last[q];
            for (int k = 0; k < nnz; k++) {
                pix[(int64_t) j * nnz + k] = k - nnz;
                pixweights[(int64_t) j * nnz + k] = 0;
            }
            signal[j] = 0;
            while (ll[j] != -1) {
                signal[ll[j]] = 0;
                for (int k = 0; k < nnz; k++) {
                    pix[(int64_t) ll[j] * nnz + k] = k - nnz;
                    pixweights[(int64_t) ll[j] * nnz + k] = 0;
                }
                j = ll[j];
            }
            n -= nnz;
            ncut++;
        }
    }
    if (ncut) trash = 1;
    /* back to global indices (precond.c:659-663), then MatInit again (670) */
    for (int64_t i = 0; i < mn; i++)
        if (pix[i] >= 0) pix[i] = lind[pix[i]];
    lcount = 0;
    rc = orc_local_shape(mn, pix, lind, &lcount, &trash);
    if (rc) return rc;
    n = lcount - nnz * trash;

    /* two-level part of build_precond (precond.c:1342-1398) */
    orc_twolvl tl;
    memset(&tl, 0, sizeof(tl));
    if (precond != 0) {
        if (Zn > n) return ORC_EINVAL;
        orc_op op = {m, nnz, trash, lcount, nb, lambda, pix, block_sizes, pixweights, invtt};
        rc = twolvl_build(&tl, &op, precond, Zn, n, signal, noise);
        if (rc < 0) return rc;
        if (z_out) memcpy(z_out, tl.Z, (size_t) Zn * n * sizeof(double));
        if (az_out) memcpy(az_out, tl.AZ, (size_t) Zn * n * sizeof(double));
        if (em1_out) memcpy(em1_out, tl.Em1, (size_t) Zn * Zn * sizeof(double));
    }
#define PRECOND_APPLY(G, CG) \
    do { \
        if (precond) \
            twolvl_apply(&tl, nnz, Winv, (G), (CG)); \
        else \
            bj_apply(n, nnz, Winv, (G), (CG)); \
    } while (0)

    /* PCG_GLS_true (pcg_true.c:20-241) */
    double *x = x_out;
    for (int i = 0; i < n; i++) x[i] = 0.0;
    double *h = (double *) malloc((size_t) (n > 0 ? n : 1) * sizeof(double));
    double *g = (double *) malloc((size_t) (n > 0 ? n : 1) * sizeof(double));
    double *gp = (double *) malloc((size_t) (n > 0 ? n : 1) * sizeof(double));
    double *AtNm1Ah = (double *) malloc((size_t) (n > 0 ? n : 1) * sizeof(double));
    double *Ah = (double *) malloc((size_t) (m > 0 ? m : 1) * sizeof(double));
    double *Cg = AtNm1Ah; /* same buffer, pcg_true.c:72 */

    orc_matvec(m, nnz, pix, pixweights, trash, x, Ah);
    ORC_PFOR
    for (int i = 0; i < m; i++) Ah[i] = signal[i] + noise[i] - Ah[i];
    rc = orc_apply_weights(nb, block_sizes, lambda, invtt, m, Ah, 1);
    if (rc < 0) return rc;
    if (tm && (rc = orc_tmpl_filter(tm, tm_binv, Ah)) < 0) return rc;
    orc_trmatvec(m, nnz, pix, pixweights, trash, lcount, Ah, g);
    PRECOND_APPLY(g, Cg);
    for (int j = 0; j < n; j++) h[j] = Cg[j];
    double g2pix = dot_seq(n, Cg, g);
    double res = dot_seq(n, g, g);
    double tol2rel = tol * tol * res;
    double res0 = res;
    int nh = 0;
    if (res_hist && maxiter > 0) res_hist[nh++] = sqrt(res) / sqrt(res0);
    /* pcg_true.c:127-136: the absolute test at k=0 sets k=K, but the loop
     * header re-initialises k=1, so the loop always runs. */
    int k, converged = 0;
    for (k = 1; k < maxiter; k++) {
        double *gt = gp;
        gp = g;
        g = gt;
        orc_matvec(m, nnz, pix, pixweights, trash, h, Ah);
        rc = orc_apply_weights(nb, block_sizes, lambda, invtt, m, Ah, 1);
        if (rc < 0) return rc;
        if (tm && (rc = orc_tmpl_filter(tm, tm_binv, Ah)) < 0) return rc;
        orc_trmatvec(m, nnz, pix, pixweights, trash, lcount, Ah, AtNm1Ah);
        double coeff = dot_seq(n, h, AtNm1Ah);
        double ro = g2pix / coeff;
        ORC_PFOR
        for (int j = 0; j < n; j++) x[j] = x[j] + ro * h[j];
        ORC_PFOR
        for (int j = 0; j < n; j++) g[j] = gp[j] - ro * AtNm1Ah[j];
        PRECOND_APPLY(g, Cg);
        double g2pixp = g2pix;
        g2pix = dot_seq(n, Cg, g);
        double g2pix_polak = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : g2pix_polak) schedule(static)
#endif
        for (int i = 0; i < n; i++) g2pix_polak += Cg[i] * (g[i] - gp[i]) * 1.0;
        res = dot_seq(n, g, g);
        if (res_hist && nh < maxiter) res_hist[nh++] = sqrt(res) / sqrt(res0);
        if (res <= tol2rel) {
            converged = 1;
            break;
        }
        if (stop_after > 0 && k >= stop_after) break;
        double gamma = g2pix_polak / g2pixp;
        ORC_PFOR
        for (int j = 0; j < n; j++) h[j] = h[j] * gamma + Cg[j];
    }
    rep->iterations = k;
    rep->converged = converged;
    rep->res0 = res0;
    rep->res = res;
    rep->n_map = n;
    rep->lcount = lcount;
    rep->trash_pix = trash;
    rep->n_cut = ncut;
    memcpy(lindices_out, lind, (size_t) lcount * sizeof(int));
    for (int q = 0; q < n / nnz; q++) {
        cond_out[q] = cond[q];
        lhits_out[q] = lhits[q];
    }
    free(lind);
    free(id_last);
    free(ll);
    free(lhits);
    free(cond);
    free(W);
    free(Winv);
    free(h);
    free(g);
    free(gp);
    free(AtNm1Ah);
    free(Ah);
    twolvl_free(&tl);
#undef PRECOND_APPLY
    return ORC_OK;
}
