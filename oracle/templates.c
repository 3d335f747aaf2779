/*
 * templates.c -- CPU restatement of MAPPRAISER's template marginalisation
 * (MTmap, mappraiser.c:387-700; PCG_GLS_templates, pcg_true.c:243-576; the
 * template classes of algebra/src/templates/templates.c).
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  The product library never links or
 * calls anything here.
 *
 * One rank.  Blocks are CES-major (block c*ndet + j = detector j of CES c) and
 * every block spans its whole CES.  The template amplitudes of a block follow
 * Tlist_init (templates.c:560-605):
 *   [poly order 0: nsweeps bins] ... [poly order npoly-1] [SSS: n_sss_bins]
 *   [HWP order 1 cos: nh] [order 1 sin: nh] ... [order nhwp sin: nh]
 * with nh = ceil(n_b / (delta_t * fs)).  A sample i of the block (i counted
 * from the CES start) belongs to
 *   poly  : sweep s with stamps[s] <= i < stamps[s+1], weight
 *           Legendre(i/fs, stamps[s]/fs, (stamps[s+1]-1)/fs, order)
 *           (expandpolydata, templates.c:693-707; Legendre/Pn, templates.c:531-552);
 *   SSS   : az bin az_binned[i], weight 1 (expandSSSdata, templates.c:710-716);
 *   HWP k : time bin floor(i / (fs * delta_t)), weights cos_k, sin_k from the
 *           recurrence of build_hwpss_w (templates.c:568-598) on cos/sin of the
 *           HWP angle (expandHWPSSdata, templates.c:719-733).
 * Samples outside [stamps[0], stamps[nsweeps]) have no polynomial template here;
 * the reference gives them bin 0 and weight 0, which adds 0 * tau[0].
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* Legendre polynomial of the given order at (2x-a-b)/(b-a), by the upward
 * three-term recurrence P_n = ((2n-1) t P_{n-1} - (n-1) P_{n-2}) / n evaluated
 * in the same operation order as Pn (templates.c:537-546); the recursive
 * calls there return the same values as this loop. */
static double legendre(double x, double a, double b, int order) {
    const double t = (2 * x - a - b) / (b - a);
    if (order == 0) return 1;
    double p0 = 1, p1 = t;
    for (int n = 2; n <= order; n++) {
        const double p2 = (double) ((2 * n - 1) * t * p1 - (n - 1) * p0) / n;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

typedef struct {
    int nb, ntot;
    int *nt;          /* per block */
    int64_t *off;     /* template offset per block */
    int64_t *t0;      /* first sample of each block */
    int *sw0;         /* offset of each CES in sweeptstamps */
    int64_t *c0;      /* offset of each CES in the per-CES sample arrays */
    int *nh;          /* per CES: HWP time bins */
} layout;

static void layout_free(layout *L) {
    free(L->nt);
    free(L->off);
    free(L->t0);
    free(L->sw0);
    free(L->c0);
    free(L->nh);
}

static int layout_build(const orc_tmpl *t, layout *L) {
    memset(L, 0, sizeof(*L));
    if (t->nces <= 0 || t->ndet <= 0 || t->npoly < 0 || t->nhwp < 0 || (t->ground != 0 && t->ground != 1))
        return ORC_EINVAL;
    if (t->ground && (t->n_sss_bins <= 0 || !t->az_binned)) return ORC_EINVAL;
    if (t->nhwp && (!(t->delta_t * t->sampling_freq > 0) || !t->hwp_cos || !t->hwp_sin)) return ORC_EINVAL;
    const int nb = t->nces * t->ndet;
    L->nb = nb;
    L->nt = (int *) calloc((size_t) nb, sizeof(int));
    L->off = (int64_t *) calloc((size_t) nb + 1, sizeof(int64_t));
    L->t0 = (int64_t *) calloc((size_t) nb + 1, sizeof(int64_t));
    L->sw0 = (int *) calloc((size_t) t->nces + 1, sizeof(int));
    L->c0 = (int64_t *) calloc((size_t) t->nces + 1, sizeof(int64_t));
    L->nh = (int *) calloc((size_t) t->nces, sizeof(int));
    if (!L->nt || !L->off || !L->t0 || !L->sw0 || !L->c0 || !L->nh) {
        layout_free(L);
        return ORC_ENOMEM;
    }
    for (int c = 0; c < t->nces; c++) {
        const int ns = t->nsweeps[c];
        if (ns < 0) return ORC_EINVAL;
        const int *st = t->sweeptstamps + L->sw0[c];
        const int len = t->block_sizes[c * t->ndet];
        if (st[ns] != len || st[0] < 0) return ORC_EINVAL; /* ces_length = stamps[nsweeps] (mappraiser.c:577) */
        for (int s = 0; s < ns; s++)
            if (st[s + 1] < st[s]) return ORC_EINVAL;
        L->sw0[c + 1] = L->sw0[c] + ns + 1;
        L->c0[c + 1] = L->c0[c] + len;
        L->nh[c] = t->nhwp ? (int) ceil(len / (t->delta_t * t->sampling_freq)) : 0;
        if (t->nhwp && len > 0 && (int) floor((double) (len - 1) / (t->sampling_freq * t->delta_t)) >= L->nh[c])
            return ORC_EINVAL; /* last time bin beyond the HWP amplitudes */
        for (int j = 0; j < t->ndet; j++) {
            const int b = c * t->ndet + j;
            if (t->block_sizes[b] != len) return ORC_EINVAL;
            L->nt[b] = t->npoly * ns + t->ground * t->n_sss_bins + 2 * t->nhwp * L->nh[c];
            L->off[b + 1] = L->off[b] + L->nt[b];
            L->t0[b + 1] = L->t0[b] + len;
        }
    }
    L->ntot = (int) L->off[nb];
    return ORC_OK;
}

int orc_bin_az(int n, const double *az, double az_min, double az_max, int nbins, int *out) {
    /* bin_az (templates.c:736-757) with miin = min */
    if (nbins <= 0) return ORC_EINVAL;
    for (int i = 0; i < n; i++) {
        const int k = (int) floor((az[i] - az_min) / ((az_max - az_min) / nbins));
        out[i] = k < nbins - 1 ? k : nbins - 1;
        if (out[i] < 0) return ORC_EINVAL;
    }
    return ORC_OK;
}

int orc_tmpl_sizes(const orc_tmpl *t, int *nt_block) {
    layout L;
    int rc = layout_build(t, &L);
    if (rc) {
        layout_free(&L);
        return rc;
    }
    if (nt_block) memcpy(nt_block, L.nt, (size_t) L.nb * sizeof(int));
    rc = L.ntot;
    layout_free(&L);
    return rc;
}

/* class weights and bins of one sample in class order: poly orders, SSS,
 * then HWP cos/sin pairs (no poly classes for a sample outside every sweep) */
typedef struct {
    int ncls;
    int64_t bin[64];
    double wt[64];
    int kind[64]; /* 0 poly, 1 SSS, 2 HWP */
} sample_cls;

static int sweep_of(const int *st, int ns, int i) {
    if (ns <= 0 || i < st[0] || i >= st[ns]) return -1;
    int lo = 0, hi = ns; /* st[lo] <= i < st[hi] */
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (st[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

static void classes_of(const orc_tmpl *t, const layout *L, int c, int b, int i, sample_cls *S) {
    const int ns = t->nsweeps[c];
    const int *st = t->sweeptstamps + L->sw0[c];
    const double fs = t->sampling_freq;
    int64_t base = L->off[b];
    int k = 0;
    const int s = sweep_of(st, ns, i);
    for (int o = 0; o < t->npoly; o++) {
        if (s >= 0) {
            S->bin[k] = base + (int64_t) o * ns + s;
            S->wt[k] = legendre((double) i / fs, (double) st[s] / fs, (double) (st[s + 1] - 1) / fs, o);
            S->kind[k] = 0;
            k++;
        }
    }
    base += (int64_t) t->npoly * ns;
    if (t->ground) {
        S->bin[k] = base + t->az_binned[L->c0[c] + i];
        S->wt[k] = 1.0;
        S->kind[k] = 1;
        k++;
        base += t->n_sss_bins;
    }
    if (t->nhwp) {
        const int nh = L->nh[c];
        const int tb = (int) floor((double) i / (t->sampling_freq * t->delta_t));
        const double c1 = t->hwp_cos[L->c0[c] + i], s1 = t->hwp_sin[L->c0[c] + i];
        double ck = c1, sk = s1;
        for (int h = 1; h <= t->nhwp; h++) {
            if (h > 1) {
                const double cn = ck * c1 - sk * s1;
                const double sn = sk * c1 + ck * s1;
                ck = cn;
                sk = sn;
            }
            S->bin[k] = base + (int64_t) (2 * (h - 1)) * nh + tb;
            S->wt[k] = ck;
            S->kind[k] = 2;
            k++;
            S->bin[k] = base + (int64_t) (2 * (h - 1) + 1) * nh + tb;
            S->wt[k] = sk;
            S->kind[k] = 2;
            k++;
        }
    }
    S->ncls = k;
}

static int check_limits(const orc_tmpl *t) {
    return (t->npoly + 1 + 2 * t->nhwp) <= 64 ? ORC_OK : ORC_EINVAL;
}

/* TVecProd (templates.c:27-103): per sample, the classes added in class order */
int orc_tmpl_tvec(const orc_tmpl *t, const double *tau, double *out) {
    layout L;
    int rc = layout_build(t, &L);
    if (!rc) rc = check_limits(t);
    if (rc) {
        layout_free(&L);
        return rc;
    }
    sample_cls S;
    for (int c = 0; c < t->nces; c++)
        for (int j = 0; j < t->ndet; j++) {
            const int b = c * t->ndet + j;
            for (int i = 0; i < t->block_sizes[b]; i++) {
                classes_of(t, &L, c, b, i, &S);
                double v = 0;
                for (int k = 0; k < S.ncls; k++) v += S.kind[k] == 1 ? tau[S.bin[k]] : S.wt[k] * tau[S.bin[k]];
                out[L.t0[b] + i] = v;
            }
        }
    layout_free(&L);
    return ORC_OK;
}

/* TrTVecProd (templates.c:106-188): out zeroed, then every class adds its
 * samples in increasing order (each bin belongs to one class, so the class
 * loop order does not change any sum) */
int orc_tmpl_trtvec(const orc_tmpl *t, const double *d, double *out) {
    layout L;
    int rc = layout_build(t, &L);
    if (!rc) rc = check_limits(t);
    if (rc) {
        layout_free(&L);
        return rc;
    }
    memset(out, 0, (size_t) L.ntot * sizeof(double));
    sample_cls S;
    for (int c = 0; c < t->nces; c++)
        for (int j = 0; j < t->ndet; j++) {
            const int b = c * t->ndet + j;
            for (int i = 0; i < t->block_sizes[b]; i++) {
                classes_of(t, &L, c, b, i, &S);
                const double v = d[L.t0[b] + i];
                for (int k = 0; k < S.ncls; k++) out[S.bin[k]] += S.kind[k] == 1 ? v : S.wt[k] * v;
            }
        }
    layout_free(&L);
    return ORC_OK;
}

/* BuildKernel (templates.c:191-436) for every block: B_b = w_b T_b^T T_b,
 * nt_b x nt_b row-major, blocks concatenated.  Entry (r, c) of classes (k, l)
 * sums over the block's samples in increasing order:
 *   poly/HWP x poly/HWP: (w * wt_k) * wt_l;  SSS x poly/HWP: w * wt_l;
 *   poly/HWP x SSS: w * wt_k;  SSS x SSS: w. */
int orc_tmpl_kernel(const orc_tmpl *t, double *B) {
    layout L;
    int rc = layout_build(t, &L);
    if (!rc) rc = check_limits(t);
    if (rc) {
        layout_free(&L);
        return rc;
    }
    int64_t boff = 0;
    sample_cls S;
    for (int c = 0; c < t->nces; c++)
        for (int j = 0; j < t->ndet; j++) {
            const int b = c * t->ndet + j;
            const int nt = L.nt[b];
            double *Bb = B + boff;
            memset(Bb, 0, (size_t) nt * nt * sizeof(double));
            const double w = t->w[b];
            for (int i = 0; i < t->block_sizes[b]; i++) {
                classes_of(t, &L, c, b, i, &S);
                for (int k = 0; k < S.ncls; k++)
                    for (int l = 0; l < S.ncls; l++) {
                        double v;
                        if (S.kind[k] != 1 && S.kind[l] != 1) v = w * S.wt[k] * S.wt[l];
                        else if (S.kind[k] == 1 && S.kind[l] != 1) v = w * S.wt[l];
                        else if (S.kind[k] != 1) v = w * S.wt[k];
                        else v = w;
                        Bb[(S.bin[k] - L.off[b]) * nt + (S.bin[l] - L.off[b])] += v;
                    }
            }
            boff += (int64_t) nt * nt;
        }
    layout_free(&L);
    return ORC_OK;
}

/* InvKernel (templates.c:449-465): the minimum-norm least-squares solution of
 * B X = I from LAPACKE_dgelss with rcond 1e-12, i.e. the pseudo-inverse with
 * singular values <= 1e-12 s_max dropped; ranks[b] = the effective rank. */
int orc_tmpl_invkernel(const orc_tmpl *t, const double *B, double *Binv, int *ranks) {
    layout L;
    int rc = layout_build(t, &L);
    if (rc) {
        layout_free(&L);
        return rc;
    }
    int64_t boff = 0;
    for (int b = 0; b < L.nb && !rc; b++) {
        const int nt = L.nt[b];
        int r = 0;
        if (nt > 0) rc = orc_pinv_rank(nt, B + boff, 1e-12, Binv + boff, &r);
        if (ranks) ranks[b] = r;
        boff += (int64_t) nt * nt;
    }
    layout_free(&L);
    return rc;
}

/* The template part of the noise weighting of PCG_GLS_templates
 * (pcg_true.c:318-371 and 426-458), on a TOD already multiplied by W:
 *   out1 = T^T tod;  out2 = Binv out1 (per block, row j summed over k in
 *   order);  Tvec = w_b * (T out2);  tod -= Tvec. */
int orc_tmpl_filter(const orc_tmpl *t, const double *Binv, double *tod) {
    layout L;
    int rc = layout_build(t, &L);
    if (rc) {
        layout_free(&L);
        return rc;
    }
    double *o1 = (double *) malloc((size_t) (L.ntot > 0 ? L.ntot : 1) * sizeof(double));
    double *o2 = (double *) malloc((size_t) (L.ntot > 0 ? L.ntot : 1) * sizeof(double));
    double *tv = (double *) malloc((size_t) (L.t0[L.nb] > 0 ? L.t0[L.nb] : 1) * sizeof(double));
    if (!o1 || !o2 || !tv) rc = ORC_ENOMEM;
    if (!rc) rc = orc_tmpl_trtvec(t, tod, o1);
    int64_t boff = 0;
    for (int b = 0; b < L.nb && !rc; b++) {
        const int nt = L.nt[b];
        const double *Bb = Binv + boff;
        for (int r = 0; r < nt; r++) {
            double v = 0;
            for (int k = 0; k < nt; k++) v += Bb[(int64_t) r * nt + k] * o1[L.off[b] + k];
            o2[L.off[b] + r] = v;
        }
        boff += (int64_t) nt * nt;
    }
    if (!rc) rc = orc_tmpl_tvec(t, o2, tv);
    for (int b = 0; b < L.nb && !rc; b++)
        for (int64_t i = L.t0[b]; i < L.t0[b + 1]; i++) {
            tv[i] = t->w[b] * tv[i];
            tod[i] = tod[i] - tv[i];
        }
    free(o1);
    free(o2);
    free(tv);
    layout_free(&L);
    return rc;
}

void orc_hwp_mod(int n, const double *angle, double *c, double *s) {
    for (int i = 0; i < n; i++) {
        c[i] = cos(angle[i]);
        s[i] = sin(angle[i]);
    }
}
