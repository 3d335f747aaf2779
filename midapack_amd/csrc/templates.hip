// templates.hip -- template marginalisation of MTmap / PCG_GLS_templates
// (mappraiser.c:387-700, pcg_true.c:243-576, algebra/src/templates/templates.c)
// on the device.
//
// With W the white-noise weights (one w_b per block, lambda 1) and T the
// template matrix, PCG_GLS_templates replaces N^-1 by
//   F = W - W T B^+ T^T W,   B = T^T W T   (block diagonal: one B_b per block),
// applied as: y = W (P h); a1 = T^T y; a2 = B^+ a1; y -= w_b (T a2).
//
// Template classes per block, in amplitude order (Tlist_init, templates.c:560-605):
// Legendre orders 0..npoly-1 per sweep (expandpolydata, templates.c:693-707),
// optionally one azimuth-binned ground class (expandSSSdata, 710-716), HWP
// harmonics 1..nhwp as cos/sin pairs per time bin of delta_t seconds
// (expandHWPSSdata and build_hwpss_w, templates.c:568-598, 719-733).
//
// MI355X layout.  Blocks span their CES (CES-major), so every per-sample table
// (sweep, azimuth bin, Legendre and HWP weights, sample orders) is stored once
// per CES and read by all its detectors from L2 / MALL.  Per PCG iteration:
//   k_tmpl_tr   T^T: one workgroup per (block, sweep | azimuth bin | time bin),
//               every class of the group accumulated in registers, fixed-order
//               workgroup sums (deterministic);
//   k_tmpl_binv a2 = B^+ a1: one wave per amplitude row, coalesced row reads;
//   k_tmpl_sub  y -= w_b T a2: one thread per CES sample looping over the CES's
//               detectors, so each sample's weights are loaded once; the classes
//               are added in class order exactly as TVecProd does.
// Setup: the tables come from the host (bin_az, the HWP cos/sin and the
// Legendre weights are the reference's own host-side expressions); every kernel
// entry B_b[r][c] is summed by one thread over its samples in increasing order
// (BuildKernel's order, so B is bit-identical to it); B^+ = V diag(1/lambda) V^T
// from rocSOLVER's batched symmetric eigensolver with dgelss's cut
// (|lambda| <= 1e-12 max |lambda| dropped), one eigensolve per CES: its
// detectors' kernels differ only by their weight w_b (B_b = w_b K_c).
#include "mgpu_internal.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

namespace mgpu {

constexpr int kTW = 256;      // threads of the T^T workgroups
constexpr int kSlots = 16;    // weighted classes per group: npoly <= 16, 2 nhwp <= 16
constexpr int kHwpSlot = 16;  // BTask weight slots >= this are HWP slots

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// sum over the kTW threads in a fixed order (the same on every call)
__device__ __forceinline__ double wg_sum(double v, double *red) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kTW / 64; k++) s += red[k];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// kernels: PCG iteration
// ---------------------------------------------------------------------------
// T^T (s_b y): one workgroup per item (a sweep, azimuth bin or time bin of a
// block; blocks in order, so the detectors of a CES follow each other and
// reuse its table entries from L2 / MALL); kind 0 sums leg[o] * v over a
// sweep for every order, kind 1 sums v over an azimuth bin (its samples listed
// in the azimuth order), kind 2 sums hc[h] * v over a time bin for every HWP
// slot, with v = s_b * y (s_b = T0[b], the weighting of apply_weights fused in;
// 1 when scale is NULL)
__global__ __launch_bounds__(kTW) void k_tmpl_tr(const TItem *__restrict__ items, const double *__restrict__ y,
                                                 const double *__restrict__ scale,
                                                 const int64_t *__restrict__ start, const int64_t *__restrict__ c0,
                                                 int ndet, const int *__restrict__ perm_az,
                                                 const double *__restrict__ leg, const double *__restrict__ hc,
                                                 int64_t ncs, int npoly, int nh2, const int64_t *__restrict__ off,
                                                 double *__restrict__ out, const int *__restrict__ state) {
    __shared__ double red[kTW / 64];
    if (state && *state) return;  // the solve has stopped: queued iterations do no work
    const TItem it = items[blockIdx.x];
    const int b = it.b;
    const int64_t cb = c0[b / ndet];
    const int nk = it.kind == 0 ? npoly : it.kind == 1 ? 1 : nh2;
    const double *__restrict__ yb = y + start[b];
    const double sc = scale ? scale[b] : 1.0;
    double acc[kSlots];
#pragma unroll
    for (int k = 0; k < kSlots; k++) acc[k] = 0.0;
    if (it.kind == 0) {
        for (int i = it.lo + (int) threadIdx.x; i < it.hi; i += kTW) {
            const double v = sc * yb[i];
#pragma unroll
            for (int k = 0; k < kSlots; k++)
                if (k < npoly) acc[k] += leg[(int64_t) k * ncs + cb + i] * v;
        }
    } else if (it.kind == 1) {
        for (int i = it.lo + (int) threadIdx.x; i < it.hi; i += kTW) acc[0] += sc * yb[perm_az[cb + i]];
    } else {
        for (int i = it.lo + (int) threadIdx.x; i < it.hi; i += kTW) {
            const double v = sc * yb[i];
#pragma unroll
            for (int k = 0; k < kSlots; k++)
                if (k < nh2) acc[k] += hc[(int64_t) k * ncs + cb + i] * v;
        }
    }
    const int64_t o = off[b] + it.rel;
#pragma unroll
    for (int k = 0; k < kSlots; k++) {
        if (k >= nk) break;
        const double s = wg_sum(acc[k], red);
        if (threadIdx.x == 0) out[o + (int64_t) k * it.stride] = s;
    }
}

// a2 = B^+ a1, one wave per amplitude row
__global__ __launch_bounds__(256) void k_tmpl_binv(const double *__restrict__ Binv, const int64_t *__restrict__ boff,
                                                   const int64_t *__restrict__ off,
                                                   const int *__restrict__ row_block, int64_t ntot,
                                                   const double *__restrict__ a1, double *__restrict__ a2,
                                                   const int *__restrict__ state) {
    if (state && *state) return;
    const int64_t r = (int64_t) blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= ntot) return;  // uniform per wave
    const int lane = threadIdx.x & 63;
    const int b = row_block[r];
    const int64_t o = off[b];
    const int nt = (int) (off[b + 1] - o);
    const double *__restrict__ row = Binv + boff[b] + (r - o) * nt;
    double s = 0.0;
    for (int k = lane; k < nt; k += 64) s += row[k] * a1[o + k];
    s = wave_sum(s);
    if (lane == 0) a2[r] = s;
}

// y[t] = v[t] - w_b (T a2)[t], v = in_scale_b q (the fused apply_weights) or q,
// for every detector of the CES at CES sample i; the classes are added to 0 in
// class order (TVecProd), then scaled by w_b (apply_weights) and subtracted
// (pcg_true.c:453-454).  INPLACE: q is y.
template <bool INPLACE>
__global__ __launch_bounds__(256) void k_tmpl_sub(const double *__restrict__ q, const double *__restrict__ in_scale,
                                                  double *__restrict__ y, const int64_t *__restrict__ start,
                                                  const double *__restrict__ T0, int ndet,
                                                  const int64_t *__restrict__ c0, const int *__restrict__ ns_c,
                                                  const int *__restrict__ nh_c, const int *__restrict__ sweep,
                                                  const int *__restrict__ az, const double *__restrict__ leg,
                                                  const double *__restrict__ hc, int64_t ncs, int npoly, int ground,
                                                  int nsss, int nh2, double fsdt, const int64_t *__restrict__ off,
                                                  const double *__restrict__ a2, const int *__restrict__ ces_of,
                                                  const int *__restrict__ len, const int *__restrict__ state) {
    if (state && *state) return;
    const int64_t g = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int c = ces_of[blockIdx.x];
    const int i = (int) (g - c0[c]);
    if (i >= len[c]) return;
    const int ns = ns_c[c], nh = nh_c[c];
    const int s = npoly ? sweep[g] : -1;
    double wl[kSlots], wh[kSlots];
#pragma unroll
    for (int k = 0; k < kSlots; k++) {
        wl[k] = (k < npoly && s >= 0) ? leg[(int64_t) k * ncs + g] : 0.0;
        wh[k] = k < nh2 ? hc[(int64_t) k * ncs + g] : 0.0;
    }
    const int a = ground ? az[g] : 0;
    const int tb = nh2 ? (int) floor((double) i / fsdt) : 0;
    const double *__restrict__ src = INPLACE ? y : q;
    for (int j = 0; j < ndet; j++) {
        const int b = c * ndet + j;
        int64_t base = off[b];
        double r = 0.0;
        if (s >= 0) {
#pragma unroll
            for (int k = 0; k < kSlots; k++)
                if (k < npoly) r += wl[k] * a2[base + (int64_t) k * ns + s];
        }
        base += (int64_t) npoly * ns;
        if (ground) {
            r += a2[base + a];
            base += nsss;
        }
#pragma unroll
        for (int k = 0; k < kSlots; k++)
            if (k < nh2) r += wh[k] * a2[base + (int64_t) k * nh + tb];
        const int64_t t = start[b] + i;
        const double tv = T0[b] * r;
        const double v = in_scale ? in_scale[b] * src[t] : src[t];
        y[t] = v - tv;
    }
}

// T a on its own (TVecProd): out[t] = (T a)[t]
__global__ __launch_bounds__(256) void k_tmpl_tvec(double *__restrict__ out, const int64_t *__restrict__ start,
                                                   int ndet, const int64_t *__restrict__ c0,
                                                   const int *__restrict__ ns_c, const int *__restrict__ nh_c,
                                                   const int *__restrict__ sweep, const int *__restrict__ az,
                                                   const double *__restrict__ leg, const double *__restrict__ hc,
                                                   int64_t ncs, int npoly, int ground, int nsss, int nh2,
                                                   double fsdt, const int64_t *__restrict__ off,
                                                   const double *__restrict__ a, const int *__restrict__ ces_of,
                                                   const int *__restrict__ len) {
    const int64_t g = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int c = ces_of[blockIdx.x];
    const int i = (int) (g - c0[c]);
    if (i >= len[c]) return;
    const int ns = ns_c[c], nh = nh_c[c];
    const int s = npoly ? sweep[g] : -1;
    const int tb = nh2 ? (int) floor((double) i / fsdt) : 0;
    for (int j = 0; j < ndet; j++) {
        const int b = c * ndet + j;
        int64_t base = off[b];
        double r = 0.0;
        if (s >= 0)
            for (int k = 0; k < npoly; k++) r += leg[(int64_t) k * ncs + g] * a[base + (int64_t) k * ns + s];
        base += (int64_t) npoly * ns;
        if (ground) {
            r += a[base + az[g]];
            base += nsss;
        }
        for (int k = 0; k < nh2; k++) r += hc[(int64_t) k * ncs + g] * a[base + (int64_t) k * nh + tb];
        out[start[b] + i] = r;
    }
}

// ---------------------------------------------------------------------------
// kernels: setup
// ---------------------------------------------------------------------------
__device__ __forceinline__ double slot_weight(int slot, const double *__restrict__ leg,
                                              const double *__restrict__ hc, int64_t ncs, int64_t g) {
    return slot < kHwpSlot ? leg[(int64_t) slot * ncs + g] : hc[(int64_t) (slot - kHwpSlot) * ncs + g];
}

// one kernel entry per thread (x), one detector of the CES per y: the sum runs
// over the entry's samples in increasing order, as BuildKernel's i loop
__global__ __launch_bounds__(64) void k_tmpl_kernel(const BTask *__restrict__ tasks, int ntask, int ces, int ndet,
                                                    const int64_t *__restrict__ c0, const int *__restrict__ perm,
                                                    const double *__restrict__ leg, const double *__restrict__ hc,
                                                    int64_t ncs, const double *__restrict__ T0,
                                                    const int64_t *__restrict__ off,
                                                    const int64_t *__restrict__ boff, double *__restrict__ B) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ntask) return;
    const BTask tk = tasks[x];
    const int b = ces * ndet + blockIdx.y;
    const int64_t cb = c0[ces];
    const double w = T0[b];
    double acc = 0.0;
    if (tk.kind == 0) {
        for (int i = tk.lo; i < tk.hi; i++) {
            const int64_t g = cb + i;
            acc += w * slot_weight(tk.ka, leg, hc, ncs, g) * slot_weight(tk.kb, leg, hc, ncs, g);
        }
    } else {
        const int *__restrict__ pm = perm + (int64_t) tk.perm * ncs + cb;
        if (tk.kind == 1) {
            for (int k = tk.lo; k < tk.hi; k++) acc += w * slot_weight(tk.kb, leg, hc, ncs, cb + pm[k]);
        } else {
            for (int k = tk.lo; k < tk.hi; k++) acc += w;
        }
    }
    const int nt = (int) (off[b + 1] - off[b]);
    double *__restrict__ Bb = B + boff[b];
    Bb[(int64_t) tk.row * nt + tk.col] = acc;
    if (tk.sym) Bb[(int64_t) tk.col * nt + tk.row] = acc;
}

// per block: the kept inverse eigenvalues (|lambda| > 1e-12 max |lambda|,
// dgelss's rcond) and the rank
__global__ __launch_bounds__(256) void k_eig_inv(double *__restrict__ D, int n, int *__restrict__ rank) {
    __shared__ double red[4];
    __shared__ int cnt[4];
    double *__restrict__ d = D + (int64_t) blockIdx.x * n;
    double mx = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) mx = fmax(mx, fabs(d[k]));
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    int kept = 0;
    for (int k = threadIdx.x; k < n; k += 256) {
        const double l = d[k];
        const bool keep = fabs(l) > 1e-12 * mx;
        d[k] = keep ? 1.0 / l : 0.0;
        kept += keep;
    }
    for (int o = 32; o > 0; o >>= 1) kept += __shfl_xor(kept, o);
    if ((threadIdx.x & 63) == 0) cnt[threadIdx.x >> 6] = kept;
    __syncthreads();
    if (threadIdx.x == 0) rank[blockIdx.x] = cnt[0] + cnt[1] + cnt[2] + cnt[3];
}

// Binv[r][c] = sum_k V[r][k] inv[k] V[c][k] for one block (blockIdx.z); V is
// rocSOLVER's column-major eigenvector matrix (eigenvector k at V + k n)
constexpr int kPT = 16;
__global__ __launch_bounds__(kPT *kPT) void k_pinv_form(const double *__restrict__ V, const double *__restrict__ inv,
                                                        int n, double *__restrict__ Binv) {
    __shared__ double sr[kPT][kPT + 1], sc[kPT][kPT + 1];
    const int z = blockIdx.z;
    const double *__restrict__ Vz = V + (int64_t) z * n * n;
    const double *__restrict__ iz = inv + (int64_t) z * n;
    const int r0 = blockIdx.y * kPT, c0 = blockIdx.x * kPT;
    const int tr = threadIdx.x / kPT, tc = threadIdx.x % kPT;
    double acc = 0.0;
    for (int k0 = 0; k0 < n; k0 += kPT) {
        // sr[kk][i] = V[r0+i][k0+kk] * inv[k0+kk]; sc[kk][j] = V[c0+j][k0+kk]
        const int kk = tr, i = tc;
        const int k = k0 + kk;
        sr[kk][i] = (k < n && r0 + i < n) ? Vz[(int64_t) k * n + r0 + i] * iz[k] : 0.0;
        sc[kk][i] = (k < n && c0 + i < n) ? Vz[(int64_t) k * n + c0 + i] : 0.0;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPT; q++) acc += sr[q][tr] * sc[q][tc];
        __syncthreads();
    }
    if (r0 + tr < n && c0 + tc < n) Binv[(int64_t) z * n * n + (int64_t) (r0 + tr) * n + c0 + tc] = acc;
}

// Binv_b = sc[j] * P_rep[j] for the j-th block of a run (blockIdx.y): the
// pseudo-inverse of B_b = w_b K from that of its CES's representative block
__global__ __launch_bounds__(256) void k_pinv_scale(const double *__restrict__ P, const int *__restrict__ rep,
                                                    const double *__restrict__ sc, int64_t nn,
                                                    double *__restrict__ Binv) {
    const int j = blockIdx.y;
    const double *__restrict__ src = P + (int64_t) rep[j] * nn;
    double *__restrict__ dst = Binv + (int64_t) j * nn;
    const double f = sc[j];
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (int64_t) gridDim.x * blockDim.x)
        dst[i] = f * src[i];
}

// ---------------------------------------------------------------------------
// host: tables
// ---------------------------------------------------------------------------
namespace {

// Legendre polynomial of the given order at (2x-a-b)/(b-a): the upward
// recurrence in the operation order of Pn (templates.c:537-552)
double legendre(double x, double a, double b, int order) {
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

// stable counting sort of positions [lo, hi) by key (key[i] in [0, nk))
void stable_by_key(const int *key, int lo, int hi, int nk, std::vector<int> &cnt, int *out) {
    cnt.assign((size_t) nk + 1, 0);
    for (int i = lo; i < hi; i++) cnt[(size_t) key[i] + 1]++;
    for (int k = 0; k < nk; k++) cnt[(size_t) k + 1] += cnt[k];
    for (int i = lo; i < hi; i++) out[lo + cnt[key[i]]++] = i;
}

struct CesTables {
    int len = 0, ns = 0, nh = 0;
    std::vector<int> stamps, sweep, az, tbs;  // tbs: nh+1 time-bin boundaries
    std::vector<int> perm[3];                 // azimuth, (sweep, az), (time bin, az) orders
    std::vector<double> leg, hc;              // order-major
};

}  // namespace

// argument check inside a worker thread: the message goes to err (the
// library's last-error slot is per thread), the caller reports it
#define TARG(cond, ...)                                                          \
    do {                                                                         \
        if (!(cond)) {                                                           \
            char b_[256];                                                        \
            std::snprintf(b_, sizeof(b_), __VA_ARGS__);                          \
            err = b_;                                                            \
            return MGPU_EINVAL;                                                  \
        }                                                                        \
    } while (0)

static int ces_tables(const mgpu_templates *t, int c, int len, double fsdt, CesTables &T, std::string &err) {
    const int npoly = t->npoly, nsss = t->n_sss_bins, nh2 = 2 * t->nhwp;
    T.len = len;
    T.ns = t->nsweeps[c];
    TARG(T.ns >= 0 && t->sweeptstamps && t->sweeptstamps[c], "CES %d: invalid sweeps", c);
    T.stamps.assign(t->sweeptstamps[c], t->sweeptstamps[c] + T.ns + 1);
    TARG(T.stamps[T.ns] == len, "CES %d: last sweep stamp %d differs from the CES length %d (mappraiser.c:577)",
             c, T.stamps[T.ns], len);
    TARG(T.stamps[0] >= 0, "CES %d: negative sweep stamp", c);
    for (int s = 0; s < T.ns; s++) TARG(T.stamps[s + 1] >= T.stamps[s], "CES %d: decreasing sweep stamps", c);
    T.sweep.assign((size_t) len, -1);
    for (int s = 0; s < T.ns; s++)
        for (int i = T.stamps[s]; i < T.stamps[s + 1]; i++) T.sweep[i] = s;
    const double fs = t->sampling_freq;
    if (npoly) {
        T.leg.assign((size_t) npoly * len, 0.0);
        for (int s = 0; s < T.ns; s++) {
            const int a = T.stamps[s], b = T.stamps[s + 1];
            for (int i = a; i < b; i++)
                for (int o = 0; o < npoly; o++)
                    T.leg[(size_t) o * len + i] =
                            legendre((double) i / fs, (double) a / fs, (double) (b - 1) / fs, o);
        }
    }
    T.az.assign((size_t) len, 0);
    if (t->ground) {
        TARG(t->az && t->az[c] && t->az_min && t->az_max, "CES %d: azimuth missing", c);
        const double lo = t->az_min[c], width = (t->az_max[c] - t->az_min[c]) / nsss;
        for (int i = 0; i < len; i++) {
            const int k = (int) floor((t->az[c][i] - lo) / width);  // bin_az, templates.c:746
            T.az[i] = std::min(nsss - 1, k);
            TARG(T.az[i] >= 0, "CES %d sample %d: azimuth below az_min", c, i);
        }
    }
    T.nh = 0;
    if (nh2) {
        TARG(t->hwp_angle && t->hwp_angle[c], "CES %d: HWP angle missing", c);
        T.nh = (int) ceil(len / (t->delta_t * t->sampling_freq));
        T.tbs.assign((size_t) T.nh + 1, len);
        int prev = -1;
        for (int i = 0; i < len; i++) {
            const int tb = (int) floor((double) i / fsdt);
            TARG(tb < T.nh, "CES %d: time bin %d beyond the %d HWP amplitudes", c, tb, T.nh);
            for (int q = prev + 1; q <= tb; q++) T.tbs[q] = i;
            prev = tb;
        }
        for (int q = prev + 1; q <= T.nh; q++) T.tbs[q] = len;
        T.hc.assign((size_t) nh2 * len, 0.0);
        for (int i = 0; i < len; i++) {
            const double c1 = std::cos(t->hwp_angle[c][i]), s1 = std::sin(t->hwp_angle[c][i]);
            double ck = c1, sk = s1;
            for (int h = 0; h < t->nhwp; h++) {
                if (h > 0) {  // build_hwpss_w's recurrence
                    const double cn = ck * c1 - sk * s1;
                    const double sn = sk * c1 + ck * s1;
                    ck = cn;
                    sk = sn;
                }
                T.hc[(size_t) (2 * h) * len + i] = ck;
                T.hc[(size_t) (2 * h + 1) * len + i] = sk;
            }
        }
    }
    // sample orders: by azimuth; by azimuth inside each sweep; inside each time bin
    std::vector<int> cnt;
    for (auto &p : T.perm) p.assign((size_t) len, 0);
    if (t->ground) {
        stable_by_key(T.az.data(), 0, len, nsss, cnt, T.perm[0].data());
        for (int s = 0; s < T.ns; s++) stable_by_key(T.az.data(), T.stamps[s], T.stamps[s + 1], nsss, cnt,
                                                     T.perm[1].data());
        for (int q = 0; q < T.nh; q++) stable_by_key(T.az.data(), T.tbs[q], T.tbs[q + 1], nsss, cnt,
                                                     T.perm[2].data());
    }
    return MGPU_OK;
}
#undef TARG

// kernel entry tasks of one CES (entries shared by its detectors)
static void ces_tasks(const mgpu_templates *t, const CesTables &T, std::vector<BTask> &tk) {
    const int npoly = t->npoly, nsss = t->n_sss_bins, nh2 = 2 * t->nhwp, ns = T.ns, nh = T.nh;
    const int pS = npoly * ns, pH = pS + t->ground * nsss;  // class bases
    auto add = [&](int kind, int row, int col, int lo, int hi, int ka, int kb, int perm, int sym) {
        if (hi > lo) tk.push_back(BTask{kind, row, col, lo, hi, ka, kb, perm, sym, 0});
    };
    // poly x poly: same sweep
    for (int s = 0; s < ns; s++)
        for (int o1 = 0; o1 < npoly; o1++)
            for (int o2 = 0; o2 < npoly; o2++)
                add(0, o1 * ns + s, o2 * ns + s, T.stamps[s], T.stamps[s + 1], o1, o2, 0, 0);
    // HWP x HWP: same time bin
    for (int q = 0; q < nh; q++)
        for (int h1 = 0; h1 < nh2; h1++)
            for (int h2 = 0; h2 < nh2; h2++)
                add(0, pH + h1 * nh + q, pH + h2 * nh + q, T.tbs[q], T.tbs[q + 1], kHwpSlot + h1, kHwpSlot + h2, 0,
                    0);
    // poly x HWP and HWP x poly: pieces where a sweep and a time bin overlap
    if (npoly && nh2) {
        for (int s = 0; s < ns; s++) {
            const int a = T.stamps[s], b = T.stamps[s + 1];
            if (b <= a) continue;
            const int q0 = (int) floor((double) a / (double) (t->sampling_freq * t->delta_t));
            for (int q = std::max(0, q0 - 1); q < nh && T.tbs[q] < b; q++) {
                const int lo = std::max(a, T.tbs[q]), hi = std::min(b, T.tbs[q + 1]);
                if (hi <= lo) continue;
                for (int o = 0; o < npoly; o++)
                    for (int h = 0; h < nh2; h++) {
                        add(0, o * ns + s, pH + h * nh + q, lo, hi, o, kHwpSlot + h, 0, 0);
                        add(0, pH + h * nh + q, o * ns + s, lo, hi, kHwpSlot + h, o, 0, 0);
                    }
            }
        }
    }
    if (t->ground) {
        // SSS x SSS: the azimuth bin's samples
        auto segments = [&](const std::vector<int> &perm, int lo, int hi, std::vector<std::pair<int, int>> &seg) {
            seg.clear();
            int k = lo;
            while (k < hi) {
                const int a = T.az[perm[k]];
                int e = k + 1;
                while (e < hi && T.az[perm[e]] == a) e++;
                seg.push_back({k, e});
                k = e;
            }
        };
        std::vector<std::pair<int, int>> seg;
        segments(T.perm[0], 0, T.len, seg);
        for (auto &p : seg) {
            const int a = T.az[T.perm[0][p.first]];
            add(2, pS + a, pS + a, p.first, p.second, 0, 0, 0, 0);
        }
        // SSS x poly (mirrored): samples of a sweep in an azimuth bin
        for (int s = 0; s < ns && npoly; s++) {
            segments(T.perm[1], T.stamps[s], T.stamps[s + 1], seg);
            for (auto &p : seg) {
                const int a = T.az[T.perm[1][p.first]];
                for (int o = 0; o < npoly; o++) add(1, pS + a, o * ns + s, p.first, p.second, 0, o, 1, 1);
            }
        }
        // SSS x HWP (mirrored): samples of a time bin in an azimuth bin
        for (int q = 0; q < nh; q++) {
            segments(T.perm[2], T.tbs[q], T.tbs[q + 1], seg);
            for (auto &p : seg) {
                const int a = T.az[T.perm[2][p.first]];
                for (int h = 0; h < nh2; h++) add(1, pS + a, pH + h * nh + q, p.first, p.second, 0, kHwpSlot + h, 2, 1);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// build
// ---------------------------------------------------------------------------
#define MGPU_ROCBLAS(expr)                                                                     \
    do {                                                                                       \
        rocblas_status st_ = (expr);                                                           \
        if (st_ != rocblas_status_success) {                                                   \
            set_error("%s:%d: %s failed: status %d", __FILE__, __LINE__, #expr, (int) st_);  \
            return MGPU_EHIP;                                                                  \
        }                                                                                      \
    } while (0)

namespace {
struct RocblasHandle {
    rocblas_handle h = nullptr;
    ~RocblasHandle() {
        if (h) (void) rocblas_destroy_handle(h);
    }
};
}  // namespace

int templates_build(Problem &pb, const mgpu_templates *t) {
    MGPU_ARG(t, "NULL templates");
    const ToeplitzPlan &T = pb.tp;
    MGPU_ARG(T.lambda == 1, "templates need the white-noise model (lambda = 1, as MTmap), the problem has lambda %d",
             T.lambda);
    MGPU_ARG(pb.opt.precision == MGPU_FP64, "templates are solved in FP64 only");
    MGPU_ARG(pb.sg.nseg == 1, "templates solve one problem at a time");
    MGPU_ARG(t->nces > 0 && T.nb % t->nces == 0, "nb = %d blocks is not a multiple of nces = %d", T.nb, t->nces);
    MGPU_ARG(t->npoly >= 0 && t->npoly <= kSlots, "npoly = %d outside [0, %d]", t->npoly, kSlots);
    MGPU_ARG(t->nhwp >= 0 && 2 * t->nhwp <= kSlots, "nhwp = %d outside [0, %d]", t->nhwp, kSlots / 2);
    MGPU_ARG(t->ground == 0 || t->ground == 1, "ground must be 0 or 1");
    MGPU_ARG(!t->ground || t->n_sss_bins > 0, "n_sss_bins must be positive");
    MGPU_ARG(!t->nhwp || (t->delta_t * t->sampling_freq > 0), "delta_t * sampling_freq must be positive");
    MGPU_ARG(t->sampling_freq > 0, "sampling_freq must be positive");
    MGPU_ARG(t->nsweeps, "NULL nsweeps");
    MGPU_RET(ensure_work(pb));  // work vectors and the timing events
    Templates &M = pb.tm;
    M = Templates{};
    hipStream_t s = pb.stream;
    const int nces = t->nces, ndet = T.nb / nces, nb = T.nb;
    const double fsdt = t->sampling_freq * t->delta_t;
    M.npoly = t->npoly;
    M.ground = t->ground;
    M.nsss = t->ground ? t->n_sss_bins : 0;
    M.nhwp = t->nhwp;
    M.fsdt = fsdt;
    M.nces = nces;
    M.ndet = ndet;
    // per-CES tables (host), the CES in parallel
    std::vector<CesTables> ct((size_t) nces);
    std::vector<int> lens((size_t) nces);
    for (int c = 0; c < nces; c++) {
        const int64_t len = T.h_start[(size_t) c * ndet + 1] - T.h_start[(size_t) c * ndet];
        for (int j = 0; j < ndet; j++) {
            const int b = c * ndet + j;
            MGPU_ARG(T.h_start[b + 1] - T.h_start[b] == len, "block %d: length differs from its CES's", b);
        }
        lens[c] = (int) len;
    }
    {
        std::vector<int> rcs((size_t) nces, MGPU_OK);
        std::vector<std::string> errs((size_t) nces);
        std::atomic<int> next{0};
        const int nth = (int) std::max(1u, std::min({16u, std::thread::hardware_concurrency(), (unsigned) nces}));
        std::vector<std::thread> th;
        for (int k = 0; k < nth; k++)
            th.emplace_back([&] {
                for (int c; (c = next++) < nces;) rcs[c] = ces_tables(t, c, lens[c], fsdt, ct[c], errs[c]);
            });
        for (auto &x : th) x.join();
        for (int c = 0; c < nces; c++) MGPU_ARG(rcs[c] == MGPU_OK, "%s", errs[c].c_str());
    }
    // amplitude layout per block (Tlist_init)
    M.h_nt.assign((size_t) nb, 0);
    M.h_off.assign((size_t) nb + 1, 0);
    M.h_boff.assign((size_t) nb + 1, 0);
    for (int b = 0; b < nb; b++) {
        const CesTables &C = ct[b / ndet];
        M.h_nt[b] = M.npoly * C.ns + M.nsss + 2 * M.nhwp * C.nh;
        M.h_off[b + 1] = M.h_off[b] + M.h_nt[b];
        M.h_boff[b + 1] = M.h_boff[b] + (int64_t) M.h_nt[b] * M.h_nt[b];
    }
    M.ntot = M.h_off[nb];
    M.nB = M.h_boff[nb];
    // per-CES tables on the device with workgroup-aligned CES starts c0w[c], so
    // that thread g of the sample kernels is CES sample g - c0w[c]; ces_of
    // names the CES of every sample-kernel workgroup
    std::vector<int> ces_of;
    std::vector<int64_t> c0w((size_t) nces + 1, 0);
    for (int c = 0; c < nces; c++) {
        const int64_t nwg = (lens[c] + 255) / 256;
        for (int64_t k = 0; k < nwg; k++) ces_of.push_back(c);
        c0w[c + 1] = c0w[c] + nwg * 256;
    }
    M.ncs = c0w[nces];
    const int64_t ncs = M.ncs;
    const int nh2 = 2 * M.nhwp;
    std::vector<int> h_ns(nces), h_nh(nces), sweep((size_t) ncs, -1), az((size_t) ncs, 0),
            perm((size_t) 3 * ncs, 0);
    std::vector<double> leg((size_t) M.npoly * ncs, 0.0), hc((size_t) nh2 * ncs, 0.0);
    for (int c = 0; c < nces; c++) {
        const CesTables &C = ct[c];
        const int64_t o = c0w[c];
        h_ns[c] = C.ns;
        h_nh[c] = C.nh;
        std::copy(C.sweep.begin(), C.sweep.end(), sweep.begin() + o);
        std::copy(C.az.begin(), C.az.end(), az.begin() + o);
        for (int k = 0; k < 3; k++) std::copy(C.perm[k].begin(), C.perm[k].end(), perm.begin() + k * ncs + o);
        for (int r = 0; r < M.npoly; r++)
            std::copy(C.leg.begin() + (size_t) r * C.len, C.leg.begin() + (size_t) (r + 1) * C.len,
                      leg.begin() + r * ncs + o);
        for (int r = 0; r < nh2; r++)
            std::copy(C.hc.begin() + (size_t) r * C.len, C.hc.begin() + (size_t) (r + 1) * C.len,
                      hc.begin() + r * ncs + o);
    }
    std::vector<int> row_block((size_t) std::max<int64_t>(M.ntot, 1), 0);
    for (int b = 0; b < nb; b++)
        for (int64_t r = M.h_off[b]; r < M.h_off[b + 1]; r++) row_block[r] = b;
    // T^T items of every block in block order: sweeps, azimuth bins, time bins
    // (rel: the output offset inside the block's amplitudes)
    std::vector<TItem> items;
    for (int b = 0; b < nb; b++) {
        const CesTables &C = ct[b / ndet];
        if (M.npoly)
            for (int sw = 0; sw < C.ns; sw++) items.push_back(TItem{b, 0, C.stamps[sw], C.stamps[sw + 1], sw, C.ns, 0});
        if (M.ground) {
            // azimuth bin a occupies a contiguous range of the azimuth order
            int k = 0;
            for (int a = 0; a < M.nsss; a++) {
                int e = k;
                while (e < C.len && C.az[C.perm[0][e]] == a) e++;
                items.push_back(TItem{b, 1, k, e, (int64_t) M.npoly * C.ns + a, 0, 0});
                k = e;
            }
        }
        if (nh2)
            for (int q = 0; q < C.nh; q++)
                items.push_back(TItem{b, 2, C.tbs[q], C.tbs[q + 1], (int64_t) M.npoly * C.ns + M.nsss + q, C.nh, 0});
    }
    M.nitems = (int) items.size();
    MGPU_RET(pb.arena.alloc(&M.d_ns, (size_t) nces));
    MGPU_RET(pb.arena.alloc(&M.d_nh, (size_t) nces));
    MGPU_RET(pb.arena.alloc(&M.d_len, (size_t) nces));
    MGPU_RET(pb.arena.alloc(&M.d_c0, (size_t) nces + 1));
    MGPU_RET(pb.arena.alloc(&M.d_sweep, (size_t) M.ncs));
    MGPU_RET(pb.arena.alloc(&M.d_az, (size_t) M.ncs));
    MGPU_RET(pb.arena.alloc(&M.d_perm, (size_t) 3 * M.ncs));
    MGPU_RET(pb.arena.alloc(&M.d_leg, (size_t) std::max<int64_t>(1, M.npoly * M.ncs)));
    MGPU_RET(pb.arena.alloc(&M.d_hc, (size_t) std::max<int64_t>(1, nh2 * M.ncs)));
    MGPU_RET(pb.arena.alloc(&M.d_off, (size_t) nb + 1));
    MGPU_RET(pb.arena.alloc(&M.d_boff, (size_t) nb + 1));
    MGPU_RET(pb.arena.alloc(&M.d_row_block, (size_t) row_block.size()));
    MGPU_RET(pb.arena.alloc(&M.d_B, (size_t) std::max<int64_t>(1, M.nB)));
    MGPU_RET(pb.arena.alloc(&M.d_Binv, (size_t) std::max<int64_t>(1, M.nB)));
    MGPU_RET(pb.arena.alloc(&M.d_a1, (size_t) std::max<int64_t>(1, M.ntot)));
    MGPU_RET(pb.arena.alloc(&M.d_a2, (size_t) std::max<int64_t>(1, M.ntot)));
    MGPU_RET(pb.arena.alloc(&M.d_items, (size_t) std::max(1, M.nitems)));
    MGPU_RET(pb.arena.alloc(&M.d_ces_of, ces_of.size()));
    M.nwg_samples = (int) ces_of.size();
    auto up = [&](void *d, const void *h, size_t bytes) -> int {
        MGPU_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
        return MGPU_OK;
    };
    MGPU_RET(up(M.d_ns, h_ns.data(), sizeof(int) * nces));
    MGPU_RET(up(M.d_nh, h_nh.data(), sizeof(int) * nces));
    MGPU_RET(up(M.d_len, lens.data(), sizeof(int) * nces));
    MGPU_RET(up(M.d_c0, c0w.data(), sizeof(int64_t) * (nces + 1)));
    MGPU_RET(up(M.d_sweep, sweep.data(), sizeof(int) * sweep.size()));
    MGPU_RET(up(M.d_az, az.data(), sizeof(int) * az.size()));
    MGPU_RET(up(M.d_perm, perm.data(), sizeof(int) * perm.size()));
    if (!leg.empty()) MGPU_RET(up(M.d_leg, leg.data(), sizeof(double) * leg.size()));
    if (!hc.empty()) MGPU_RET(up(M.d_hc, hc.data(), sizeof(double) * hc.size()));
    MGPU_RET(up(M.d_off, M.h_off.data(), sizeof(int64_t) * (nb + 1)));
    MGPU_RET(up(M.d_boff, M.h_boff.data(), sizeof(int64_t) * (nb + 1)));
    MGPU_RET(up(M.d_row_block, row_block.data(), sizeof(int) * row_block.size()));
    if (!items.empty()) MGPU_RET(up(M.d_items, items.data(), sizeof(TItem) * items.size()));
    MGPU_RET(up(M.d_ces_of, ces_of.data(), sizeof(int) * ces_of.size()));
    MGPU_HIP(hipEventRecord(pb.ev0, s));
    // kernels B_b: every entry summed by one thread in sample order
    if (M.nB > 0) MGPU_HIP(hipMemsetAsync(M.d_B, 0, sizeof(double) * M.nB, s));
    for (int c = 0; c < nces; c++) {
        std::vector<BTask> tk;
        ces_tasks(t, ct[c], tk);
        if (tk.empty()) continue;
        BTask *d_tk = nullptr;
        MGPU_RET(pb.arena.alloc(&d_tk, tk.size()));
        MGPU_RET(up(d_tk, tk.data(), sizeof(BTask) * tk.size()));
        // the task kernels read the per-CES tables at the aligned CES start
        k_tmpl_kernel<<<dim3(grid_for((int64_t) tk.size(), 64), ndet), 64, 0, s>>>(
                d_tk, (int) tk.size(), c, ndet, M.d_c0, M.d_perm, M.d_leg, M.d_hc, M.ncs, T.d_T0, M.d_off, M.d_boff,
                M.d_B);
        MGPU_LAUNCHED();
        MGPU_HIP(hipStreamSynchronize(s));
        pb.arena.release(d_tk);
    }
    // Pseudo-inverses.  Every entry of B_b is a sum of w_b times products of the
    // CES's template values, so B_b = w_b K_c for all detectors b of CES c: one
    // eigensolve per CES (its first block of non-zero weight, the representative
    // r), and B_b^+ = (w_r / w_b) B_r^+ with the rank of B_r (dgelss's cut is
    // relative, so it keeps the same eigenvalues); a block of weight 0 has
    // B_b = 0, B_b^+ = 0 and rank 0.  One batched eigensolve per run of
    // consecutive CES with the same template count, each run's working copies
    // kept under kEigBatchBytes.
    M.h_rank.assign((size_t) nb, 0);
    if (M.nB > 0) {
        constexpr size_t kEigBatchBytes = size_t(4) << 30;
        std::vector<double> w((size_t) nb);
        MGPU_HIP(hipMemcpyAsync(w.data(), T.d_T0, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
        MGPU_HIP(hipStreamSynchronize(s));
        RocblasHandle rh;
        MGPU_ROCBLAS(rocblas_create_handle(&rh.h));
        MGPU_ROCBLAS(rocblas_set_stream(rh.h, s));
        int *d_rank = nullptr, *d_info = nullptr, *d_rep = nullptr;
        double *d_sc = nullptr;
        MGPU_RET(pb.arena.alloc(&d_rank, (size_t) nces));
        MGPU_RET(pb.arena.alloc(&d_info, (size_t) nces));
        MGPU_RET(pb.arena.alloc(&d_rep, (size_t) nb));
        MGPU_RET(pb.arena.alloc(&d_sc, (size_t) nb));
        for (int c = 0; c < nces;) {
            const int b0 = c * ndet;
            const int n = M.h_nt[b0];
            const size_t per_ces = (size_t) 2 * n * n * sizeof(double);  // eigenvectors and pseudo-inverse
            int e = c + 1;
            while (e < nces && M.h_nt[(size_t) e * ndet] == n && (size_t) (e + 1 - c) * per_ces <= kEigBatchBytes) e++;
            const int nrep = e - c, nblk = nrep * ndet;
            const int c_first = c;
            c = e;
            if (n == 0) continue;
            const int64_t nn = (int64_t) n * n;
            double *V = nullptr, *D = nullptr, *E = nullptr, *Pinv = nullptr;
            MGPU_RET(pb.arena.alloc(&V, (size_t) nrep * nn));
            MGPU_RET(pb.arena.alloc(&D, (size_t) nrep * n));
            MGPU_RET(pb.arena.alloc(&E, (size_t) nrep * n));
            MGPU_RET(pb.arena.alloc(&Pinv, (size_t) nrep * nn));
            std::vector<int> rep_blk((size_t) nrep), rep_of((size_t) nblk);
            std::vector<double> sc((size_t) nblk);
            for (int k = 0; k < nrep; k++) {
                const int cb = (c_first + k) * ndet;
                int r = cb;
                for (int j2 = 0; j2 < ndet; j2++)
                    if (w[(size_t) cb + j2] != 0.0) {
                        r = cb + j2;
                        break;
                    }
                rep_blk[k] = r;
                for (int j2 = 0; j2 < ndet; j2++) {
                    const int b = cb + j2;
                    rep_of[(size_t) (b - b0)] = k;
                    sc[(size_t) (b - b0)] = (w[b] != 0.0 && w[r] != 0.0) ? w[r] / w[b] : 0.0;
                }
                MGPU_HIP(hipMemcpyAsync(V + k * nn, M.d_B + M.h_boff[r], sizeof(double) * nn, hipMemcpyDeviceToDevice,
                                        s));
            }
            MGPU_HIP(hipMemcpyAsync(d_rep, rep_of.data(), sizeof(int) * nblk, hipMemcpyHostToDevice, s));
            MGPU_HIP(hipMemcpyAsync(d_sc, sc.data(), sizeof(double) * nblk, hipMemcpyHostToDevice, s));
            MGPU_ROCBLAS(rocsolver_dsyevd_strided_batched(rh.h, rocblas_evect_original, rocblas_fill_upper, n, V, n,
                                                          (rocblas_stride) nn, D, n, E, n, d_info, nrep));
            std::vector<int> info((size_t) nrep), rank((size_t) nrep);
            MGPU_HIP(hipMemcpyAsync(info.data(), d_info, sizeof(int) * nrep, hipMemcpyDeviceToHost, s));
            MGPU_HIP(hipStreamSynchronize(s));
            for (int k = 0; k < nrep; k++)
                MGPU_ARG(info[k] == 0, "eigensolve of template kernel %d failed (info %d)", rep_blk[k], info[k]);
            k_eig_inv<<<nrep, 256, 0, s>>>(D, n, d_rank);
            MGPU_LAUNCHED();
            const unsigned tiles = (unsigned) ((n + kPT - 1) / kPT);
            k_pinv_form<<<dim3(tiles, tiles, nrep), kPT * kPT, 0, s>>>(V, D, n, Pinv);
            MGPU_LAUNCHED();
            k_pinv_scale<<<dim3((unsigned) std::min<int64_t>(grid_for(nn, 256), 1024), nblk), 256, 0, s>>>(
                    Pinv, d_rep, d_sc, nn, M.d_Binv + M.h_boff[b0]);
            MGPU_LAUNCHED();
            MGPU_HIP(hipMemcpyAsync(rank.data(), d_rank, sizeof(int) * nrep, hipMemcpyDeviceToHost, s));
            MGPU_HIP(hipStreamSynchronize(s));
            for (int b = b0; b < b0 + nblk; b++)
                M.h_rank[(size_t) b] = w[b] != 0.0 ? rank[(size_t) rep_of[(size_t) (b - b0)]] : 0;
            pb.arena.release(V);
            pb.arena.release(D);
            pb.arena.release(E);
            pb.arena.release(Pinv);
        }
    }
    MGPU_HIP(hipEventRecord(pb.ev1, s));
    MGPU_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    MGPU_HIP(hipEventElapsedTime(&ms, pb.ev0, pb.ev1));
    M.build_ms = ms;
    M.on = true;
    return MGPU_OK;
}

// ---------------------------------------------------------------------------
// application
// ---------------------------------------------------------------------------
static int launch_tr(Problem &pb, const double *d_y, const double *scale, double *d_out) {
    Templates &M = pb.tm;
    if (M.nitems == 0) return MGPU_OK;
    k_tmpl_tr<<<(unsigned) M.nitems, kTW, 0, pb.stream>>>(M.d_items, d_y, scale, pb.tp.d_start, M.d_c0, M.ndet,
                                                          M.d_perm, M.d_leg, M.d_hc, M.ncs, M.npoly, 2 * M.nhwp,
                                                          M.d_off, d_out, pb.live_state);
    MGPU_LAUNCHED();
    return MGPU_OK;
}

int templates_trtvec(Problem &pb, const double *d_y, double *d_out) {
    MGPU_ARG(pb.tm.on, "no templates attached");
    return launch_tr(pb, d_y, nullptr, d_out);
}

int templates_tvec(Problem &pb, const double *d_a, double *d_out) {
    Templates &M = pb.tm;
    MGPU_ARG(M.on, "no templates attached");
    k_tmpl_tvec<<<M.nwg_samples, 256, 0, pb.stream>>>(d_out, pb.tp.d_start, M.ndet, M.d_c0, M.d_ns, M.d_nh,
                                                      M.d_sweep, M.d_az, M.d_leg, M.d_hc, M.ncs, M.npoly, M.ground,
                                                      M.nsss, 2 * M.nhwp, M.fsdt, M.d_off, d_a, M.d_ces_of,
                                                      M.d_len);
    MGPU_LAUNCHED();
    return MGPU_OK;
}

// y = s q - W T B^+ T^T (s q), s = W (in_scale) or 1
static int apply_tm(Problem &pb, const double *d_q, const double *in_scale, double *d_y) {
    Templates &M = pb.tm;
    MGPU_ARG(M.on, "no templates attached");
    hipStream_t s = pb.stream;
    pb.prof.begin(MGPU_KT_TMPL, s);
    MGPU_RET(launch_tr(pb, d_q, in_scale, M.d_a1));
    if (M.ntot > 0) {  // (with no template class k_tmpl_sub reads no amplitude)
        k_tmpl_binv<<<grid_for(M.ntot, 4), 256, 0, s>>>(M.d_Binv, M.d_boff, M.d_off, M.d_row_block, M.ntot, M.d_a1,
                                                        M.d_a2, pb.live_state);
        MGPU_LAUNCHED();
    }
    if (d_q == d_y)
        k_tmpl_sub<true><<<M.nwg_samples, 256, 0, s>>>(nullptr, in_scale, d_y, pb.tp.d_start, pb.tp.d_T0, M.ndet,
                                                       M.d_c0, M.d_ns, M.d_nh, M.d_sweep, M.d_az, M.d_leg, M.d_hc,
                                                       M.ncs, M.npoly, M.ground, M.nsss, 2 * M.nhwp, M.fsdt, M.d_off,
                                                       M.d_a2, M.d_ces_of, M.d_len, pb.live_state);
    else
        k_tmpl_sub<false><<<M.nwg_samples, 256, 0, s>>>(d_q, in_scale, d_y, pb.tp.d_start, pb.tp.d_T0, M.ndet,
                                                        M.d_c0, M.d_ns, M.d_nh, M.d_sweep, M.d_az, M.d_leg, M.d_hc,
                                                        M.ncs, M.npoly, M.ground, M.nsss, 2 * M.nhwp, M.fsdt,
                                                        M.d_off, M.d_a2, M.d_ces_of, M.d_len, pb.live_state);
    MGPU_LAUNCHED();
    pb.prof.end(s);
    return MGPU_OK;
}

int templates_filter(Problem &pb, double *d_y) { return apply_tm(pb, d_y, nullptr, d_y); }

int templates_noise(Problem &pb, const double *d_q, double *d_y) {
    MGPU_ARG(pb.tp.lambda == 1, "templates need lambda = 1");
    return apply_tm(pb, d_q, pb.tp.d_T0, d_y);
}

}  // namespace mgpu

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

using mgpu::Problem;

int mappraiser_gpu_templates_set(mgpu_problem *h, const mgpu_templates *t) {
    MGPU_ARG(h && t, "NULL argument");
    Problem &pb = *reinterpret_cast<Problem *>(h);
    MGPU_HIP(hipSetDevice(pb.device));
    const int rc = mgpu::templates_build(pb, t);
    if (rc != MGPU_OK) pb.tm.on = false;
    return rc;
}

int mappraiser_gpu_templates_info(mgpu_problem *h, long long *ntot, int *nt_block, int *ranks) {
    MGPU_ARG(h, "NULL argument");
    Problem &pb = *reinterpret_cast<Problem *>(h);
    MGPU_ARG(pb.tm.on, "no templates attached");
    if (ntot) *ntot = pb.tm.ntot;
    if (nt_block) std::copy(pb.tm.h_nt.begin(), pb.tm.h_nt.end(), nt_block);
    if (ranks) std::copy(pb.tm.h_rank.begin(), pb.tm.h_rank.end(), ranks);
    return MGPU_OK;
}

int mappraiser_gpu_templates_apply(mgpu_problem *h, int op, const double *in, double *out) {
    MGPU_ARG(h && in && out, "NULL argument");
    Problem &pb = *reinterpret_cast<Problem *>(h);
    MGPU_HIP(hipSetDevice(pb.device));
    MGPU_ARG(pb.tm.on, "no templates attached");
    MGPU_ARG(op == MGPU_TMPL_T || op == MGPU_TMPL_TT || op == MGPU_TMPL_FILTER, "unknown template operation %d",
             op);
    MGPU_RET(mgpu::ensure_work(pb));
    hipStream_t s = pb.stream;
    const int64_t m = pb.pt.m, nt = pb.tm.ntot;
    const size_t in_n = (size_t) (op == MGPU_TMPL_T ? nt : m), out_n = (size_t) (op == MGPU_TMPL_TT ? nt : m);
    if (op == MGPU_TMPL_T) {
        MGPU_HIP(hipMemcpyAsync(pb.tm.d_a2, in, sizeof(double) * in_n, hipMemcpyHostToDevice, s));
        MGPU_HIP(hipMemsetAsync(pb.d_tod, 0, sizeof(double) * m, s));
        MGPU_RET(mgpu::templates_tvec(pb, pb.tm.d_a2, pb.d_tod));
        MGPU_HIP(hipMemcpyAsync(out, pb.d_tod, sizeof(double) * out_n, hipMemcpyDeviceToHost, s));
    } else if (op == MGPU_TMPL_TT) {
        MGPU_HIP(hipMemcpyAsync(pb.d_tod, in, sizeof(double) * in_n, hipMemcpyHostToDevice, s));
        MGPU_RET(mgpu::templates_trtvec(pb, pb.d_tod, pb.tm.d_a1));
        MGPU_HIP(hipMemcpyAsync(out, pb.tm.d_a1, sizeof(double) * out_n, hipMemcpyDeviceToHost, s));
    } else {
        MGPU_HIP(hipMemcpyAsync(pb.d_tod, in, sizeof(double) * in_n, hipMemcpyHostToDevice, s));
        MGPU_RET(mgpu::templates_filter(pb, pb.d_tod));
        MGPU_HIP(hipMemcpyAsync(out, pb.d_tod, sizeof(double) * out_n, hipMemcpyDeviceToHost, s));
    }
    MGPU_HIP(hipStreamSynchronize(s));
    return MGPU_OK;
}

int mappraiser_gpu_templates_kernel(mgpu_problem *h, double *B, double *Binv) {
    MGPU_ARG(h, "NULL argument");
    Problem &pb = *reinterpret_cast<Problem *>(h);
    MGPU_HIP(hipSetDevice(pb.device));
    MGPU_ARG(pb.tm.on, "no templates attached");
    const size_t n = (size_t) pb.tm.nB;
    if (B && n) MGPU_HIP(hipMemcpyAsync(B, pb.tm.d_B, sizeof(double) * n, hipMemcpyDeviceToHost, pb.stream));
    if (Binv && n) MGPU_HIP(hipMemcpyAsync(Binv, pb.tm.d_Binv, sizeof(double) * n, hipMemcpyDeviceToHost, pb.stream));
    MGPU_HIP(hipStreamSynchronize(pb.stream));
    return MGPU_OK;
}

}  // extern "C"
