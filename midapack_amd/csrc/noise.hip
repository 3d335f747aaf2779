// noise.hip -- the noise-model builder on the device: per Toeplitz block, the
// first row of the banded inverse noise covariance from a noise timestream.
//
// Reference: compute_invtt / noise2invtt (TOAST_interface/utils.py:435-601).
// For a block of n samples:
//   1. periodogram: Welch with one segment of n samples (nperseg = n, utils.py
//      527), Hann window (periodic), linear detrend, one-sided density scaling;
//   2. fit of the log-PSD model a + log10(1 + ((f + fmin)/fknee)^alpha) to
//      log10 psd[1:] (utils.py:530-538: p0 = (-7, -1, 0.1, 1e-6), bounds
//      a in [-20, 0], alpha in [-10, 0], fknee in [0, 20], fmin in [0, 1]);
//   3. psdm1 = 10^-a / (1 + ((f + fmin)/fknee)^alpha) on rfftfreq(n);
//   4. inv_tt = irfft(psdm1, n); result = window[:lambda] * inv_tt[:lambda],
//      window = ifftshift(chebwin(2 lambda, 150 dB)) (or the general Gaussian
//      of utils.py:569-573).
// The reference fits with scipy's trust-region-reflective least squares; here
// a bounded Levenberg-Marquardt runs one workgroup per block on the device.
// Both minimise the same objective; results agree to the fit tolerance.
//
// FFTs: batched radix-2 Stockham passes in HBM (one launch per pass) for
// power-of-two n; other n use Bluestein's chirp-z transform on the next
// power of two >= 2n - 1.  The forward transform of the periodogram and the
// inverse transform of psdm1 (real and even, so its inverse DFT is its
// forward DFT divided by n) share the same machinery.  This is set-up work
// (once per noise model), not the PCG hot path.
#include "mgpu_internal.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>

namespace mgpu {

namespace {

// one radix-2 Stockham pass over `batch` transforms of length M = 2^logm:
// sub-transform size ns -> 2 ns; sign -1 forward, +1 inverse (no scaling)
__global__ void k_fft_pass(const double2 *__restrict__ in, double2 *__restrict__ out, int logm, int ns, int sign,
                           int batch) {
    const int64_t M = (int64_t) 1 << logm, half = M >> 1;
    const int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= half * batch) return;
    const int64_t b = gid / half, j = gid % half;
    const double2 *x = in + b * M;
    double2 *y = out + b * M;
    const int64_t r = j % ns;
    double s, c;
    sincospi((double) sign * (double) r / (double) ns, &s, &c);
    const double2 v0 = x[j], u = x[j + half];
    const double2 v1 = make_double2(u.x * c - u.y * s, u.x * s + u.y * c);
    const int64_t d = (j / ns) * ns * 2 + r;
    y[d] = make_double2(v0.x + v1.x, v0.y + v1.y);
    y[d + ns] = make_double2(v0.x - v1.x, v0.y - v1.y);
}

// exp(sign * pi * i * j^2 / n), with j^2 reduced mod 2n exactly
__device__ __forceinline__ double2 chirp(int64_t j, int64_t n, int sign) {
    const int64_t e = (j * j) % (2 * n);
    double s, c;
    sincospi((double) sign * (double) e / (double) n, &s, &c);
    return make_double2(c, s);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Bluestein, step 1: A_t = x_t w_t (t < n), 0 beyond; B = conj chirp, wrapped
__global__ void k_blue_prep(const double *__restrict__ x, const int64_t *__restrict__ xoff,
                            const int *__restrict__ nlen, int logm, int batch, double2 *__restrict__ A,
                            double2 *__restrict__ B) {
    const int64_t M = (int64_t) 1 << logm;
    const int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= M * batch) return;
    const int64_t b = gid / M, t = gid % M;
    const int64_t n = nlen[b];
    A[gid] = t < n ? cmul(make_double2(x[xoff[b] + t], 0.0), chirp(t, n, -1)) : make_double2(0.0, 0.0);
    double2 v = make_double2(0.0, 0.0);
    if (t < n)
        v = chirp(t, n, +1);
    else if (t > M - n)
        v = chirp(M - t, n, +1);
    B[gid] = v;
}

__global__ void k_pointwise(double2 *__restrict__ A, const double2 *__restrict__ B, int64_t total) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) A[i] = cmul(A[i], B[i]);
}

// Bluestein, last step: X_k = w_k conv_k / M for k < nout
__global__ void k_blue_post(const double2 *__restrict__ C, const int *__restrict__ nlen, int logm, int batch,
                            int64_t nout_max, double2 *__restrict__ X) {
    const int64_t M = (int64_t) 1 << logm;
    const int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nout_max * batch) return;
    const int64_t b = gid / nout_max, k = gid % nout_max;
    const int64_t n = nlen[b];
    if (k >= n) return;
    const double2 v = cmul(C[b * M + k], chirp(k, n, -1));
    X[b * nout_max + k] = make_double2(v.x / (double) M, v.y / (double) M);
}

// power of two: load the real input as complex
__global__ void k_real_to_c(const double *__restrict__ x, const int64_t *__restrict__ xoff, int logm, int batch,
                            double2 *__restrict__ A) {
    const int64_t M = (int64_t) 1 << logm;
    const int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= M * batch) return;
    const int64_t b = gid / M, t = gid % M;
    A[gid] = make_double2(x[xoff[b] + t], 0.0);
}

struct Buf {
    Arena &ar;
    double2 *p = nullptr;
    explicit Buf(Arena &a) : ar(a) {}
    ~Buf() {
        if (p) ar.release(p);
    }
};

// forward DFT of `batch` real sequences of lengths nlen[b] (all with the same
// transform size M = 2^logm: n = M for powers of two, else the Bluestein
// size) read at x + xoff[b]; X[b * nout_max + k] = DFT_k for k < n.
int dft_real(Problem &pb, const double *d_x, const int64_t *d_xoff, const int *d_nlen, int batch, int logm,
             bool pow2, int64_t nout_max, double2 *d_X) {
    hipStream_t s = pb.stream;
    const int64_t M = (int64_t) 1 << logm, total = M * batch;
    Buf a(pb.arena), b(pb.arena), c(pb.arena);
    MGPU_RET(pb.arena.alloc(&a.p, (size_t) total));
    MGPU_RET(pb.arena.alloc(&b.p, (size_t) total));
    auto fft = [&](double2 *&src, double2 *&tmp, int sign) -> int {
        for (int ns = 1; ns < M; ns <<= 1) {
            k_fft_pass<<<grid_for(total / 2, 256), 256, 0, s>>>(src, tmp, logm, ns, sign, batch);
            MGPU_LAUNCHED();
            std::swap(src, tmp);
        }
        return MGPU_OK;
    };
    if (pow2) {
        k_real_to_c<<<grid_for(total, 256), 256, 0, s>>>(d_x, d_xoff, logm, batch, a.p);
        MGPU_LAUNCHED();
        MGPU_RET(fft(a.p, b.p, -1));
        // a.p holds the spectrum (pointers swapped in place)
        MGPU_HIP(hipMemcpy2DAsync(d_X, sizeof(double2) * nout_max, a.p, sizeof(double2) * M,
                                  sizeof(double2) * std::min<int64_t>(nout_max, M), batch, hipMemcpyDeviceToDevice, s));
        MGPU_HIP(hipStreamSynchronize(s));  // before the work buffers go back to the block cache
        return MGPU_OK;
    }
    MGPU_RET(pb.arena.alloc(&c.p, (size_t) total));
    k_blue_prep<<<grid_for(total, 256), 256, 0, s>>>(d_x, d_xoff, d_nlen, logm, batch, a.p, c.p);
    MGPU_LAUNCHED();
    MGPU_RET(fft(a.p, b.p, -1));
    MGPU_RET(fft(c.p, b.p, -1));
    k_pointwise<<<grid_for(total, 256), 256, 0, s>>>(a.p, c.p, total);
    MGPU_LAUNCHED();
    MGPU_RET(fft(a.p, b.p, +1));
    k_blue_post<<<grid_for(nout_max * batch, 256), 256, 0, s>>>(a.p, d_nlen, logm, batch, nout_max, d_X);
    MGPU_LAUNCHED();
    MGPU_HIP(hipStreamSynchronize(s));  // before the work buffers go back to the block cache
    return MGPU_OK;
}

// detrended, Hann-windowed segment: xw[t] = w_t (x_t - (c0 + c1 t)), the line
// fitted by least squares (scipy.signal.detrend type 'linear'); one workgroup per block
__global__ __launch_bounds__(256) void k_detrend_window(const double *__restrict__ noise,
                                                        const int64_t *__restrict__ start,
                                                        const int *__restrict__ nlen, const int *__restrict__ blocks,
                                                        const int64_t *__restrict__ xoff, double *__restrict__ xw,
                                                        double *__restrict__ wsum2) {
    __shared__ double red[256];
    const int bi = blockIdx.x;
    const int64_t n = nlen[bi];
    const double *x = noise + start[blocks[bi]];
    // sums for the line fit over t = 0..n-1
    double sx = 0.0, stx = 0.0;
    for (int64_t t = threadIdx.x; t < n; t += 256) {
        sx += x[t];
        stx += (double) t * x[t];
    }
    auto tree = [&](double v) {
        red[threadIdx.x] = v;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int) threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        const double r = red[0];
        __syncthreads();
        return r;
    };
    sx = tree(sx);
    stx = tree(stx);
    const double dn = (double) n;
    const double st = dn * (dn - 1.0) / 2.0, stt = (dn - 1.0) * dn * (2.0 * dn - 1.0) / 6.0;
    const double det = dn * stt - st * st;
    const double c1 = (dn * stx - st * sx) / det, c0 = (sx - c1 * st) / dn;
    double w2 = 0.0;
    for (int64_t t = threadIdx.x; t < n; t += 256) {
        double sn, cs;
        sincospi(2.0 * (double) t / dn, &sn, &cs);
        const double w = 0.5 - 0.5 * cs;  // periodic Hann
        xw[xoff[bi] + t] = w * (x[t] - (c0 + c1 * (double) t));
        w2 += w * w;
    }
    w2 = tree(w2);
    if (threadIdx.x == 0) wsum2[bi] = w2;
}

// log10 of the one-sided density periodogram, k = 1 .. n/2
__global__ void k_logpsd(const double2 *__restrict__ X, const int *__restrict__ nlen, int64_t nout_max, int batch,
                         const double *__restrict__ wsum2, double fsamp, double *__restrict__ y) {
    const int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t half = nout_max / 2 + 1;
    if (gid >= half * batch) return;
    const int64_t b = gid / half, k = gid % half;
    const int64_t n = nlen[b];
    if (k > n / 2) return;
    const double2 v = X[b * nout_max + k];
    double p = (v.x * v.x + v.y * v.y) / (fsamp * wsum2[b]);
    const bool nyq = (n % 2 == 0) && k == n / 2;
    if (k > 0 && !nyq) p *= 2.0;
    y[b * half + k] = log10(p);
}

// model pieces at frequency f: h = log10(1 + g), g = u^alpha, u = (f + fmin) / fknee
struct Model {
    double h, dh_alpha, dh_fknee, dh_fmin;
};
__device__ __forceinline__ Model model_at(double f, double alpha, double fknee, double fmin) {
    const double u = (f + fmin) / fknee;
    const double lu = log(u);
    const double g = exp(alpha * lu);
    const double k = g / ((1.0 + g) * 2.302585092994045684);  // g / ((1 + g) ln 10)
    Model m;
    m.h = log10(1.0 + g);
    m.dh_alpha = k * lu;
    m.dh_fknee = -k * alpha / fknee;
    m.dh_fmin = k * alpha / (f + fmin);
    return m;
}

constexpr int kFitIter = 400;

// bounded Levenberg-Marquardt fit of a + log10(1 + ((f+fmin)/fknee)^alpha) to y[1..n/2];
// one workgroup per block; params[4b..4b+3] = (a, alpha, fknee, fmin)
__global__ __launch_bounds__(256) void k_fit(const double *__restrict__ y, const int *__restrict__ nlen,
                                             int64_t nout_max, double fsamp, double *__restrict__ params) {
    __shared__ double red[256];
    __shared__ double sh[20];
    const int b = blockIdx.x;
    const int64_t n = nlen[b], half = nout_max / 2 + 1, K = n / 2;
    const double *yb = y + b * half;
    const double lo[4] = {-20.0, -10.0, 1e-12, 0.0}, hi[4] = {0.0, 0.0, 20.0, 1.0};
    double p[4] = {-7.0, -1.0, 0.1, 1e-6};
    auto tree = [&](double v) {
        red[threadIdx.x] = v;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int) threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        const double r = red[0];
        __syncthreads();
        return r;
    };
    auto cost_at = [&](const double *q) {
        double c = 0.0;
        for (int64_t k = 1 + threadIdx.x; k <= K; k += 256) {
            const double f = (double) k * fsamp / (double) n;
            const double r = q[0] + model_at(f, q[1], q[2], q[3]).h - yb[k];
            c += r * r;
        }
        return tree(c);
    };
    double cost = cost_at(p);
    double mu = 1e-3;
    for (int it = 0; it < kFitIter; it++) {
        // normal equations J^T J (10 entries) and J^T r (4)
        double acc[14] = {0};
        for (int64_t k = 1 + threadIdx.x; k <= K; k += 256) {
            const double f = (double) k * fsamp / (double) n;
            const Model md = model_at(f, p[1], p[2], p[3]);
            const double r = p[0] + md.h - yb[k];
            const double J[4] = {1.0, md.dh_alpha, md.dh_fknee, md.dh_fmin};
            int e = 0;
            for (int i = 0; i < 4; i++)
                for (int j = i; j < 4; j++) acc[e++] += J[i] * J[j];
            for (int i = 0; i < 4; i++) acc[10 + i] += J[i] * r;
        }
        for (int e = 0; e < 14; e++) {
            const double v = tree(acc[e]);
            if (threadIdx.x == 0) sh[e] = v;
        }
        __syncthreads();
        // thread 0: damped step (J^T J + mu diag) d = -J^T r, projected on the bounds
        if (threadIdx.x == 0) {
            double A[4][4], g[4];
            int e = 0;
            for (int i = 0; i < 4; i++)
                for (int j = i; j < 4; j++) {
                    A[i][j] = sh[e];
                    A[j][i] = sh[e];
                    e++;
                }
            for (int i = 0; i < 4; i++) g[i] = -sh[10 + i];
            for (int i = 0; i < 4; i++) A[i][i] += mu * (A[i][i] > 0.0 ? A[i][i] : 1.0);
            // active set: a parameter on a bound whose descent direction points
            // outside stays fixed (its row and column leave the system)
            for (int i = 0; i < 4; i++) {
                const bool act = (p[i] <= lo[i] && g[i] < 0.0) || (p[i] >= hi[i] && g[i] > 0.0);
                if (act) {
                    for (int j = 0; j < 4; j++) {
                        A[i][j] = 0.0;
                        A[j][i] = 0.0;
                    }
                    A[i][i] = 1.0;
                    g[i] = 0.0;
                }
            }
            // Gaussian elimination with partial pivoting (4x4)
            for (int c = 0; c < 4; c++) {
                int pr = c;
                for (int r = c + 1; r < 4; r++)
                    if (fabs(A[r][c]) > fabs(A[pr][c])) pr = r;
                for (int j = 0; j < 4; j++) {
                    const double t = A[c][j];
                    A[c][j] = A[pr][j];
                    A[pr][j] = t;
                }
                const double tg = g[c];
                g[c] = g[pr];
                g[pr] = tg;
                for (int r = c + 1; r < 4; r++) {
                    const double fct = A[c][c] != 0.0 ? A[r][c] / A[c][c] : 0.0;
                    for (int j = c; j < 4; j++) A[r][j] -= fct * A[c][j];
                    g[r] -= fct * g[c];
                }
            }
            double d[4];
            for (int r = 3; r >= 0; r--) {
                double v = g[r];
                for (int j = r + 1; j < 4; j++) v -= A[r][j] * d[j];
                d[r] = A[r][r] != 0.0 ? v / A[r][r] : 0.0;
            }
            double dn = 0.0, pn = 0.0;
            for (int i = 0; i < 4; i++) {
                const double q = fmin(fmax(p[i] + d[i], lo[i]), hi[i]);
                sh[15 + i] = q;
                dn += (q - p[i]) * (q - p[i]);
                pn += p[i] * p[i];
            }
            sh[19] = sqrt(dn) <= 1e-14 * (sqrt(pn) + 1e-14) ? 1.0 : 0.0;
        }
        __syncthreads();
        double q[4] = {sh[15], sh[16], sh[17], sh[18]};
        const bool tiny = sh[19] != 0.0;
        const double cq = cost_at(q);
        if (cq < cost) {
            const double rel = (cost - cq) / fmax(cost, 1e-300);
            for (int i = 0; i < 4; i++) p[i] = q[i];
            cost = cq;
            mu = fmax(mu * 0.3, 1e-12);
            if (rel < 1e-15 || tiny) break;
        } else {
            mu *= 10.0;
            if (mu > 1e12 || tiny) break;
        }
    }
    if (threadIdx.x == 0)
        for (int i = 0; i < 4; i++) params[4 * (int64_t) b + i] = p[i];
}

// psdm1 of the fitted model on rfftfreq(n), extended to the full even sequence of length n
__global__ void k_psdm1_even(const double *__restrict__ params, const int *__restrict__ nlen, int64_t nmax,
                             int batch, double fsamp, double *__restrict__ out) {
    const int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nmax * batch) return;
    const int64_t b = gid / nmax, t = gid % nmax;
    const int64_t n = nlen[b];
    if (t >= n) return;
    const int64_t k = t <= n / 2 ? t : n - t;
    const double *p = params + 4 * b;
    const double f = (double) k * fsamp / (double) n;
    const double u = (f + p[3]) / p[2];
    out[b * nmax + t] = pow(10.0, -p[0]) / (1.0 + pow(u, p[1]));
}

// invtt row: window[j] * Re(DFT(psdm1_even))[j] / n for j < lambda
__global__ void k_invtt_rows(const double2 *__restrict__ X, const int *__restrict__ nlen, int64_t nout_max,
                             const int *__restrict__ blocks, int batch, int lambda,
                             const double *__restrict__ window, double *__restrict__ invtt) {
    const int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t) lambda * batch) return;
    const int64_t b = gid / lambda, j = gid % lambda;
    invtt[(int64_t) blocks[b] * lambda + j] = window[j] * (X[b * nout_max + j].x / (double) nlen[b]);
}

}  // namespace

// Symmetric Dolph-Chebyshev window of M points, attenuation at dB, by its
// frequency-domain definition: W_k = T_{M-1}(beta cos(pi k / M)) with
// beta = cosh(acosh(10^(at/20)) / (M-1)); the window is the real DFT of W
// (with a half-sample phase for even M), mirrored and normalised to a unit
// maximum.  Direct DFT over a cosine table: set-up work, cached per lambda.
static std::vector<double> chebwin_sym(int M, double at) {
    const double order = M - 1.0;
    const double beta = std::cosh(std::acosh(std::pow(10.0, std::fabs(at) / 20.0)) / order);
    std::vector<double> pr((size_t) M), pi((size_t) M), ct((size_t) M), st((size_t) M);
    for (int k = 0; k < M; k++) {
        ct[k] = std::cos(2.0 * M_PI * k / M);
        st[k] = std::sin(2.0 * M_PI * k / M);
    }
    for (int k = 0; k < M; k++) {
        const double x = beta * std::cos(M_PI * k / M);
        double v;
        if (x > 1.0)
            v = std::cosh(order * std::acosh(x));
        else if (x < -1.0)
            v = (2.0 * (M % 2) - 1.0) * std::cosh(order * std::acosh(-x));
        else
            v = std::cos(order * std::acos(x));
        if (M % 2) {
            pr[k] = v;
            pi[k] = 0.0;
        } else {
            pr[k] = v * std::cos(M_PI * k / M);
            pi[k] = v * std::sin(M_PI * k / M);
        }
    }
    const int n = M % 2 ? (M + 1) / 2 : M / 2 + 1;
    std::vector<double> w((size_t) n);
    for (int t = 0; t < n; t++) {  // real part of the forward DFT at t
        double acc = 0.0;
        for (int k = 0; k < M; k++) {
            const size_t e = (size_t) (((int64_t) k * t) % M);
            acc += pr[k] * ct[e] + pi[k] * st[e];
        }
        w[t] = acc;
    }
    std::vector<double> full;
    full.reserve((size_t) M);
    for (int t = n - 1; t >= 1; t--) full.push_back(w[t]);
    for (int t = (M % 2 ? 0 : 1); t < n; t++) full.push_back(w[t]);
    double mx = 0.0;
    for (double v : full) mx = std::max(mx, v);
    for (double &v : full) v /= mx;
    return full;
}

// window[:lambda] of ifftshift(get_window(kind, 2 lambda)) (utils.py:562-577):
// the DFT-even ("fftbins") window of 2 lambda points is the symmetric window of
// 2 lambda + 1 points without its last point
static const std::vector<double> &apod_window(int lambda, int kind) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::vector<double>> cache;
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find({lambda, kind});
    if (it != cache.end()) return it->second;
    const int N = 2 * lambda;
    std::vector<double> full;
    if (kind == MGPU_APOD_CHEBWIN) {
        full = chebwin_sym(N + 1, 150.0);
    } else {  // general Gaussian, p = 1, sigma = lambda / 3 (q_apo = 3)
        const double sigma = lambda / 3.0, c = lambda;  // centre of N + 1 points
        full.resize((size_t) N + 1);
        for (int t = 0; t <= N; t++) full[t] = std::exp(-0.5 * ((t - c) / sigma) * ((t - c) / sigma));
    }
    std::vector<double> w((size_t) lambda);
    for (int j = 0; j < lambda; j++) w[j] = full[(size_t) ((j + lambda) % N)];
    return cache[{lambda, kind}] = w;
}

int noise_invtt(Problem &pb, const double *d_noise, int nb, const int *block_sizes, int lambda, double fsamp,
                int window_kind, double *d_invtt, double *h_params) {
    hipStream_t s = pb.stream;
    MGPU_ARG(nb > 0 && lambda >= 1 && fsamp > 0.0, "nb=%d lambda=%d fsamp=%g", nb, lambda, fsamp);
    std::vector<int64_t> start((size_t) nb + 1, 0);
    for (int b = 0; b < nb; b++) {
        MGPU_ARG(block_sizes[b] >= 4, "block %d has %d samples", b, block_sizes[b]);
        MGPU_ARG(lambda <= block_sizes[b] / 2, "Bandwidth cannot be larger than timestream (utils.py:561): "
                                               "lambda=%d, block %d has %d samples", lambda, b, block_sizes[b]);
        start[b + 1] = start[b] + block_sizes[b];
    }
    MGPU_ARG(window_kind == MGPU_APOD_CHEBWIN || window_kind == MGPU_APOD_GAUSSIAN, "unknown window %d",
             window_kind);
    const std::vector<double> &win = apod_window(lambda, window_kind);
    double *d_win, *d_params;
    int64_t *d_start;
    MGPU_RET(pb.arena.alloc(&d_win, (size_t) lambda));
    MGPU_RET(pb.arena.alloc(&d_params, (size_t) 4 * nb));
    MGPU_RET(pb.arena.alloc(&d_start, (size_t) nb + 1));
    MGPU_HIP(hipMemcpyAsync(d_win, win.data(), sizeof(double) * lambda, hipMemcpyHostToDevice, s));
    MGPU_HIP(hipMemcpyAsync(d_start, start.data(), sizeof(int64_t) * (nb + 1), hipMemcpyHostToDevice, s));
    // group blocks by transform size (power of two n, or the Bluestein size)
    std::map<std::pair<int, int>, std::vector<int>> groups;  // (logm, pow2) -> blocks
    for (int b = 0; b < nb; b++) {
        const int64_t n = block_sizes[b];
        const bool p2 = (n & (n - 1)) == 0;
        int logm = 0;
        while (((int64_t) 1 << logm) < (p2 ? n : 2 * n - 1)) logm++;
        groups[{logm, p2 ? 1 : 0}].push_back(b);
    }
    for (auto &kv : groups) {
        const int logm = kv.first.first;
        const bool p2 = kv.first.second != 0;
        const std::vector<int> &bl = kv.second;
        // batches bounded to ~2 GiB of complex work per buffer
        const int64_t M = (int64_t) 1 << logm;
        const int per = (int) std::max<int64_t>(1, ((int64_t) 1 << 27) / M);
        for (size_t g0 = 0; g0 < bl.size(); g0 += (size_t) per) {
            const int batch = (int) std::min<size_t>((size_t) per, bl.size() - g0);
            std::vector<int> blk(bl.begin() + g0, bl.begin() + g0 + batch), nlen((size_t) batch);
            std::vector<int64_t> xoff((size_t) batch);
            int64_t nmax = 0, tot = 0;
            for (int i = 0; i < batch; i++) {
                nlen[i] = block_sizes[blk[i]];
                nmax = std::max<int64_t>(nmax, nlen[i]);
            }
            for (int i = 0; i < batch; i++) {
                xoff[i] = tot;
                tot += nmax;
            }
            int *d_blk, *d_nlen;
            int64_t *d_xoff;
            double *d_xw, *d_w2, *d_y, *d_even;
            double2 *d_X;
            MGPU_RET(pb.arena.alloc(&d_blk, (size_t) batch));
            MGPU_RET(pb.arena.alloc(&d_nlen, (size_t) batch));
            MGPU_RET(pb.arena.alloc(&d_xoff, (size_t) batch));
            MGPU_RET(pb.arena.alloc(&d_xw, (size_t) tot));
            MGPU_RET(pb.arena.alloc(&d_w2, (size_t) batch));
            MGPU_RET(pb.arena.alloc(&d_y, (size_t) (nmax / 2 + 1) * batch));
            MGPU_RET(pb.arena.alloc(&d_even, (size_t) tot));
            MGPU_RET(pb.arena.alloc(&d_X, (size_t) tot));
            MGPU_HIP(hipMemcpyAsync(d_blk, blk.data(), sizeof(int) * batch, hipMemcpyHostToDevice, s));
            MGPU_HIP(hipMemcpyAsync(d_nlen, nlen.data(), sizeof(int) * batch, hipMemcpyHostToDevice, s));
            MGPU_HIP(hipMemcpyAsync(d_xoff, xoff.data(), sizeof(int64_t) * batch, hipMemcpyHostToDevice, s));
            MGPU_HIP(hipMemsetAsync(d_xw, 0, sizeof(double) * tot, s));
            k_detrend_window<<<batch, 256, 0, s>>>(d_noise, d_start, d_nlen, d_blk, d_xoff, d_xw, d_w2);
            MGPU_LAUNCHED();
            MGPU_RET(dft_real(pb, d_xw, d_xoff, d_nlen, batch, logm, p2, nmax, d_X));
            k_logpsd<<<grid_for((nmax / 2 + 1) * batch, 256), 256, 0, s>>>(d_X, d_nlen, nmax, batch, d_w2, fsamp, d_y);
            MGPU_LAUNCHED();
            k_fit<<<batch, 256, 0, s>>>(d_y, d_nlen, nmax, fsamp, d_params + 4 * (int64_t) 0 + 0);
            MGPU_LAUNCHED();
            // params of this batch were written at 4 * local index: move them to their blocks below
            std::vector<double> hp((size_t) 4 * batch);
            MGPU_HIP(hipMemcpyAsync(hp.data(), d_params, sizeof(double) * 4 * batch, hipMemcpyDeviceToHost, s));
            MGPU_HIP(hipStreamSynchronize(s));
            if (h_params)
                for (int i = 0; i < batch; i++)
                    for (int k = 0; k < 4; k++) h_params[4 * (int64_t) blk[i] + k] = hp[4 * (size_t) i + k];
            k_psdm1_even<<<grid_for(tot, 256), 256, 0, s>>>(d_params, d_nlen, nmax, batch, fsamp, d_even);
            MGPU_LAUNCHED();
            MGPU_RET(dft_real(pb, d_even, d_xoff, d_nlen, batch, logm, p2, nmax, d_X));
            k_invtt_rows<<<grid_for((int64_t) lambda * batch, 256), 256, 0, s>>>(d_X, d_nlen, nmax, d_blk, batch,
                                                                                 lambda, d_win, d_invtt);
            MGPU_LAUNCHED();
            MGPU_HIP(hipStreamSynchronize(s));
            for (void *p : {(void *) d_blk, (void *) d_nlen, (void *) d_xoff, (void *) d_xw, (void *) d_w2,
                            (void *) d_y, (void *) d_even, (void *) d_X})
                pb.arena.release(p);
        }
    }
    for (void *p : {(void *) d_win, (void *) d_params, (void *) d_start}) pb.arena.release(p);
    return MGPU_OK;
}

}  // namespace mgpu
