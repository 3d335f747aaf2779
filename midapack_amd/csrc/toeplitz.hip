// toeplitz.hip -- N^-1 = block-diagonal banded Toeplitz product with per-block
// mean removal (gfx950).
//
// Reference semantics (one rank):
//   apply_weights  pcg_true.c:578-600: lambda == 1 -> tod *= T_b[0], else stbmmProd.
//   stbmmProd      toeplitz_wizard.c:80-111: V_b -= mean(V_b) for every block,
//                  then mpi_stbmm (toeplitz_block.c:115-636), which on one rank
//                  multiplies every block [idv, idv+n) by its banded symmetric
//                  Toeplitz matrix T_b (band T_b[0..lambda-1]) with zero
//                  boundaries: (T V)[i] = sum_{|i-j|<lambda, j in block} T[|i-j|] V[j].
//                  The FFT shift-and-overlap of stmm_core (toeplitz.c:731-865)
//                  computes exactly this product; blocks with n < lambda-1 are
//                  left as they are (stmm_main, toeplitz.c:912-913).
//
// MI355X design.  Each tile is one workgroup.  A real tile of N = 2*NC samples
// (NC complex points) is packed z[n] = x[2n] + i x[2n+1] and held in registers,
// V complex values per lane (NT = NC/V lanes).  A mixed-radix Stockham FFT
// (radix 4..32 butterflies in registers) exchanges data through LDS between
// passes; when the complex tile does not fit the 160 KB LDS twice (FP64 at
// NC = 16384) the real and imaginary planes are exchanged one after the other.
// The last forward pass and the first inverse pass give each lane the pair of
// butterflies {j, NC/R - j}, so the real-FFT untangling, the multiplication by
// the real spectrum H of the circulant embedding and the re-tangling collapse
// to one in-register step
//     Z'[k] = a_k Z[k] + i b_k conj(Z[NC-k]),
//     a_k = (Hs - Hd sin(2 pi k/N))/NC,  b_k = Hd cos(2 pi k/N)/NC,
//     Hs = (H[k] + H[k+NC])/2,  Hd = (H[k] - H[k+NC])/2,
// and the inverse transform is conj(FFT(conj(Z'))), reusing the forward passes.
// Tiles read y - mean (zero outside the block) straight from HBM in the first
// pass and write only their central L = N - 2(lambda-1) outputs.  Only the
// filter spectrum differs from the reference's complex multiply: H is real and
// even (T_circ is real and symmetric), so the imaginary rounding noise that
// FFTW leaves in C_fft (and that toeplitz.c:587-588 multiplies in) is dropped.
#include "mgpu_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <atomic>
#include <unordered_map>

namespace mgpu {

// ---------------------------------------------------------------------------
// compile-time schedule
// ---------------------------------------------------------------------------
// Forward schedule: the last pass has radix V/2 (two paired butterflies per
// lane); the other LOGNC - (LOGV-1) bits are spread as evenly as possible over
// passes of at most LOGR bits.  The inverse schedule is the reverse.
constexpr int n_passes(int LOGNC, int LOGV, int LOGR) {
    return 1 + (LOGNC - (LOGV - 1) + LOGR - 1) / LOGR;
}

constexpr int fwd_bits(int LOGNC, int LOGV, int LOGR, int p) {
    return (p == n_passes(LOGNC, LOGV, LOGR) - 1)
                   ? LOGV - 1
                   : (LOGNC - (LOGV - 1)) / (n_passes(LOGNC, LOGV, LOGR) - 1) +
                             ((p < (LOGNC - (LOGV - 1)) % (n_passes(LOGNC, LOGV, LOGR) - 1)) ? 1 : 0);
}

// bits of pass p of the forward (INV=false) or inverse (reversed) schedule
constexpr int sched_bits(int LOGNC, int LOGV, int LOGR, bool INV, int p) {
    return INV ? fwd_bits(LOGNC, LOGV, LOGR, n_passes(LOGNC, LOGV, LOGR) - 1 - p)
               : fwd_bits(LOGNC, LOGV, LOGR, p);
}

constexpr int sched_logns(int LOGNC, int LOGV, int LOGR, bool INV, int p) {
    return p == 0 ? 0
                  : sched_logns(LOGNC, LOGV, LOGR, INV, p - 1) + sched_bits(LOGNC, LOGV, LOGR, INV, p - 1);
}

#ifndef TILE14_F64_LOGV
#define TILE14_F64_LOGV 5
#endif
#ifndef TILE14_F64_LOGR
#define TILE14_F64_LOGR 4
#endif
#ifndef TILE14_F32_LOGV
#define TILE14_F32_LOGV 5
#endif
#ifndef TILE14_F32_LOGR
#define TILE14_F32_LOGR 4
#endif

// coefficient pairs of the spectral step loaded together (FP64 tiles of 16+
// values per lane, which run with a scheduling fence between groups)
#ifndef MGPU_SPEC_GROUP
#define MGPU_SPEC_GROUP 4
#endif

// one padding element every 32 (FP64 tiles of 32768 samples: 0.559 ms against
// 0.576 ms with one every 16, 0.630 ms with one every 8; FP32 unchanged)
#ifndef MGPU_LDS_PAD_SHIFT
#define MGPU_LDS_PAD_SHIFT 5
#endif
constexpr int kPadShift = MGPU_LDS_PAD_SHIFT;  // one padding element every 2^kPadShift in LDS (bank spread)
__device__ __forceinline__ int lpad(int i) { return i + (i >> kPadShift); }
constexpr int lds_elems(int NC) { return NC + (NC >> kPadShift); }

// ---------------------------------------------------------------------------
// in-register DFT of size 2^LR (forward), natural order in and out
// ---------------------------------------------------------------------------
// cos/sin(2 pi j / 32), j = 0..31
__device__ constexpr double kCos32[32] = {
        1.0,
        0.98078528040323044913,
        0.92387953251128675613,
        0.83146961230254523708,
        0.70710678118654752440,
        0.55557023301960222474,
        0.38268343236508977173,
        0.19509032201612826785,
        0.0,
        -0.19509032201612826785,
        -0.38268343236508977173,
        -0.55557023301960222474,
        -0.70710678118654752440,
        -0.83146961230254523708,
        -0.92387953251128675613,
        -0.98078528040323044913,
        -1.0,
        -0.98078528040323044913,
        -0.92387953251128675613,
        -0.83146961230254523708,
        -0.70710678118654752440,
        -0.55557023301960222474,
        -0.38268343236508977173,
        -0.19509032201612826785,
        0.0,
        0.19509032201612826785,
        0.38268343236508977173,
        0.55557023301960222474,
        0.70710678118654752440,
        0.83146961230254523708,
        0.92387953251128675613,
        0.98078528040323044913,
};

template <int N>
__device__ constexpr int bitrev(int k) {
    int r = 0;
    for (int b = 1; b < N; b <<= 1) {
        r = (r << 1) | (k & 1);
        k >>= 1;
    }
    return r;
}

// Radix-2 decimation-in-time butterfly of a stage of length LEN at (a, b):
// a <- a + w b, b <- a - w b with w = exp(-2 pi i J / LEN).  A general w costs
// six FMA-unit operations: a + w b as two fused chains, a - w b = 2a - (a + w b).
template <typename R, int LEN, int J>
__device__ __forceinline__ void bfly_dit(R &ar, R &ai, R &br, R &bi) {
    constexpr int e = (32 / LEN) * J;  // index into the 32-point table
    R r0, i0, r1, i1;
    if constexpr (J == 0) {
        r0 = ar + br;
        i0 = ai + bi;
        r1 = ar - br;
        i1 = ai - bi;
    } else if constexpr (4 * J == LEN) {  // w = -i: w b = bi - i br
        r0 = ar + bi;
        i0 = ai - br;
        r1 = ar - bi;
        i1 = ai + br;
    } else if constexpr (8 * J == LEN) {  // w = (1 - i)/sqrt2: w b = c (br + bi) + i c (bi - br)
        const R c = (R) kCos32[4];
        const R u = br + bi, v = bi - br;
        r0 = fma(c, u, ar);
        i0 = fma(c, v, ai);
        r1 = fma(-c, u, ar);
        i1 = fma(-c, v, ai);
    } else if constexpr (8 * J == 3 * LEN) {  // w = -(1 + i)/sqrt2: w b = -c (br - bi) - i c (br + bi)
        const R c = (R) kCos32[4];
        const R u = br - bi, v = br + bi;
        r0 = fma(-c, u, ar);
        i0 = fma(-c, v, ai);
        r1 = fma(c, u, ar);
        i1 = fma(c, v, ai);
    } else {
        // w = cos x - i sin x: w b = (c br + s bi) + i (c bi - s br)
        const R c = (R) kCos32[e], s = (R) kCos32[(e + 24) & 31];  // sin(x) = cos(x - pi/2)
        r0 = fma(c, br, fma(s, bi, ar));
        i0 = fma(c, bi, fma(-s, br, ai));
        r1 = fma((R) 2, ar, -r0);
        i1 = fma((R) 2, ai, -i0);
    }
    ar = r0;
    ai = i0;
    br = r1;
    bi = i1;
}

// every butterfly of the stage of length LEN (groups START, START + LEN, ...)
template <typename R, int RAD, int LEN, int START, int J>
struct DitStage {
    static __device__ __forceinline__ void run(R *re, R *im) {
        if constexpr (START < RAD) {
            if constexpr (J < LEN / 2) {
                constexpr int i0 = START + J, i1 = START + J + LEN / 2;
                bfly_dit<R, LEN, J>(re[i0], im[i0], re[i1], im[i1]);
                DitStage<R, RAD, LEN, START, J + 1>::run(re, im);
            } else {
                DitStage<R, RAD, LEN, START + LEN, 0>::run(re, im);
            }
        }
    }
};

// stages LEN, 2 LEN, ..., RAD
template <typename R, int RAD, int LEN>
struct DitAll {
    static __device__ __forceinline__ void run(R *re, R *im) {
        if constexpr (LEN <= RAD) {
            DitStage<R, RAD, LEN, 0, 0>::run(re, im);
            DitAll<R, RAD, 2 * LEN>::run(re, im);
        }
    }
};

// forward DFT of RAD points at re/im[0..RAD), natural order in and out: the
// inputs are renamed into bit-reversed order (no data movement), then the
// radix-2 DIT stages run in place (RAD = 32: 388 operations, 16: 148)
template <typename R, int RAD>
__device__ __forceinline__ void dft(R *re, R *im) {
    if constexpr (RAD > 1) {
        R tr[RAD], ti[RAD];
#pragma unroll
        for (int k = 0; k < RAD; k++) {
            tr[k] = re[bitrev<RAD>(k)];
            ti[k] = im[bitrev<RAD>(k)];
        }
        DitAll<R, RAD, 2>::run(tr, ti);
#pragma unroll
        for (int k = 0; k < RAD; k++) {
            re[k] = tr[k];
            im[k] = ti[k];
        }
    }
}

// ---------------------------------------------------------------------------
// Stockham pass pieces
// ---------------------------------------------------------------------------
template <int NT, bool PAIRED>
__device__ __forceinline__ int slot_j(int t, int b) {
    if constexpr (!PAIRED) {
        return t + b * NT;
    } else {
        return b == 0 ? t : (t == 0 ? NT : 2 * NT - t);
    }
}

template <typename R>
struct Cx {
    R re, im;
};

// first twiddle of every slot of a pass (loaded ahead of the pass, so that its
// latency overlaps the preceding LDS exchange)
template <typename R, int LOGNC, int LOGV, int LR, int LOGNS, bool PAIRED>
__device__ __forceinline__ void pass_twiddles(const Cx<R> *__restrict__ tw, int t, Cx<R> *w) {
    constexpr int V = 1 << LOGV, RAD = 1 << LR, B = V / RAD, NT = (1 << LOGNC) / V;
    if constexpr (LOGNS > 0) {
#pragma unroll
        for (int b = 0; b < B; b++) {
            const int j = slot_j<NT, PAIRED>(t, b);
            const int k = j & ((1 << LOGNS) - 1);
            w[b] = tw[k << (LOGNC - LOGNS - LR)];
        }
    }
}

// twiddle + butterfly of every slot of a pass (radix 2^LR, sub-FFT size 2^LOGNS)
template <typename R, int LOGNC, int LOGV, int LR, int LOGNS, bool PAIRED>
__device__ __forceinline__ void pass_compute(R *re, R *im, const Cx<R> *w, int t) {
    constexpr int V = 1 << LOGV, RAD = 1 << LR, B = V / RAD;
#pragma unroll
    for (int b = 0; b < B; b++) {
        R *vr = re + b * RAD;
        R *vi = im + b * RAD;
        if constexpr (LOGNS > 0) {
            const Cx<R> w1 = w[b];
            R wr = w1.re, wi = w1.im;
#pragma unroll
            for (int r = 1; r < RAD; r++) {
                const R xr = vr[r], xi = vi[r];
                vr[r] = fma(xr, wr, -(xi * wi));
                vi[r] = fma(xr, wi, xi * wr);
                if (r + 1 < RAD) {
                    const R nr = fma(wr, w1.re, -(wi * w1.im));
                    const R ni = fma(wr, w1.im, wi * w1.re);
                    wr = nr;
                    wi = ni;
                }
            }
        }
        dft<R, RAD>(vr, vi);
    }
}

// write a pass's outputs (one plane) to LDS.  With sub-FFT size Ns a multiple
// of the pad period P = 2^kPadShift the padded address is linear in r
// (lpad(base + r Ns) = lpad(base) + r Ns (P+1)/P), so one address per slot
// plus instruction offsets serve all r.
template <typename R, int LOGNC, int LOGV, int LR, int LOGNS, bool PAIRED>
__device__ __forceinline__ void pass_store(R *__restrict__ plane, const R *v, int t) {
    constexpr int V = 1 << LOGV, RAD = 1 << LR, B = V / RAD, NT = (1 << LOGNC) / V;
    constexpr int NS = 1 << LOGNS;
#pragma unroll
    for (int b = 0; b < B; b++) {
        const int j = slot_j<NT, PAIRED>(t, b);
        const int k = j & (NS - 1);
        const int base = ((j >> LOGNS) << (LOGNS + LR)) + k;
        if constexpr (NS >= (1 << kPadShift) || (LOGNS == 0 && LR <= kPadShift)) {
            // LOGNS == 0, RAD <= 2^kPadShift: base = j RAD, so base % 2^kPadShift + r never crosses a pad boundary
            R *p = plane + lpad(base);
#pragma unroll
            for (int r = 0; r < RAD; r++) p[r * (NS + (NS >> kPadShift))] = v[b * RAD + r];
        } else {
#pragma unroll
            for (int r = 0; r < RAD; r++) plane[lpad(base + (r << LOGNS))] = v[b * RAD + r];
        }
    }
}

// read the inputs of a pass (one plane) from LDS; the stride NC/RAD is a
// multiple of the pad period, so the padded address is linear in r
template <typename R, int LOGNC, int LOGV, int LR, bool PAIRED>
__device__ __forceinline__ void pass_load(const R *__restrict__ plane, R *v, int t) {
    constexpr int V = 1 << LOGV, RAD = 1 << LR, B = V / RAD, NT = (1 << LOGNC) / V;
    constexpr int S = (1 << LOGNC) >> LR;
    static_assert(S % (1 << kPadShift) == 0, "pass stride must be a multiple of the pad period");
#pragma unroll
    for (int b = 0; b < B; b++) {
        const int j = slot_j<NT, PAIRED>(t, b);
        const R *p = plane + lpad(j);
#pragma unroll
        for (int r = 0; r < RAD; r++) v[b * RAD + r] = p[r * (S + (S >> kPadShift))];
    }
}

template <typename R, int LOGNC, int LOGV, int LOGR, bool TWO, bool INV, int P>
struct Passes {
    // w: the first twiddles of pass P's slots (loaded by the caller)
    static __device__ __forceinline__ void run(R *re, R *im, R *lds, const Cx<R> *__restrict__ tw, int t,
                                               const Cx<R> *w) {
        constexpr int NP = n_passes(LOGNC, LOGV, LOGR);
        constexpr int LR = sched_bits(LOGNC, LOGV, LOGR, INV, P);
        constexpr int LOGNS = sched_logns(LOGNC, LOGV, LOGR, INV, P);
        constexpr bool PAIRED = INV ? (P == 0) : (P == NP - 1);
        pass_compute<R, LOGNC, LOGV, LR, LOGNS, PAIRED>(re, im, w, t);
        if constexpr (P + 1 < NP) {
            constexpr int LRN = sched_bits(LOGNC, LOGV, LOGR, INV, P + 1);
            constexpr int LOGNSN = sched_logns(LOGNC, LOGV, LOGR, INV, P + 1);
            constexpr bool PAIREDN = INV ? false : (P + 1 == NP - 1);
            constexpr int NCP = lds_elems(1 << LOGNC);
            Cx<R> wn[(1 << LOGV) >> LRN];
            pass_twiddles<R, LOGNC, LOGV, LRN, LOGNSN, PAIREDN>(tw, t, wn);
            if constexpr (TWO) {
                pass_store<R, LOGNC, LOGV, LR, LOGNS, PAIRED>(lds, re, t);
                pass_store<R, LOGNC, LOGV, LR, LOGNS, PAIRED>(lds + NCP, im, t);
                __syncthreads();
                pass_load<R, LOGNC, LOGV, LRN, PAIREDN>(lds, re, t);
                pass_load<R, LOGNC, LOGV, LRN, PAIREDN>(lds + NCP, im, t);
                __syncthreads();
            } else {
                pass_store<R, LOGNC, LOGV, LR, LOGNS, PAIRED>(lds, re, t);
                __syncthreads();
                pass_load<R, LOGNC, LOGV, LRN, PAIREDN>(lds, re, t);
                __syncthreads();
                pass_store<R, LOGNC, LOGV, LR, LOGNS, PAIRED>(lds, im, t);
                __syncthreads();
                pass_load<R, LOGNC, LOGV, LRN, PAIREDN>(lds, im, t);
                __syncthreads();
            }
            Passes<R, LOGNC, LOGV, LOGR, TWO, INV, P + 1>::run(re, im, lds, tw, t, wn);
        }
    }
};

// ---------------------------------------------------------------------------
// the tile kernel
// ---------------------------------------------------------------------------
// the lane index as a value the compiler cannot tie to the one computed at the
// kernel's start: the inverse transform derives its exchange coordinates from
// it again instead of keeping them live (and spilled) across the spectral step
__device__ __forceinline__ int lane_again() {
    int t = (int) threadIdx.x;
    __asm__ volatile("" : "+v"(t));
    return t;
}

// slices of the symmetric mode's output-mean sum (see k_block_mean)
constexpr int kEdgeSlices = 8;

// output mean of block b of length n: the slice sums added in slice order
__device__ __forceinline__ double mean_out_of(const double *__restrict__ edge_part, int b, int64_t n) {
    double s = 0.0;
    for (int k = 0; k < kEdgeSlices; k++) s += edge_part[(int64_t) b * kEdgeSlices + k];
    return s / (double) n;
}

struct TileArgs {
    const void *y;           // TOD in (R)
    void *z;                 // TOD out (R)
    const TileDesc *tiles;
    const void *coef;        // Cx<R> [nfilters * NC]
    const void *tw;          // Cx<R> [NC]
    const double *mean;      // [nb]
    const double *edge_part; // [nb * kEdgeSlices] output-mean slice sums (0 unless MGPU_OFFSET_SYMMETRIC)
    int lambda;
    int L;
    const int *block_seg;    // [nb] segment of each block (-1: none), with
    const int *state;        // the PCG state per segment: stopped segments are skipped (nullptr: none)
};

template <typename R>
__device__ __forceinline__ void spectral_pair(R &pr, R &pi, R &qr, R &qi, Cx<R> cp, Cx<R> cq) {
    // Z'[k] = a_k P + i b_k conj(Q); Z'[k'] = a_k' Q + i b_k' conj(P); then conjugate
    const R npr = fma(cp.re, pr, cp.im * qi);
    const R npi = -fma(cp.re, pi, cp.im * qr);
    const R nqr = fma(cq.re, qr, cq.im * pi);
    const R nqi = -fma(cq.re, qi, cq.im * pr);
    pr = npr;
    pi = npi;
    qr = nqr;
    qi = nqi;
}

template <typename R>
__device__ __forceinline__ void spectral_self(R &pr, R &pi, Cx<R> c) {
    const R npr = fma(c.re, pr, c.im * pi);
    const R npi = -fma(c.re, pi, c.im * pr);
    pr = npr;
    pi = npi;
}

template <typename R, int LOGNC, int LOGV, int LOGR, bool TWO, int MINW>
__global__ __launch_bounds__((1 << LOGNC) >> LOGV, MINW) void k_tiles(TileArgs a) {
    constexpr int NC = 1 << LOGNC, V = 1 << LOGV, NT = NC / V;
    constexpr int NP = n_passes(LOGNC, LOGV, LOGR);
    constexpr int LR0 = fwd_bits(LOGNC, LOGV, LOGR, 0);
    static_assert(LOGV - 1 <= LOGR, "paired pass radix exceeds the maximum radix");
    // FP64 tiles with 16+ values per lane run close to the register budget:
    // keep the scheduler from hoisting every load of a phase at once
    constexpr bool SCHED_FENCE = sizeof(R) == 8 && LOGV >= 4;
    constexpr int RAD0 = 1 << LR0, B0 = V / RAD0;
    constexpr int RL = V / 2;  // radix of the paired passes
    extern __shared__ unsigned char smem_raw[];
    R *lds = reinterpret_cast<R *>(smem_raw);

    const int t = threadIdx.x;
    const TileDesc d = a.tiles[blockIdx.x];
    if (a.state) {  // the tile's segment has stopped (uniform over the workgroup)
        const int sg = a.block_seg[d.block];
        if (sg >= 0 && a.state[sg]) return;
    }
    const R *__restrict__ y = reinterpret_cast<const R *>(a.y);
    R *__restrict__ z = reinterpret_cast<R *>(a.z);
    const Cx<R> *__restrict__ tw = reinterpret_cast<const Cx<R> *>(a.tw);
    const Cx<R> *__restrict__ cf = reinterpret_cast<const Cx<R> *>(a.coef) + (int64_t) d.filter * NC;
    const R mean = (R) a.mean[d.block];
    const R mean_out = (R) mean_out_of(a.edge_part, d.block, d.blk_hi - d.blk_lo);
    const int64_t in0 = d.out_lo - (a.lambda - 1);

    // block-relative 32-bit offsets from a uniform base: loads are clamped into
    // the block and masked, so the first pass has no divergent branches
    const R *__restrict__ yb = y + d.blk_lo;
    R *__restrict__ zb = z + d.blk_lo;
    const int nblk = (int) (d.blk_hi - d.blk_lo);
    const int rel0 = (int) (in0 - d.blk_lo);

    // interior tile (the whole input window inside the block, uniform per
    // workgroup): plain loads and stores without clamping or masking
    const bool interior = rel0 >= 0 && rel0 + 2 * NC <= nblk;

    R re[V], im[V];
    // first forward pass inputs: z[n] = x[2n] + i x[2n+1], n = j + r NC/RAD0
    if (interior) {
        const R *__restrict__ yt = yb + rel0;
#pragma unroll
        for (int b = 0; b < B0; b++) {
            const int j = t + b * NT;
#pragma unroll
            for (int r = 0; r < RAD0; r++) {
                const int n = j + r * (NC / RAD0);
                re[b * RAD0 + r] = yt[2 * n] - mean;
                im[b * RAD0 + r] = yt[2 * n + 1] - mean;
            }
            if constexpr (SCHED_FENCE) __builtin_amdgcn_sched_barrier(0);
        }
    } else {
#pragma unroll
        for (int b = 0; b < B0; b++) {
            const int j = t + b * NT;
#pragma unroll
            for (int r = 0; r < RAD0; r++) {
                const int n = j + r * (NC / RAD0);
                const int i0 = rel0 + 2 * n, i1 = i0 + 1;
                const R v0 = yb[(unsigned) min(max(i0, 0), nblk - 1)];
                const R v1 = yb[(unsigned) min(max(i1, 0), nblk - 1)];
                re[b * RAD0 + r] = ((unsigned) i0 < (unsigned) nblk) ? v0 - mean : (R) 0;
                im[b * RAD0 + r] = ((unsigned) i1 < (unsigned) nblk) ? v1 - mean : (R) 0;
            }
            if constexpr (SCHED_FENCE) __builtin_amdgcn_sched_barrier(0);
        }
    }
    Cx<R> w0[1];  // the first pass of either direction has no twiddles (sub-FFT size 1)
    static_assert(sched_logns(LOGNC, LOGV, LOGR, false, 0) == 0 && sched_logns(LOGNC, LOGV, LOGR, true, 0) == 0,
                  "first passes are twiddle-free");
    Passes<R, LOGNC, LOGV, LOGR, TWO, false, 0>::run(re, im, lds, tw, t, w0);

    // spectral step on the paired slots: slot 0 holds k = t + r 2NT, slot 1 holds
    // k = j1 + r 2NT with j1 = 2NT - t (t > 0) or NT (t == 0)
    if (t != 0 && !SCHED_FENCE) {
#pragma unroll
        for (int r = 0; r < RL; r++) {
            const int k = t + r * 2 * NT;
            spectral_pair<R>(re[r], im[r], re[RL + RL - 1 - r], im[RL + RL - 1 - r], cf[k], cf[NC - k]);
        }
    } else if (t != 0) {
        // the coefficients are loaded SG pairs at a time, the next group's loads
        // issued before the current group's arithmetic (with a scheduling fence
        // after every pair their L2 latencies would add up one after the other)
        constexpr int SG = RL < MGPU_SPEC_GROUP ? RL : MGPU_SPEC_GROUP;
        static_assert(RL % SG == 0, "coefficient groups must tile the paired slots");
        Cx<R> cp[2][SG], cq[2][SG];
#pragma unroll
        for (int u = 0; u < SG; u++) {
            const int k = t + u * 2 * NT;
            cp[0][u] = cf[k];
            cq[0][u] = cf[NC - k];
        }
#pragma unroll
        for (int g0 = 0; g0 < RL; g0 += SG) {
            const int cur = (g0 / SG) & 1;
            if (g0 + SG < RL) {
#pragma unroll
                for (int u = 0; u < SG; u++) {
                    const int k = t + (g0 + SG + u) * 2 * NT;
                    cp[cur ^ 1][u] = cf[k];
                    cq[cur ^ 1][u] = cf[NC - k];
                }
            }
#pragma unroll
            for (int u = 0; u < SG; u++) {
                const int r = g0 + u;
                spectral_pair<R>(re[r], im[r], re[RL + RL - 1 - r], im[RL + RL - 1 - r], cp[cur][u], cq[cur][u]);
            }
            if constexpr (SCHED_FENCE) __builtin_amdgcn_sched_barrier(0);
        }
    } else {
        spectral_self<R>(re[0], im[0], cf[0]);
        spectral_self<R>(re[RL / 2], im[RL / 2], cf[(RL / 2) * 2 * NT]);
#pragma unroll
        for (int r = 1; r < RL / 2; r++) {
            const int k = r * 2 * NT;
            spectral_pair<R>(re[r], im[r], re[RL - r], im[RL - r], cf[k], cf[NC - k]);
        }
#pragma unroll
        for (int r = 0; r < RL / 2; r++) {
            const int k = NT + r * 2 * NT;
            spectral_pair<R>(re[RL + r], im[RL + r], re[RL + RL - 1 - r], im[RL + RL - 1 - r], cf[k], cf[NC - k]);
        }
    }

    Passes<R, LOGNC, LOGV, LOGR, TWO, true, 0>::run(re, im, lds, tw, t, w0);

    // last inverse pass (radix RAD0, not paired): position n = j + r NC/RAD0 holds
    // conj(y[2n] + i y[2n+1]); tile sample i -> block offset rel0 + i
    const int lo = a.lambda - 1, hi = a.lambda - 1 + a.L;
    if (interior) {
        // tile sample i is an output iff (unsigned) (i - lo) < L; it lies inside the block
        // Whether a slot (b, r) holds outputs is decided for the whole workgroup
        // (its samples span [2(bNT + r NC/RAD0), +2NT)): inside the output
        // range every lane stores both values without a per-lane test
        R *__restrict__ zt = zb + rel0;
        const unsigned L = (unsigned) a.L;
#pragma unroll
        for (int b = 0; b < B0; b++) {
            const int j = t + b * NT;
#pragma unroll
            for (int r = 0; r < RAD0; r++) {
                const int i0 = 2 * (j + r * (NC / RAD0));
                const int s_lo = 2 * (b * NT + r * (NC / RAD0)), s_hi = s_lo + 2 * NT;  // uniform
                if (s_lo >= lo && s_hi <= hi) {
                    zt[i0] = re[b * RAD0 + r] - mean_out;
                    zt[i0 + 1] = -im[b * RAD0 + r] - mean_out;
                } else if (s_hi > lo && s_lo < hi) {
                    if ((unsigned) (i0 - lo) < L) zt[i0] = re[b * RAD0 + r] - mean_out;
                    if ((unsigned) (i0 + 1 - lo) < L) zt[i0 + 1] = -im[b * RAD0 + r] - mean_out;
                }
            }
        }
    } else {
#pragma unroll
        for (int b = 0; b < B0; b++) {
            const int j = t + b * NT;
#pragma unroll
            for (int r = 0; r < RAD0; r++) {
                const int n = j + r * (NC / RAD0);
                const int i0 = 2 * n;
                const int o0 = rel0 + i0;
                if (i0 >= lo && i0 < hi && o0 < nblk) zb[(unsigned) o0] = re[b * RAD0 + r] - mean_out;
                if (i0 + 1 >= lo && i0 + 1 < hi && o0 + 1 < nblk)
                    zb[(unsigned) (o0 + 1)] = -im[b * RAD0 + r] - mean_out;
            }
        }
    }
    (void) NP;
}

// ---------------------------------------------------------------------------
// N = 32768 tiles with wave-local inner passes (k_tiles_wl)
// ---------------------------------------------------------------------------
// NC = 16384 complex points on 512 lanes of 32 values.  Forward transform:
//   A   lane t: 32-point DFT over r of z[t + 512 r], then twiddle w_NC^(t k1)
//   E1  workgroup exchange: wave w gathers Y[k1][t] (all t) for its four k1
//   B1  lane (k1, t1): 32-point DFT over t2 of Y[k1][t1 + 16 t2], then twiddle
//       w_512^(t1 k2a)
//   E2  exchange inside the wave: the lane gathers its two (k1, k2a) columns
//   B2  16-point DFT over t1: X[k1 + 32 k2a + 1024 k2b], k2b = 0..15
// A lane's two columns are each other's mirror (X[k], X[NC - k]), so the
// real-FFT untangling and the spectral step run in registers as in k_tiles.
// The inverse runs B2, E2, B1, E1, A backwards (twiddles before the DFTs,
// conj trick).  Only E1 synchronises the workgroup: between the two E1
// exchanges of a tile the waves run independently, so one wave's LDS traffic
// overlaps another's arithmetic on the same SIMD.  The spectral coefficients
// are read from a copy permuted into the lanes' order (k_coef_wl): coalesced.
constexpr int kWlNC = 16384, kWlNT = 512;
constexpr int kWlPlaneStride = 33;                // E1: element (t, k1) at t * 33 + k1
constexpr int kWlRow = 17, kWlCol = 32 * kWlRow;  // E2: (kidx, k2a, t1) at kidx * 544 + k2a * 17 + t1
constexpr int kWlRegion = 4 * kWlCol;             // elements of one wave's E2 region
constexpr int kWlLdsElems = 8 * kWlRegion;        // >= 512 * 33, the E1 plane

// the four k1 of wave w (wave 0: 0, 16, 1, 31; wave j: 2j, 32-2j, 2j+1, 31-2j)
__device__ __forceinline__ int wl_k1(int w, int kidx) {
    if (w == 0) return kidx == 0 ? 0 : kidx == 1 ? 16 : kidx == 2 ? 1 : 31;
    const int a = 2 * w, b = 2 * w + 1;
    return kidx == 0 ? a : kidx == 1 ? 32 - a : kidx == 2 ? b : 32 - b;
}

// the two mirrored columns (kidx in the wave's k1 list, k2a) of lane l of wave w;
// lane 63 of wave 0 holds the two self-mirrored columns (0, 0) and (0, 16)
__device__ __forceinline__ void wl_pair(int w, int l, int &kp, int &ap, int &kq, int &aq) {
    if (w > 0) {
        const int h = l >> 5, c = l & 31;
        kp = 2 * h;          // k1 = a (h = 0) or b (h = 1)
        kq = 2 * h + 1;      // 32 - a or 32 - b
        ap = c;
        aq = 31 - c;
    } else if (l < 32) {     // k1 = 1 and its mirror 31
        kp = 2;
        kq = 3;
        ap = l;
        aq = 31 - l;
    } else if (l < 48) {     // k1 = 16 with itself
        kp = kq = 1;
        ap = l - 32;
        aq = 31 - (l - 32);
    } else if (l < 63) {     // k1 = 0 with itself
        kp = kq = 0;
        ap = l - 47;
        aq = 32 - (l - 47);
    } else {
        kp = kq = 0;
        ap = 0;
        aq = 16;
    }
}

// spectral index k of the i-th coefficient of lane tid (i < 16: first column,
// k2b = i; i >= 16: second column, k2b = i - 16)
__device__ __forceinline__ int wl_coef_k(int tid, int i) {
    int kp, ap, kq, aq;
    wl_pair(tid >> 6, tid & 63, kp, ap, kq, aq);
    const int w = tid >> 6;
    return i < 16 ? wl_k1(w, kp) + 32 * ap + 1024 * i : wl_k1(w, kq) + 32 * aq + 1024 * (i - 16);
}

// coefficient tables of every filter in the lanes' order: entry i * 512 + tid
__global__ void k_coef_wl(const double *__restrict__ c64, int nf, double *__restrict__ w64, float *__restrict__ w32) {
    const int64_t idx = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t) nf * kWlNC) return;
    const int f = (int) (idx / kWlNC), j = (int) (idx % kWlNC);
    const int i = j / kWlNT, tid = j % kWlNT;
    const int64_t src = ((int64_t) f * kWlNC + wl_coef_k(tid, i)) * 2;
    w64[2 * idx] = c64[src];
    w64[2 * idx + 1] = c64[src + 1];
    w32[2 * idx] = (float) c64[src];
    w32[2 * idx + 1] = (float) c64[src + 1];
}

// v[k] *= w^k for k = 1..N-1, w = tw[e] (recurrence, as in pass_compute)
template <typename R, int N>
__device__ __forceinline__ void wl_twiddle(R *re, R *im, Cx<R> w1) {
    R wr = w1.re, wi = w1.im;
#pragma unroll
    for (int k = 1; k < N; k++) {
        const R xr = re[k], xi = im[k];
        re[k] = fma(xr, wr, -(xi * wi));
        im[k] = fma(xr, wi, xi * wr);
        if (k + 1 < N) {
            const R nr = fma(wr, w1.re, -(wi * w1.im));
            const R ni = fma(wr, w1.im, wi * w1.re);
            wr = nr;
            wi = ni;
        }
    }
}

// keep the compiler from moving LDS accesses across a point inside a wave
__device__ __forceinline__ void wl_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename R, int MINW>
__global__ __launch_bounds__(kWlNT, MINW) void k_tiles_wl(TileArgs a) {
    constexpr int NC = kWlNC, NT = kWlNT;
    // scheduling fences between phases in both precisions: they keep the loads of
    // a later phase from being hoisted over the current one (FP32 at 128 VGPRs:
    // 21 registers spilled instead of 33, 2 % faster)
    constexpr bool SCHED_FENCE = true;
    extern __shared__ unsigned char smem_raw[];
    R *lds = reinterpret_cast<R *>(smem_raw);

    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const TileDesc d = a.tiles[blockIdx.x];
    if (a.state) {  // the tile's segment has stopped (uniform over the workgroup)
        const int sg = a.block_seg[d.block];
        if (sg >= 0 && a.state[sg]) return;
    }
    const R *__restrict__ y = reinterpret_cast<const R *>(a.y);
    R *__restrict__ z = reinterpret_cast<R *>(a.z);
    const Cx<R> *__restrict__ tw = reinterpret_cast<const Cx<R> *>(a.tw);
    const Cx<R> *__restrict__ cw = reinterpret_cast<const Cx<R> *>(a.coef) + (int64_t) d.filter * NC + t;
    const R mean = (R) a.mean[d.block];
    const int64_t in0 = d.out_lo - (a.lambda - 1);
    const R *__restrict__ yb = y + d.blk_lo;
    R *__restrict__ zb = z + d.blk_lo;
    const int nblk = (int) (d.blk_hi - d.blk_lo);
    const int rel0 = (int) (in0 - d.blk_lo);
    const bool interior = rel0 >= 0 && rel0 + 2 * NC <= nblk;

    // lane roles: (kidx, t1) in B1 (its two columns in B2 are set before E2)
    const int kidx = l >> 4, t1 = l & 15;
    R *__restrict__ reg = lds + w * kWlRegion;

    R re[32], im[32];
    // ---- A: z[n] = x[2n] + i x[2n+1], n = t + 512 r ----
    if (interior) {
        const R *__restrict__ yt = yb + rel0;
#pragma unroll
        for (int r = 0; r < 32; r++) {
            const int n = t + r * NT;
            re[r] = yt[2 * n] - mean;
            im[r] = yt[2 * n + 1] - mean;
        }
    } else {
#pragma unroll
        for (int r = 0; r < 32; r++) {
            const int n = t + r * NT;
            const int i0 = rel0 + 2 * n, i1 = i0 + 1;
            const R v0 = yb[(unsigned) min(max(i0, 0), nblk - 1)];
            const R v1 = yb[(unsigned) min(max(i1, 0), nblk - 1)];
            re[r] = ((unsigned) i0 < (unsigned) nblk) ? v0 - mean : (R) 0;
            im[r] = ((unsigned) i1 < (unsigned) nblk) ? v1 - mean : (R) 0;
        }
    }
    if constexpr (SCHED_FENCE) __builtin_amdgcn_sched_barrier(0);
    dft<R, 32>(re, im);
    if constexpr (SCHED_FENCE) __builtin_amdgcn_sched_barrier(0);
    wl_twiddle<R, 32>(re, im, tw[t]);
    // ---- E1 ----
    const int k1 = wl_k1(w, kidx);
#pragma unroll
    for (int p = 0; p < 2; p++) {
        R *v = p ? im : re;
#pragma unroll
        for (int q = 0; q < 32; q++) lds[t * kWlPlaneStride + q] = v[q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 32; q++) v[q] = lds[(t1 + 16 * q) * kWlPlaneStride + k1];
        __syncthreads();
    }
    // ---- B1 ----
    dft<R, 32>(re, im);
    if constexpr (SCHED_FENCE) __builtin_amdgcn_sched_barrier(0);
    wl_twiddle<R, 32>(re, im, tw[32 * t1]);
    // ---- E2 (inside the wave) ----
    int kp, ap, kq, aq;
    wl_pair(w, l, kp, ap, kq, aq);
#pragma unroll
    for (int p = 0; p < 2; p++) {
        R *v = p ? im : re;
#pragma unroll
        for (int q = 0; q < 32; q++) reg[kidx * kWlCol + q * kWlRow + t1] = v[q];
        wl_wave_sync();
#pragma unroll
        for (int q = 0; q < 16; q++) {
            v[q] = reg[kp * kWlCol + ap * kWlRow + q];
            v[16 + q] = reg[kq * kWlCol + aq * kWlRow + q];
        }
        wl_wave_sync();
    }
    // ---- B2 and the spectral step ----
    dft<R, 16>(re, im);
    dft<R, 16>(re + 16, im + 16);
    if (w != 0 || l != 63) {
        // P[k2b] (k) mirrors Q[15 - k2b] (NC - k); coefficients 4 pairs at a time
#pragma unroll
        for (int g = 0; g < 16; g += 4) {
            Cx<R> cp[4], cq[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                cp[u] = cw[(g + u) * NT];
                cq[u] = cw[(16 + 15 - (g + u)) * NT];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int b = g + u;
                spectral_pair<R>(re[b], im[b], re[16 + 15 - b], im[16 + 15 - b], cp[u], cq[u]);
            }
            if constexpr (SCHED_FENCE) __builtin_amdgcn_sched_barrier(0);
        }
    } else {
        // (0, 0): k = 1024 k2b, k2b <-> 16 - k2b; (0, 16): k = 512 + 1024 k2b, k2b <-> 15 - k2b
        spectral_self<R>(re[0], im[0], cw[0]);
        spectral_self<R>(re[8], im[8], cw[8 * NT]);
#pragma unroll
        for (int b = 1; b < 8; b++)
            spectral_pair<R>(re[b], im[b], re[16 - b], im[16 - b], cw[b * NT], cw[(16 - b) * NT]);
#pragma unroll
        for (int b = 0; b < 8; b++)
            spectral_pair<R>(re[16 + b], im[16 + b], re[31 - b], im[31 - b], cw[(16 + b) * NT], cw[(31 - b) * NT]);
    }
    // ---- inverse B2: 16-point DFTs over k2b ----
    dft<R, 16>(re, im);
    dft<R, 16>(re + 16, im + 16);
    {
        // the lane's coordinates again for the inverse transform (same values)
        const int t = lane_again(), w = t >> 6, l = t & 63, kidx = l >> 4, t1 = l & 15;
        R *__restrict__ reg = lds + w * kWlRegion;
        int kp, ap, kq, aq;
        wl_pair(w, l, kp, ap, kq, aq);
        const int k1 = wl_k1(w, kidx);
        // ---- inverse E2 ----
#pragma unroll
        for (int p = 0; p < 2; p++) {
            R *v = p ? im : re;
#pragma unroll
            for (int q = 0; q < 16; q++) {
                reg[kp * kWlCol + ap * kWlRow + q] = v[q];
                reg[kq * kWlCol + aq * kWlRow + q] = v[16 + q];
            }
            wl_wave_sync();
#pragma unroll
            for (int q = 0; q < 32; q++) v[q] = reg[kidx * kWlCol + q * kWlRow + t1];
            wl_wave_sync();
        }
        // ---- inverse B1: twiddle, 32-point DFT over k2a ----
        wl_twiddle<R, 32>(re, im, tw[32 * t1]);
        if constexpr (SCHED_FENCE) __builtin_amdgcn_sched_barrier(0);
        dft<R, 32>(re, im);
        // ---- inverse E1 (every wave has finished reading its E2 region) ----
        __syncthreads();
#pragma unroll
        for (int p = 0; p < 2; p++) {
            R *v = p ? im : re;
#pragma unroll
            for (int q = 0; q < 32; q++) lds[(t1 + 16 * q) * kWlPlaneStride + k1] = v[q];
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 32; q++) v[q] = lds[t * kWlPlaneStride + q];
            if (p == 0) __syncthreads();
        }
        // ---- inverse A: twiddle, 32-point DFT over k1 -> conj(y) at n = t + 512 r ----
        wl_twiddle<R, 32>(re, im, tw[t]);
        if constexpr (SCHED_FENCE) __builtin_amdgcn_sched_barrier(0);
        dft<R, 32>(re, im);
        // loaded here rather than at the start: one register less over the whole tile
        const R mean_out = (R) mean_out_of(a.edge_part, d.block, d.blk_hi - d.blk_lo);
        const int lo = a.lambda - 1, hi = a.lambda - 1 + a.L;
        if (interior) {
            R *__restrict__ zt = zb + rel0;
            const unsigned L = (unsigned) a.L;
#pragma unroll
            for (int r = 0; r < 32; r++) {
                const int i0 = 2 * (t + r * NT);
                const int s_lo = 2 * r * NT, s_hi = s_lo + 2 * NT;  // samples of slot r (uniform)
                if (s_lo >= lo && s_hi <= hi) {
                    zt[i0] = re[r] - mean_out;
                    zt[i0 + 1] = -im[r] - mean_out;
                } else if (s_hi > lo && s_lo < hi) {
                    if ((unsigned) (i0 - lo) < L) zt[i0] = re[r] - mean_out;
                    if ((unsigned) (i0 + 1 - lo) < L) zt[i0 + 1] = -im[r] - mean_out;
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < 32; r++) {
                const int i0 = 2 * (t + r * NT);
                const int o0 = rel0 + i0;
                if (i0 >= lo && i0 < hi && o0 < nblk) zb[(unsigned) o0] = re[r] - mean_out;
                if (i0 + 1 >= lo && i0 + 1 < hi && o0 + 1 < nblk) zb[(unsigned) (o0 + 1)] = -im[r] - mean_out;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// mean, short blocks, tail, lambda == 1, direct fallback
// ---------------------------------------------------------------------------
__device__ __forceinline__ int find_block_t(const int64_t *__restrict__ start, int nb, int64_t t) {
    if (t >= start[nb]) return nb;
    int lo = 0, hi = nb;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (start[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

template <typename R>
__global__ __launch_bounds__(256) void k_partials(const R *__restrict__ y, int64_t m,
                                                  const int64_t *__restrict__ start,
                                                  const int *__restrict__ chunk_block, int nb,
                                                  double *__restrict__ partials) {
    __shared__ double red[4];
    const int64_t c = blockIdx.x, c0 = c * kChunk;
    const int64_t c1 = min(c0 + (int64_t) kChunk, m);
    int b = chunk_block[c];
    for (; b < nb && start[b] < c1; b++) {
        const int64_t lo = start[b], hi = start[b + 1];
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < kChunk / 256; i++) {
            const int64_t t = c0 + threadIdx.x + (int64_t) i * 256;
            if (t < m && t >= lo && t < hi) s += (double) y[t];
        }
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        __syncthreads();
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) partials[c + b] = ((red[0] + red[1]) + red[2]) + red[3];
    }
}

__device__ __forceinline__ double tree256(double v, double *red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// mean[b] = (sum of the block's partials, fixed order) / n_b.
// With edge_part (MGPU_OFFSET_SYMMETRIC) also the mean of T (v - mean) over the
// block: T 1 equals its interior value c = T0 + 2 sum_k T_k except within
// lambda-1 of the block ends, where it is c + e_i with
// e_i = -tail(i) - tail(n-1-i), tail(j) = sum_{k>j} T_k; since v - mean sums
// to zero, mean(T (v - mean)) = (1/n) sum_i e_i (v_i - mean), a sum over the
// 2(lambda-1) edge samples only.  That sum is split over kEdgeSlices
// workgroups per block (blockIdx.y), each writing its slice's sum to
// edge_part[b * kEdgeSlices + slice]; the kernels that subtract the output mean
// add the slices in slice order and divide by the block length (mean_out_of).
template <typename R>
__global__ __launch_bounds__(256) void k_block_mean(const double *__restrict__ partials,
                                                    const int64_t *__restrict__ start, int nb,
                                                    double *__restrict__ mean, const R *__restrict__ v,
                                                    const double *__restrict__ tail,
                                                    const int64_t *__restrict__ tail_off,
                                                    const int *__restrict__ block_filter,
                                                    const int *__restrict__ block_kind,
                                                    const int *__restrict__ block_lambda,
                                                    double *__restrict__ edge_part,
                                                    const int *__restrict__ block_seg, const int *__restrict__ state) {
    __shared__ double red[256];
    const int b = blockIdx.x, sl = blockIdx.y;
    if (state && block_seg[b] >= 0 && state[block_seg[b]]) return;  // stopped segment
    const int64_t lo = start[b], hi = start[b + 1];
    const int64_t first = lo / kChunk + b, last = (hi - 1) / kChunk + b;
    double s = 0.0;
    for (int64_t i = first + threadIdx.x; i <= last; i += 256) s += partials[i];
    const double mu = tree256(s, red) / (double) (hi - lo);
    if (sl == 0 && threadIdx.x == 0) mean[b] = mu;
    if (!edge_part) return;
    double part = 0.0;
    if (block_kind[b] == 0) {  // short blocks: identity product, output mean 0
        const int n = (int) (hi - lo), L1 = block_lambda[b] - 1;
        const double *tl = tail + tail_off[block_filter[b]];
        const int a_end = min(L1, n);
        const int b_beg = max(n - L1, a_end);
        const int tot_e = a_end + (n - b_beg);
        const int per = (tot_e + kEdgeSlices - 1) / kEdgeSlices;
        const int i0 = sl * per, i1 = min(tot_e, i0 + per);
        double acc = 0.0;
        for (int i = i0 + (int) threadIdx.x; i < i1; i += 256) {
            const int k = i < a_end ? i : b_beg + (i - a_end);
            const double e = -(k < L1 ? tl[k] : 0.0) - (n - 1 - k < L1 ? tl[n - 1 - k] : 0.0);
            acc += e * ((double) v[lo + k] - mu);
        }
        part = tree256(acc, red);
    }
    if (threadIdx.x == 0) edge_part[(int64_t) b * kEdgeSlices + sl] = part;
}

// z = y - mean over the listed (short) blocks
template <typename R>
__global__ void k_mean_only(const R *__restrict__ y, R *__restrict__ z, const int64_t *__restrict__ start,
                            const int *__restrict__ blocks, const double *__restrict__ mean) {
    const int b = blocks[blockIdx.y];
    const int64_t lo = start[b], hi = start[b + 1];
    const R mu = (R) mean[b];
    for (int64_t t = lo + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < hi;
         t += (int64_t) gridDim.x * blockDim.x)
        z[t] = y[t] - mu;
}

template <typename R>
__global__ void k_copy_tail(const R *__restrict__ y, R *__restrict__ z, int64_t lo, int64_t hi) {
    int64_t t = lo + (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < hi) z[t] = y[t];
}

// lambda == 1: z = T_b[0] * y inside blocks (no mean removal), copy after them
// block of sample t for the threads of one workgroup: the block of the
// workgroup's first sample is found once (uniform, scalar loads), each lane then
// steps over the few block boundaries inside the workgroup
__device__ __forceinline__ int wg_block_of(const int64_t *__restrict__ start, int nb, int64_t t) {
    int b = find_block_t(start, nb, (int64_t) blockIdx.x * blockDim.x);
    while (b < nb && t >= start[b + 1]) b++;
    return b;
}

template <typename R>
__global__ __launch_bounds__(256) void k_diag(const R *__restrict__ y, R *__restrict__ z, int64_t m,
                                              const int64_t *__restrict__ start, int nb,
                                              const double *__restrict__ T0) {
    int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int b = wg_block_of(start, nb, t);
    z[t] = b < nb ? (R) T0[b] * y[t] : y[t];
}

// lambda == 1 with stbmmProd's offsets (toeplitz_wizard.c:86-95 ahead of the
// diagonal product): z = T_b[0] * (y - mean_b) inside blocks, copy after them
template <typename R>
__global__ __launch_bounds__(256) void k_diag_mean(const R *__restrict__ y, R *__restrict__ z, int64_t m,
                                                   const int64_t *__restrict__ start, int nb,
                                                   const double *__restrict__ T0, const double *__restrict__ mean) {
    int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int b = wg_block_of(start, nb, t);
    z[t] = b < nb ? (R) (T0[b] * ((double) y[t] - mean[b])) : y[t];
}

// direct banded product for bands too wide for an LDS tile (slow path)
template <typename R>
__global__ void k_direct(const R *__restrict__ y, R *__restrict__ z, const int64_t *__restrict__ start,
                         const int *__restrict__ blocks, const double *__restrict__ mean,
                         const double *__restrict__ edge_part, const double *__restrict__ invtt,
                         const int64_t *__restrict__ inv_off, const int *__restrict__ block_lambda) {
    const int b = blocks[blockIdx.y];
    const int64_t lo = start[b], hi = start[b + 1];
    const double mu = mean[b], mo = mean_out_of(edge_part, b, hi - lo);
    const double *T = invtt + inv_off[b];
    const int lambda = block_lambda[b];
    for (int64_t i = lo + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < hi;
         i += (int64_t) gridDim.x * blockDim.x) {
        const int64_t j0 = max(lo, i - (lambda - 1)), j1 = min(hi, i + lambda);
        double acc = 0.0;
        for (int64_t j = j0; j < j1; j++) acc += T[j > i ? j - i : i - j] * ((double) y[j] - mu);
        z[i] = (R) (acc - mo);
    }
}

// ---------------------------------------------------------------------------
// filter spectra: H = FFT_N(T_circ) (toeplitz.c:441-470), then (a_k, b_k)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_circ_fft(const double *__restrict__ invtt, const int64_t *__restrict__ src_off,
                                                    int lambda, int log2n, double2 *__restrict__ buf) {
    const int f = blockIdx.x;
    const int N = 1 << log2n;
    double2 *a = buf + (int64_t) f * N;
    const double *T = invtt + src_off[f];
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        int r = (int) (__builtin_bitreverse32((unsigned) i) >> (32 - log2n));
        double v = 0.0;
        if (r == 0)
            v = T[0];
        else if (r < lambda)
            v = T[r];
        else if (N - r < lambda)
            v = T[N - r];
        a[i] = make_double2(v, 0.0);
    }
    __syncthreads();
    for (int s = 1; s <= log2n; s++) {
        const int half = 1 << (s - 1);
        for (int idx = threadIdx.x; idx < N / 2; idx += blockDim.x) {
            const int grp = idx >> (s - 1), j = idx & (half - 1);
            const int i0 = (grp << s) + j, i1 = i0 + half;
            double sn, cs;
            sincospi(-(double) j / (double) half, &sn, &cs);
            const double2 u = a[i0], v = a[i1];
            const double vr = v.x * cs - v.y * sn, vi = v.x * sn + v.y * cs;
            a[i0] = make_double2(u.x + vr, u.y + vi);
            a[i1] = make_double2(u.x - vr, u.y - vi);
        }
        __syncthreads();
    }
}

__global__ void k_coef(const double2 *__restrict__ buf, int log2n, int nf, double *__restrict__ c64,
                       float *__restrict__ c32) {
    const int NC = 1 << (log2n - 1), N = 1 << log2n;
    const int64_t idx = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t) nf * NC) return;
    const int f = (int) (idx / NC), k = (int) (idx % NC);
    const double2 *H = buf + (int64_t) f * N;
    const double hs = 0.5 * (H[k].x + H[k + NC].x), hd = 0.5 * (H[k].x - H[k + NC].x);
    double sn, cs;
    sincospi(2.0 * (double) k / (double) N, &sn, &cs);
    const double inv = 1.0 / (double) NC;
    const double ak = (hs - hd * sn) * inv, bk = (hd * cs) * inv;
    c64[2 * idx] = ak;
    c64[2 * idx + 1] = bk;
    c32[2 * idx] = (float) ak;
    c32[2 * idx + 1] = (float) bk;
}

__global__ void k_twiddles(int NC, double *__restrict__ t64, float *__restrict__ t32) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= NC) return;
    double sn, cs;
    sincospi(-2.0 * (double) m / (double) NC, &sn, &cs);
    t64[2 * m] = cs;
    t64[2 * m + 1] = sn;
    t32[2 * m] = (float) cs;
    t32[2 * m + 1] = (float) sn;
}

// ---------------------------------------------------------------------------
// host side: plan + dispatch
// ---------------------------------------------------------------------------
static int max_log2n(bool fp32) { return fp32 ? 16 : 15; }

// Workgroup order of the tiles: workgroup b runs on XCD b % 8, so XCD x is
// dealt a contiguous run of tiles (in sample order).  Neighbouring tiles,
// which read half of each other's input, then run on one XCD at about the
// same time and share those reads (and their filter's spectrum) in its L2.
static std::vector<TileDesc> xcd_order(const std::vector<TileDesc> &t) { return deal_to_xcds(t); }
static int min_log2n() { return 10; }

// tile length choice: minimise N (log2 N + c) / L over the sizes that fit LDS
static int choose_log2n(int lambda, int64_t nmax, bool fp32, int forced);

int toeplitz_tile_log2(int lambda, int64_t nmax, int forced) {
    return lambda > 1 ? choose_log2n(lambda, nmax, false, forced) : 0;
}

static int choose_log2n(int lambda, int64_t nmax, bool fp32, int forced) {
    if (forced > 0) return forced;
    int best = -1;
    double best_cost = 1e300;
    for (int l2 = min_log2n(); l2 <= max_log2n(fp32); l2++) {
        const double N = (double) (1 << l2);
        const double L = N - 2.0 * (lambda - 1);
        if (L < 1) continue;
        // no benefit in tiles much longer than the longest block
        const double useful = std::min(L, (double) nmax);
        const double cost = N * (l2 + 8.0) / useful;
        if (cost < best_cost) {
            best_cost = cost;
            best = l2;
        }
    }
    return best;
}

// spectral tables of one tile length for the filters whose bands start at
// d_inv + src_off[i] (band lambda): FFT of the circulant embedding, then the
// (a_k, b_k) coefficients in both precisions and the twiddles
static int class_spectra(Problem &pb, const double *d_inv, const std::vector<int64_t> &src_off, int lambda, int l2,
                         double *c64, float *c32, double *t64, float *t32) {
    hipStream_t s = pb.stream;
    const int nf = (int) src_off.size();
    const int N = 1 << l2, NC = N / 2;
    int64_t *d_src;
    double2 *d_buf;
    MGPU_RET(pb.arena.alloc(&d_src, src_off.size()));
    MGPU_HIP(hipMemcpyAsync(d_src, src_off.data(), sizeof(int64_t) * src_off.size(), hipMemcpyHostToDevice, s));
    MGPU_RET(pb.arena.alloc(&d_buf, (size_t) nf * N));
    k_circ_fft<<<nf, 1024, 0, s>>>(d_inv, d_src, lambda, l2, d_buf);
    MGPU_LAUNCHED();
    double *c64_tmp = nullptr;
    if (!c64) MGPU_RET(pb.arena.alloc(&c64_tmp, (size_t) nf * NC * 2));
    k_coef<<<grid_for((int64_t) nf * NC, 256), 256, 0, s>>>(d_buf, l2, nf, c64 ? c64 : c64_tmp, c32);
    MGPU_LAUNCHED();
    double *t64_tmp = nullptr;
    if (!t64) MGPU_RET(pb.arena.alloc(&t64_tmp, (size_t) NC * 2));
    k_twiddles<<<grid_for(NC, 256), 256, 0, s>>>(NC, t64 ? t64 : t64_tmp, t32);
    MGPU_LAUNCHED();
    MGPU_HIP(hipStreamSynchronize(s));
    for (void *p : {(void *) d_buf, (void *) d_src, (void *) c64_tmp, (void *) t64_tmp})
        if (p) pb.arena.release(p);
    return MGPU_OK;
}

int toeplitz_plan(Problem &pb, const int *block_sizes, int nb, int lambda, const double *d_invtt,
                  const double *h_invtt, const int *block_lambda, const int *block_log2n) {
    ToeplitzPlan &T = pb.tp;
    hipStream_t s = pb.stream;
    MGPU_ARG(nb > 0, "nb_blocks=%d must be positive", nb);
    MGPU_ARG(lambda >= 1, "lambda=%d must be >= 1", lambda);
    // band of every block and the offset of its band in invtt
    std::vector<int> lam((size_t) nb, lambda);
    if (block_lambda)
        for (int b = 0; b < nb; b++) lam[b] = block_lambda[b];
    std::vector<int64_t> inv_off((size_t) nb + 1, 0);
    int lmax = 0, lmin = INT32_MAX;
    for (int b = 0; b < nb; b++) {
        MGPU_ARG(lam[b] >= 1, "block %d: lambda=%d must be >= 1", b, lam[b]);
        inv_off[b + 1] = inv_off[b] + lam[b];
        lmax = std::max(lmax, lam[b]);
        lmin = std::min(lmin, lam[b]);
    }
    MGPU_ARG(lmin == lmax || lmin > 1, "bands of 1 (diagonal noise) cannot be mixed with wider bands");
    T.lambda = lmax;
    T.mixed_lambda = lmin != lmax;
    T.nb = nb;
    T.m = pb.pt.m;
    T.h_start.assign(nb + 1, 0);
    for (int b = 0; b < nb; b++) {
        MGPU_ARG(block_sizes[b] > 0, "local_blocks_sizes[%d]=%d must be positive", b, block_sizes[b]);
        T.h_start[b + 1] = T.h_start[b] + block_sizes[b];
    }
    T.covered = T.h_start[nb];
    MGPU_ARG(T.covered <= T.m, "sum of local_blocks_sizes (%lld) exceeds m (%lld)", (long long) T.covered,
             (long long) T.m);
    MGPU_RET(pb.arena.alloc(&T.d_start, (size_t) nb + 1));
    MGPU_HIP(hipMemcpyAsync(T.d_start, T.h_start.data(), sizeof(int64_t) * (nb + 1), hipMemcpyHostToDevice, s));
    MGPU_RET(pb.arena.alloc(&T.d_block_lambda, (size_t) nb));
    MGPU_HIP(hipMemcpyAsync(T.d_block_lambda, lam.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
    // host copy of invtt (dedupe filters, T0)
    const int64_t ninv = inv_off[nb];
    std::vector<double> hinv;
    const double *inv = h_invtt;
    if (!inv) {
        hinv.resize((size_t) ninv);
        MGPU_HIP(hipMemcpyAsync(hinv.data(), d_invtt, sizeof(double) * hinv.size(), hipMemcpyDeviceToHost, s));
        MGPU_HIP(hipStreamSynchronize(s));
        inv = hinv.data();
    }
    std::vector<double> T0(nb);
    for (int b = 0; b < nb; b++) T0[b] = inv[inv_off[b]];
    MGPU_RET(pb.arena.alloc(&T.d_T0, (size_t) nb));
    MGPU_HIP(hipMemcpyAsync(T.d_T0, T0.data(), sizeof(double) * nb, hipMemcpyHostToDevice, s));

    T.nchunks = (T.m + kChunk - 1) / kChunk;
    MGPU_RET(pb.arena.alloc(&T.d_partials, (size_t) (T.nchunks + nb)));
    {
        // first block overlapping each chunk of kChunk samples (nb past the covered range)
        std::vector<int> cb((size_t) T.nchunks);
        int b = 0;
        for (int64_t c = 0; c < T.nchunks; c++) {
            const int64_t t0 = c * kChunk;
            while (b < nb && T.h_start[b + 1] <= t0) b++;
            cb[(size_t) c] = b;
        }
        MGPU_RET(pb.arena.alloc(&T.d_chunk_block, (size_t) T.nchunks));
        MGPU_HIP(hipMemcpyAsync(T.d_chunk_block, cb.data(), sizeof(int) * cb.size(), hipMemcpyHostToDevice, s));
        MGPU_HIP(hipStreamSynchronize(s));
    }
    MGPU_RET(pb.arena.alloc(&T.d_mean, (size_t) nb));
    MGPU_RET(pb.arena.alloc(&T.d_edge_part, (size_t) nb * kEdgeSlices));
    MGPU_HIP(hipMemsetAsync(T.d_mean, 0, sizeof(double) * nb, s));
    MGPU_HIP(hipMemsetAsync(T.d_edge_part, 0, sizeof(double) * nb * kEdgeSlices, s));
    T.offset_mode = pb.opt.offset_mode;
    MGPU_ARG(T.offset_mode >= MGPU_OFFSET_REFERENCE && T.offset_mode <= MGPU_OFFSET_NONE, "unknown offset_mode %d",
             T.offset_mode);
    if (lmax == 1) {
        MGPU_HIP(hipStreamSynchronize(s));
        return MGPU_OK;
    }

    const bool fp32 = pb.opt.precision == MGPU_MIXED;
    // dedupe filters by band and content
    std::vector<int> filter_src;  // first block of every filter
    T.h_block_filter.assign(nb, -1);
    {
        std::unordered_map<uint64_t, std::vector<int>> by_hash;
        for (int b = 0; b < nb; b++) {
            const double *Tb = inv + inv_off[b];
            const int lb = lam[b];
            // hash of the band and <= 256 evenly spaced values (candidates are confirmed by memcmp)
            uint64_t h = 1469598103934665603ull ^ (uint64_t) lb;
            const int step = lb > 256 ? lb / 256 : 1;
            for (int i = 0; i < lb; i += step) {
                uint64_t w;
                std::memcpy(&w, Tb + i, sizeof(w));
                h = (h ^ w) * 1099511628211ull;
                h ^= h >> 29;
            }
            int fid = -1;
            auto &cand = by_hash[h];
            for (int f : cand)
                if (lam[filter_src[f]] == lb &&
                    std::memcmp(inv + inv_off[filter_src[f]], Tb, sizeof(double) * lb) == 0) {
                    fid = f;
                    break;
                }
            if (fid < 0) {
                fid = (int) filter_src.size();
                filter_src.push_back(b);
                cand.push_back(fid);
            }
            T.h_block_filter[b] = fid;
        }
    }
    T.nfilters = (int) filter_src.size();
    // block-edge tails for the symmetric offset projection: tail_f[j] = sum_{k>j} T_f[k]
    {
        const int nf = T.nfilters;
        std::vector<int64_t> toff((size_t) nf + 1, 0);
        for (int f = 0; f < nf; f++) toff[f + 1] = toff[f] + lam[filter_src[f]];
        std::vector<double> tails((size_t) toff[nf], 0.0);
        for (int f = 0; f < nf; f++) {
            const int lf = lam[filter_src[f]];
            const double *Tf = inv + inv_off[filter_src[f]];
            double *tl = tails.data() + toff[f];
            double acc = 0.0;
            for (int j = lf - 1; j >= 0; j--) {
                tl[j] = acc;
                acc += Tf[j];
            }
        }
        std::vector<int> kind(nb);
        for (int b = 0; b < nb; b++) kind[b] = block_sizes[b] < lam[b] - 1 ? 1 : 0;
        MGPU_RET(pb.arena.alloc(&T.d_tail, tails.size()));
        MGPU_RET(pb.arena.alloc(&T.d_tail_off, (size_t) nf));
        MGPU_RET(pb.arena.alloc(&T.d_block_filter, (size_t) nb));
        MGPU_RET(pb.arena.alloc(&T.d_block_kind, (size_t) nb));
        MGPU_HIP(hipMemcpyAsync(T.d_tail, tails.data(), sizeof(double) * tails.size(), hipMemcpyHostToDevice, s));
        MGPU_HIP(hipMemcpyAsync(T.d_tail_off, toff.data(), sizeof(int64_t) * nf, hipMemcpyHostToDevice, s));
        MGPU_HIP(hipMemcpyAsync(T.d_block_filter, T.h_block_filter.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
        MGPU_HIP(hipMemcpyAsync(T.d_block_kind, kind.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
        MGPU_HIP(hipStreamSynchronize(s));
    }
    // tile length of every block: the caller's, or the choice for its band and
    // the longest block of that band; blocks without a tile length take the direct product
    std::map<int, int64_t> nmax_of;
    for (int b = 0; b < nb; b++) nmax_of[lam[b]] = std::max<int64_t>(nmax_of[lam[b]], block_sizes[b]);
    std::vector<int> l2b((size_t) nb, -1);
    for (int b = 0; b < nb; b++) {
        int l2 = block_log2n ? block_log2n[b]
                             : choose_log2n(lam[b], nmax_of[lam[b]], false, std::min(pb.opt.tile_log2, max_log2n(false)));
        if (l2 < min_log2n() || l2 > max_log2n(false) || (1 << l2) - 2 * (lam[b] - 1) < 1) l2 = -1;
        l2b[b] = l2;
    }
    // classes: (band, tile length) in increasing order; blocks shorter than their
    // band are only mean-subtracted
    T.h_short_blocks.clear();
    std::vector<int> direct_blocks;
    std::map<std::pair<int, int>, std::vector<int>> class_blocks;
    for (int b = 0; b < nb; b++) {
        if (block_sizes[b] < lam[b] - 1)
            T.h_short_blocks.push_back(b);
        else if (l2b[b] < 0)
            direct_blocks.push_back(b);
        else
            class_blocks[std::make_pair(lam[b], l2b[b])].push_back(b);
    }
    MGPU_ARG(direct_blocks.empty() || !T.mixed_lambda,
             "bands wider than a tile (direct product) cannot be batched with other bands");
    // the spectra need invtt on the device
    const double *d_inv = d_invtt;
    double *d_inv_tmp = nullptr;
    if (!d_inv && !class_blocks.empty()) {
        MGPU_RET(pb.arena.alloc(&d_inv_tmp, (size_t) ninv));
        MGPU_HIP(hipMemcpyAsync(d_inv_tmp, inv, sizeof(double) * (size_t) ninv, hipMemcpyHostToDevice, s));
        d_inv = d_inv_tmp;
    }
    T.cls.clear();
    for (auto &kv : class_blocks) {
        T.cls.emplace_back();
        TileClass &C = T.cls.back();
        C.lambda = kv.first.first;
        C.log2n = kv.first.second;
        const std::vector<int> &bl = kv.second;
        const int N = 1 << C.log2n, NC = N / 2;
        C.L = N - 2 * (C.lambda - 1);
        // the class's filters, numbered in order of first use
        std::unordered_map<int, int> local;
        std::vector<int64_t> src_off;
        int64_t nmax_c = 0;
        for (int b : bl) {
            nmax_c = std::max<int64_t>(nmax_c, block_sizes[b]);
            const int f = T.h_block_filter[b];
            if (!local.count(f)) {
                local[f] = (int) src_off.size();
                src_off.push_back(inv_off[filter_src[f]]);
            }
        }
        C.nfilters = (int) src_off.size();
        std::vector<TileDesc> tiles;
        for (int b : bl) {
            const int64_t nt = (block_sizes[b] + C.L - 1) / C.L;
            for (int64_t k = 0; k < nt; k++)
                tiles.push_back(TileDesc{T.h_start[b], T.h_start[b + 1], T.h_start[b] + k * C.L, b,
                                         local[T.h_block_filter[b]]});
        }
        C.ntiles = (int) tiles.size();
        MGPU_RET(pb.arena.alloc(&C.d_tiles, tiles.size()));
        C.h_tiles = tiles;
        tiles = xcd_order(tiles);
        MGPU_HIP(hipMemcpyAsync(C.d_tiles, tiles.data(), sizeof(TileDesc) * tiles.size(), hipMemcpyHostToDevice, s));
        MGPU_RET(pb.arena.alloc(&C.d_coef64, (size_t) C.nfilters * NC * 2));
        MGPU_RET(pb.arena.alloc(&C.d_coef32, (size_t) C.nfilters * NC * 2));
        MGPU_RET(pb.arena.alloc(&C.d_tw64, (size_t) NC * 2));
        MGPU_RET(pb.arena.alloc(&C.d_tw32, (size_t) NC * 2));
        MGPU_RET(class_spectra(pb, d_inv, src_off, C.lambda, C.log2n, C.d_coef64, C.d_coef32, C.d_tw64, C.d_tw32));
        if (NC == kWlNC) {  // the wave-local tile kernel reads the coefficients in its lanes' order
            MGPU_RET(pb.arena.alloc(&C.d_coefW64, (size_t) C.nfilters * NC * 2));
            MGPU_RET(pb.arena.alloc(&C.d_coefW32, (size_t) C.nfilters * NC * 2));
            k_coef_wl<<<grid_for((int64_t) C.nfilters * NC, 256), 256, 0, s>>>(C.d_coef64, C.nfilters, C.d_coefW64,
                                                                                 C.d_coefW32);
            MGPU_LAUNCHED();
        }
        // FP32 products of a mixed solve: a longer tile when it costs fewer FFT
        // operations per useful sample (FP32 halves the registers and LDS a tile needs)
        const int l2f = fp32 ? choose_log2n(C.lambda, nmax_c, true, pb.opt.tile_log2) : -1;
        if (l2f > C.log2n && l2f <= max_log2n(true)) {
            const int Nf = 1 << l2f, NCf = Nf / 2;
            C.log2n32 = l2f;
            C.L32 = Nf - 2 * (C.lambda - 1);
            std::vector<TileDesc> tf;
            for (int b : bl) {
                const int64_t nt = (block_sizes[b] + C.L32 - 1) / C.L32;
                for (int64_t k = 0; k < nt; k++)
                    tf.push_back(TileDesc{T.h_start[b], T.h_start[b + 1], T.h_start[b] + k * C.L32, b,
                                          local[T.h_block_filter[b]]});
            }
            C.ntiles32 = (int) tf.size();
            MGPU_RET(pb.arena.alloc(&C.d_tiles32, tf.size()));
            C.h_tiles32 = tf;
            tf = xcd_order(tf);
            MGPU_HIP(hipMemcpyAsync(C.d_tiles32, tf.data(), sizeof(TileDesc) * tf.size(), hipMemcpyHostToDevice, s));
            MGPU_RET(pb.arena.alloc(&C.d_coef32L, (size_t) C.nfilters * NCf * 2));
            MGPU_RET(pb.arena.alloc(&C.d_tw32L, (size_t) NCf * 2));
            MGPU_RET(class_spectra(pb, d_inv, src_off, C.lambda, l2f, nullptr, C.d_coef32L, nullptr, C.d_tw32L));
        }
    }
    if (d_inv_tmp) pb.arena.release(d_inv_tmp);
    if (!direct_blocks.empty()) {
        // keep invtt on device for the direct product
        double *d_inv_keep;
        MGPU_RET(pb.arena.alloc(&d_inv_keep, (size_t) ninv));
        MGPU_HIP(hipMemcpyAsync(d_inv_keep, inv, sizeof(double) * (size_t) ninv, hipMemcpyHostToDevice, s));
        T.d_direct_invtt = d_inv_keep;
        MGPU_RET(pb.arena.alloc(&T.d_inv_off, (size_t) nb));
        MGPU_HIP(hipMemcpyAsync(T.d_inv_off, inv_off.data(), sizeof(int64_t) * nb, hipMemcpyHostToDevice, s));
        T.h_direct_blocks = direct_blocks;
        MGPU_RET(pb.arena.alloc(&T.d_direct_blocks, direct_blocks.size()));
        MGPU_HIP(hipMemcpyAsync(T.d_direct_blocks, direct_blocks.data(), sizeof(int) * direct_blocks.size(),
                                hipMemcpyHostToDevice, s));
    }
    if (!T.h_short_blocks.empty()) {
        MGPU_RET(pb.arena.alloc(&T.d_short_blocks, T.h_short_blocks.size()));
        MGPU_HIP(hipMemcpyAsync(T.d_short_blocks, T.h_short_blocks.data(), sizeof(int) * T.h_short_blocks.size(),
                                hipMemcpyHostToDevice, s));
    }
    MGPU_HIP(hipStreamSynchronize(s));
    return MGPU_OK;
}

// ---------------------------------------------------------------------------
// tile kernel variants and their autotuning
// ---------------------------------------------------------------------------
// A variant fixes the values per lane (2^LOGV), the largest radix of a pass
// (2^LOGR), whether the real and imaginary planes are exchanged through LDS
// together (TWO) and the minimum resident waves per SIMD requested from the
// register allocator (MINW; 4 with 512-lane tiles lets two workgroups share a
// CU).  Variant 1 of every length is the default shape.
struct TileShape {
    int logv, logr;
    bool two;
    int minw;
    bool wave_local;  // k_tiles_wl (N = 32768): inner passes exchange inside the wave
};

static const TileShape kShapes64_9[] = {{3, 3, true, 1, false}};
static const TileShape kShapes64_mid[] = {{4, 4, true, 1, false}, {3, 3, true, 1, false}};
static const TileShape kShapes64_14[] = {{5, 4, false, 1, false}, {5, 5, false, 1, false}, {4, 4, false, 1, false},
                                         {5, 5, false, 1, true}};
static const TileShape kShapes32_14[] = {{5, 4, true, 1, false}, {5, 5, true, 1, false}, {4, 4, true, 1, false},
                                         {5, 4, false, 4, false}, {5, 5, false, 4, true}};
static const TileShape kShapes32_15[] = {{5, 4, false, 1, false}, {6, 5, false, 1, false}};

static const TileShape *tile_shapes(bool fp32, int log2n, int *n) {
    const int lognc = log2n - 1;
#define MGPU_SHAPES(a)                          \
    do {                                        \
        *n = (int) (sizeof(a) / sizeof(a[0])); \
        return a;                               \
    } while (0)
    if (lognc == 9) MGPU_SHAPES(kShapes64_9);
    if (lognc >= 10 && lognc <= 13) MGPU_SHAPES(kShapes64_mid);
    if (lognc == 14) {
        if (fp32) MGPU_SHAPES(kShapes32_14);
        MGPU_SHAPES(kShapes64_14);
    }
    if (lognc == 15 && fp32) MGPU_SHAPES(kShapes32_15);
#undef MGPU_SHAPES
    *n = 0;
    return nullptr;
}

int tile_variant_count(bool fp32, int log2n) {
    int n = 0;
    tile_shapes(fp32, log2n, &n);
    return n;
}

const char *tile_variant_desc(bool fp32, int log2n, int variant) {
    static thread_local char buf[160];
    int n = 0;
    const TileShape *sh = tile_shapes(fp32, log2n, &n);
    if (!sh || variant < 1 || variant > n) return "";
    const TileShape &v = sh[variant - 1];
    const int nc = 1 << (log2n - 1);
    std::snprintf(buf, sizeof(buf),
                  "%s N=%d: %d complex values per lane, %d lanes, radix <= %d, %s LDS exchange%s, >= %d waves/SIMD",
                  fp32 ? "fp32" : "fp64", 1 << log2n, 1 << v.logv, nc >> v.logv, 1 << v.logr,
                  v.two ? "two-plane" : "one-plane", v.wave_local ? " (inner passes inside the wave)" : "", v.minw);
    return buf;
}

// the tiles a precision's products use: the FP64 set, or the FP32 long-tile set
struct TileUse {
    int log2n, L, ntiles;
    const TileDesc *tiles;
    const void *coef, *tw;
};
// the tiles of a launch: during a segmented PCG (live), those of the running segments
template <typename R>
static TileUse tile_use(const TileClass &C, bool live) {
    if (sizeof(R) == 4 && C.log2n32 > 0 && C.fp32_long) {
        if (live && C.ntiles32_run >= 0) return {C.log2n32, C.L32, C.ntiles32_run, C.d_tiles32_run, C.d_coef32L, C.d_tw32L};
        return {C.log2n32, C.L32, C.ntiles32, C.d_tiles32, C.d_coef32L, C.d_tw32L};
    }
    const bool run = live && C.ntiles_run >= 0;
    const int n = run ? C.ntiles_run : C.ntiles;
    const TileDesc *t = run ? C.d_tiles_run : C.d_tiles;
    if (sizeof(R) == 4) return {C.log2n, C.L, n, t, C.d_coef32, C.d_tw32};
    return {C.log2n, C.L, n, t, C.d_coef64, C.d_tw64};
}

template <typename R, int LOGNC, int LOGV, int LOGR, bool TWO, int MINW>
static int launch_tiles_cfg(Problem &pb, const TileClass &C, const R *d_in, R *d_out, const void *coef,
                            const void *tw, const ToeplitzScratch &sc) {
    const TileUse u = tile_use<R>(C, pb.live_state != nullptr);
    if (u.ntiles == 0) return MGPU_OK;
    constexpr int NC = 1 << LOGNC, NT = NC / (1 << LOGV);
    constexpr size_t lds = sizeof(R) * (size_t) lds_elems(NC) * (TWO ? 2 : 1);
    static_assert(lds <= 160 * 1024, "tile does not fit the LDS");
    auto kern = k_tiles<R, LOGNC, LOGV, LOGR, TWO, MINW>;
    static std::atomic<unsigned> attr_devices{0};  // devices on which the LDS limit was raised
    if (pb.device < 32 && !(attr_devices.load(std::memory_order_acquire) & (1u << pb.device))) {
        MGPU_HIP(hipFuncSetAttribute((const void *) kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
        attr_devices.fetch_or(1u << pb.device, std::memory_order_release);
    }
    TileArgs a{d_in, d_out, u.tiles, coef, tw, sc.mean, sc.edge_part, C.lambda, u.L, pb.sg.d_block_seg,
               pb.sg.d_block_seg ? pb.live_state : nullptr};
    hipLaunchKernelGGL(kern, dim3(u.ntiles), dim3(NT), lds, pb.stream, a);
    MGPU_LAUNCHED();
    return MGPU_OK;
}

template <typename R, int MINW>
static int launch_tiles_wl(Problem &pb, const TileClass &C, const R *d_in, R *d_out, const ToeplitzScratch &sc) {
    const void *coef = sizeof(R) == 4 ? (const void *) C.d_coefW32 : (const void *) C.d_coefW64;
    const void *tw = sizeof(R) == 4 ? (const void *) C.d_tw32 : (const void *) C.d_tw64;
    MGPU_ARG(coef && C.log2n == 15 && !(sizeof(R) == 4 && C.fp32_long), "wave-local tiles need N = 32768");
    const TileUse u = tile_use<R>(C, pb.live_state != nullptr);
    if (u.ntiles == 0) return MGPU_OK;
    constexpr size_t lds = sizeof(R) * (size_t) kWlLdsElems;
    static_assert(lds <= 160 * 1024, "tile does not fit the LDS");
    auto kern = k_tiles_wl<R, MINW>;
    static std::atomic<unsigned> attr_devices{0};  // devices on which the LDS limit was raised
    if (pb.device < 32 && !(attr_devices.load(std::memory_order_acquire) & (1u << pb.device))) {
        MGPU_HIP(hipFuncSetAttribute((const void *) kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
        attr_devices.fetch_or(1u << pb.device, std::memory_order_release);
    }
    TileArgs a{d_in, d_out, u.tiles, coef, tw, sc.mean, sc.edge_part, C.lambda, C.L, pb.sg.d_block_seg,
               pb.sg.d_block_seg ? pb.live_state : nullptr};
    hipLaunchKernelGGL(kern, dim3(u.ntiles), dim3(kWlNT), lds, pb.stream, a);
    MGPU_LAUNCHED();
    return MGPU_OK;
}

template <typename R>
static int launch_tiles(Problem &pb, const TileClass &C, const R *d_in, R *d_out, int variant,
                        const ToeplitzScratch &sc) {
    constexpr bool F32 = sizeof(R) == 4;
    const TileUse u = tile_use<R>(C, pb.live_state != nullptr);
    const void *coef = u.coef, *tw = u.tw;
#define MGPU_TL(LOGNC, LOGV, LOGR, TWO, MINW) \
    return launch_tiles_cfg<R, LOGNC, LOGV, LOGR, TWO, MINW>(pb, C, d_in, d_out, coef, tw, sc)
    switch (u.log2n - 1) {  // LOGNC
        case 9: MGPU_TL(9, 3, 3, true, 1);
        case 10:
            if (variant == 2) MGPU_TL(10, 3, 3, true, 1);
            MGPU_TL(10, 4, 4, true, 1);
        case 11:
            if (variant == 2) MGPU_TL(11, 3, 3, true, 1);
            MGPU_TL(11, 4, 4, true, 1);
        case 12:
            if (variant == 2) MGPU_TL(12, 3, 3, true, 1);
            MGPU_TL(12, 4, 4, true, 1);
        case 13:
            if (variant == 2) MGPU_TL(13, 3, 3, true, 1);
            MGPU_TL(13, 4, 4, true, 1);
        case 14:
            if constexpr (F32) {
                if (variant == 2) MGPU_TL(14, 5, 5, true, 1);
                if (variant == 3) MGPU_TL(14, 4, 4, true, 1);
                if (variant == 4) MGPU_TL(14, 5, 4, false, 4);
                if (variant == 5) return launch_tiles_wl<R, 4>(pb, C, d_in, d_out, sc);
                MGPU_TL(14, 5, 4, true, 1);
            } else {
                if (variant == 2) MGPU_TL(14, 5, 5, false, 1);
                if (variant == 3) MGPU_TL(14, 4, 4, false, 1);
                if (variant == 4) return launch_tiles_wl<R, 1>(pb, C, d_in, d_out, sc);
                MGPU_TL(14, 5, 4, false, 1);
            }
        case 15:
            if constexpr (F32) {
                if (variant == 2) MGPU_TL(15, 6, 5, false, 1);
                MGPU_TL(15, 5, 4, false, 1);
            }
            set_error("unsupported FP64 tile length 2^16");
            return MGPU_EINVAL;
        default: set_error("unsupported tile length 2^%d", u.log2n); return MGPU_EINVAL;
    }
#undef MGPU_TL
}

// Fixed tile variant of a (precision, tile length): the one the products use
// unless the caller forces another or opts into autotuning.  FP64 results
// therefore depend only on the inputs (variants differ in pass structure and
// round differently).  The entries are the fastest variants measured on
// MI355X for the lengths the workloads use (the variant tables above).
static int default_variant(bool fp32, int log2n) {
    const int lognc = log2n - 1;
    if (!fp32 && lognc == 14) return 2;
    if (fp32 && lognc == 14) return 5;
    return 1;
}

// Process-wide autotuning results (opt-in, tile_variant = MGPU_TILE_AUTOTUNE):
// (device, precision, tile lengths) -> variant.  The mutex is held while a
// key is being timed, so concurrent first products of one key wait for the
// result instead of timing under each other's load.
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int>, int> g_tune;

// Choose the tile variant of precision R for a tile class: the forced one, the
// fixed default, or (MGPU_TILE_AUTOTUNE) the cached result or the fastest of
// 1 warm-up + 3 timed launches of each variant on the caller's buffers (d_out
// is overwritten by every launch; the caller launches the winner afterwards).
template <typename R>
static int pick_variant(Problem &pb, TileClass &T, const R *d_in, R *d_out, const ToeplitzScratch &sc) {
    const int slot = sizeof(R) == 4 ? 1 : 0;
    if (T.variant[slot]) return MGPU_OK;
    // FP32 products use the FP64 tile length unless autotuning picks the long set
    if (slot == 1) T.fp32_long = false;
    const int log2n = tile_use<R>(T, false).log2n;
    const int nv = tile_variant_count(slot == 1, log2n);
    MGPU_ARG(nv > 0, "unsupported tile length 2^%d", log2n);
    // a forced variant applies to the precision of the PCG iterations (FP32 under
    // MGPU_MIXED); the FP64 products of a mixed solve (right-hand side, residual
    // refreshes) use the fixed FP64 default
    const int main_slot = pb.opt.precision == MGPU_MIXED ? 1 : 0;
    if (pb.opt.tile_variant > 0 && slot == main_slot) {
        MGPU_ARG(pb.opt.tile_variant <= nv, "tile_variant=%d: the %s tile of length 2^%d has %d variant(s)",
                 pb.opt.tile_variant, slot ? "fp32" : "fp64", log2n, nv);
        T.variant[slot] = pb.opt.tile_variant;
        return MGPU_OK;
    }
    if (pb.opt.tile_variant != MGPU_TILE_AUTOTUNE || slot != main_slot) {
        MGPU_ARG(pb.opt.tile_variant >= 0 || pb.opt.tile_variant == MGPU_TILE_AUTOTUNE,
                 "tile_variant=%d: use 0 (fixed default), k > 0 or MGPU_TILE_AUTOTUNE", pb.opt.tile_variant);
        T.variant[slot] = std::min(default_variant(slot == 1, log2n), nv);
        return MGPU_OK;
    }
    // autotuning; FP32 with a long-tile set: the candidates are every variant of both lengths
    const bool two_len = slot == 1 && T.log2n32 > 0;
    const int nv_long = two_len ? tile_variant_count(true, T.log2n32) : 0;
    if (nv == 1 && !two_len) {
        T.variant[slot] = 1;
        return MGPU_OK;
    }
    const auto key = std::make_tuple(pb.device, slot, log2n * 100 + (two_len ? T.log2n32 : 0));
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
        T.variant[slot] = it->second % 1000;
        if (two_len) T.fp32_long = it->second >= 1000;
        return MGPU_OK;
    }
    hipEvent_t e0, e1;
    MGPU_HIP(hipEventCreate(&e0));
    MGPU_HIP(hipEventCreate(&e1));
    int best = 1;
    bool best_long = false;
    float best_ms = 0.f;
    int rc = MGPU_OK;
    for (int c = 1; c <= nv + nv_long && rc == MGPU_OK; c++) {
        const bool lng = c > nv;
        const int v = lng ? c - nv : c;
        if (two_len) T.fp32_long = lng;
        rc = launch_tiles<R>(pb, T, d_in, d_out, v, sc);
        if (rc != MGPU_OK) break;
        if (hipEventRecord(e0, pb.stream) != hipSuccess) {
            set_error("tile autotuning: hipEventRecord failed");
            rc = MGPU_EHIP;
            break;
        }
        for (int k = 0; k < 3 && rc == MGPU_OK; k++) rc = launch_tiles<R>(pb, T, d_in, d_out, v, sc);
        // per useful output: the two lengths have different tile counts but the same outputs
        if (rc != MGPU_OK) break;
        float ms = 0.f;
        if (hipEventRecord(e1, pb.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            set_error("tile autotuning: event timing failed");
            rc = MGPU_EHIP;
            break;
        }
        if (c == 1 || ms < best_ms) {
            best_ms = ms;
            best = v;
            best_long = lng;
        }
    }
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    if (rc != MGPU_OK) return rc;
    T.variant[slot] = best;
    if (two_len) T.fp32_long = best_long;
    g_tune[key] = best + (best_long ? 1000 : 0);
    return MGPU_OK;
}

// scratch of one product: per-chunk partial sums, block means, slice sums of
// the symmetric mode's output mean (the plan's own buffers, or a caller's)
template <typename R>
static int toeplitz_apply_t(Problem &pb, const R *d_in, R *d_out, bool compute_partials,
                            const ToeplitzScratch &sc) {
    ToeplitzPlan &T = pb.tp;
    hipStream_t s = pb.stream;
    if (T.lambda == 1) {
        if (pb.stbmm_offsets && T.offset_mode != MGPU_OFFSET_NONE) {
            // stbmmProd with lambda 1 (the two-level products): block means removed
            // (with lambda 1, C T C = T C, so the symmetric mode is the same product)
            k_partials<R><<<(unsigned) T.nchunks, 256, 0, s>>>(d_in, T.m, T.d_start, T.d_chunk_block, T.nb,
                                                               sc.partials);
            MGPU_LAUNCHED();
            k_block_mean<R><<<dim3(T.nb, 1), 256, 0, s>>>(sc.partials, T.d_start, T.nb, sc.mean, d_in, nullptr,
                                                          nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
            MGPU_LAUNCHED();
            k_diag_mean<R><<<grid_for(T.m, 256), 256, 0, s>>>(d_in, d_out, T.m, T.d_start, T.nb, T.d_T0, sc.mean);
        } else {
            k_diag<R><<<grid_for(T.m, 256), 256, 0, s>>>(d_in, d_out, T.m, T.d_start, T.nb, T.d_T0);
        }
        MGPU_LAUNCHED();
        return MGPU_OK;
    }
    if (T.offset_mode != MGPU_OFFSET_NONE) {  // MGPU_OFFSET_NONE: means stay 0
        pb.prof.begin(MGPU_KT_MEAN, s);
        if (compute_partials) {
            k_partials<R><<<(unsigned) T.nchunks, 256, 0, s>>>(d_in, T.m, T.d_start, T.d_chunk_block, T.nb,
                                                               sc.partials);
            MGPU_LAUNCHED();
        }
        const bool sym = T.offset_mode == MGPU_OFFSET_SYMMETRIC;
        k_block_mean<R><<<dim3(T.nb, sym ? kEdgeSlices : 1), 256, 0, s>>>(
                sc.partials, T.d_start, T.nb, sc.mean, d_in, T.d_tail, T.d_tail_off, T.d_block_filter, T.d_block_kind,
                T.d_block_lambda,
                sym ? sc.edge_part : nullptr, pb.sg.d_block_seg, pb.sg.d_block_seg ? pb.live_state : nullptr);
        MGPU_LAUNCHED();
        pb.prof.end(s);
    }
    // one launch per tile class (band and tile length), in one profiled span
    for (TileClass &C : T.cls) MGPU_RET(pick_variant<R>(pb, C, d_in, d_out, sc));
    if (!T.cls.empty()) {
        pb.prof.begin(MGPU_KT_TILES, s);
        for (TileClass &C : T.cls) MGPU_RET(launch_tiles<R>(pb, C, d_in, d_out, C.variant[sizeof(R) == 4 ? 1 : 0], sc));
        pb.prof.end(s);
    }
    if (!T.h_direct_blocks.empty()) {
        k_direct<R><<<dim3(64, (unsigned) T.h_direct_blocks.size()), 256, 0, s>>>(
                d_in, d_out, T.d_start, T.d_direct_blocks, sc.mean, sc.edge_part, T.d_direct_invtt, T.d_inv_off,
                T.d_block_lambda);
        MGPU_LAUNCHED();
    }
    if (!T.h_short_blocks.empty()) {
        k_mean_only<R><<<dim3(16, (unsigned) T.h_short_blocks.size()), 256, 0, s>>>(d_in, d_out, T.d_start,
                                                                                  T.d_short_blocks, sc.mean);
        MGPU_LAUNCHED();
    }
    if (T.covered < T.m) {
        k_copy_tail<R><<<grid_for(T.m - T.covered, 256), 256, 0, s>>>(d_in, d_out, T.covered, T.m);
        MGPU_LAUNCHED();
    }
    return MGPU_OK;
}

static ToeplitzScratch plan_scratch(const ToeplitzPlan &T) { return {T.d_partials, T.d_mean, T.d_edge_part}; }

int toeplitz_apply(Problem &pb, const double *d_in, double *d_out, bool compute_partials) {
    return toeplitz_apply_t<double>(pb, d_in, d_out, compute_partials, plan_scratch(pb.tp));
}

int toeplitz_apply_f32(Problem &pb, const float *d_in, float *d_out, bool compute_partials) {
    return toeplitz_apply_t<float>(pb, d_in, d_out, compute_partials, plan_scratch(pb.tp));
}

size_t toeplitz_scratch_doubles(const ToeplitzPlan &T) {
    return (size_t) (T.nchunks + T.nb) + (size_t) T.nb + (size_t) T.nb * kEdgeSlices;
}

__global__ void k_fill_zero(double *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// N^-1 without any allocation (mappraiser_gpu_apply_weights_noalloc)
// ---------------------------------------------------------------------------
// The product's scratch (per-chunk partial sums, block means, output-mean
// slices: toeplitz_scratch_doubles(T) doubles) lives in this file-scope device
// array, one copy per GPU, held until the process ends.  Host code never
// passes the array's name to a kernel: it uses the address that
// hipGetSymbolAddress returns for the current device, kept per device.  Every
// call on a GPU shares that GPU's copy; calls that overlap on one GPU race.
constexpr size_t kNoallocScratchDoubles = (size_t) 1 << 20;  // 8 MiB
static __device__ __align__(256) double scratch[kNoallocScratchDoubles];
constexpr int kNoallocDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kNoallocDevices] = {};

int noalloc_scratch_init(int device) {
    MGPU_ARG(device >= 0 && device < kNoallocDevices, "device %d out of range of the scratch table", device);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device]) return MGPU_OK;
    int cur = -1;
    MGPU_HIP(hipGetDevice(&cur));
    MGPU_ARG(cur == device, "scratch set-up for device %d while device %d is current", device, cur);
    void *p = nullptr;
    MGPU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MGPU_ARG(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = static_cast<double *>(p);
    return MGPU_OK;
}

int toeplitz_apply_noalloc(Problem &pb, const double *d_in, double *d_out, hipStream_t stream) {
    ToeplitzPlan &T = pb.tp;
    const size_t need = toeplitz_scratch_doubles(T);
    MGPU_ARG(need <= kNoallocScratchDoubles,
             "apply_weights_noalloc: this problem needs %zu scratch doubles (%lld samples, %d blocks), the "
             "library's array holds %zu",
             need, (long long) T.m, T.nb, kNoallocScratchDoubles);
    MGPU_RET(noalloc_scratch_init(pb.device));
    double *sc = g_scratch_ptr[pb.device];
    // the tile variant comes from the handle's choice or the fixed default (no
    // autotuning here: it would time launches with events)
    const int keep_variant_opt = pb.opt.tile_variant;
    if (pb.opt.tile_variant == MGPU_TILE_AUTOTUNE)
        for (const TileClass &C : T.cls)
            if (C.variant[0] == 0) pb.opt.tile_variant = 0;
    const bool keep_prof = pb.prof.on;  // no profiling events either
    pb.prof.on = false;
    hipStream_t keep_stream = pb.stream;
    if (stream) pb.stream = stream;
    const int rc = toeplitz_apply_scratch(pb, d_in, d_out, sc);
    pb.stream = keep_stream;
    pb.prof.on = keep_prof;
    pb.opt.tile_variant = keep_variant_opt;
    return rc;
}

int toeplitz_apply_scratch(Problem &pb, const double *d_in, double *d_out, double *d_scratch) {
    const ToeplitzPlan &T = pb.tp;
    ToeplitzScratch sc;
    sc.partials = d_scratch;
    sc.mean = sc.partials + (T.nchunks + T.nb);
    sc.edge_part = sc.mean + T.nb;
    // means and output-mean slices that the product reads but no kernel of this
    // mode writes must be zero (the array keeps the previous call's values):
    // cleared with a kernel so that a captured graph holds no memset node
    const int64_t nz = (int64_t) T.nb * (1 + kEdgeSlices);
    if (T.lambda > 1 && T.offset_mode != MGPU_OFFSET_SYMMETRIC) {
        k_fill_zero<<<grid_for(nz, 256), 256, 0, pb.stream>>>(sc.mean, nz);
        MGPU_LAUNCHED();
    }
    return toeplitz_apply_t<double>(pb, d_in, d_out, true, sc);
}

}  // namespace mgpu
