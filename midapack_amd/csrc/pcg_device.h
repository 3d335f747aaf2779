// pcg_device.h -- device-side helpers shared by the PCG kernels (pcg.hip)
// and the two-level preconditioner (twolvl.hip): per-segment scalar slots,
// stop flags, workgroup reductions in a fixed order and the map-domain
// workgroup -> (segment, pixel) mapping.
#pragma once
#include "mgpu_internal.h"

namespace mgpu {

// scalar slots of one segment in d_scal (segment r at r * S_NSCAL)
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
    S_TRUE_R2,  // ||b - A x||^2 of the returned x (mixed precision)
    S_TRUE_B2,  // ||b||^2
    S_PEND,     // mixed precision: 1 while an FP32-updated residual of this segment
                // met the stop test (or, with adaptive refreshes, drifted far enough
                // from its last FP64 value) and awaits its FP64 refresh
    S_RREF,     // adaptive refreshes: ||g||^2 of the last FP64 residual
    S_KREF,     // adaptive refreshes: iteration of the last FP64 residual
    S_NSCAL = 18
};

// Adaptive refreshes (MGPU_REFRESH_ADAPTIVE): a segment asks for an FP64
// residual once its FP32-updated ||g||^2 has fallen below kAdaptDrop times the
// last FP64 one (the norm by 10x), or kAdaptGap iterations after it.  The
// recurrence's drift grows with the largest residual since the last
// replacement, so replacing it at every decade keeps the drift a small
// fraction of the current residual (reliable updates, Sleijpen and van der
// Vorst 1996).
constexpr double kAdaptDrop = 1e-2;
constexpr double kAdaptGap = 64.0;

// device flags: [0] every segment stopped, [1] segment 0 converged, [2] largest
// iteration count reached, [3] number of stopped segments, [4] never set,
// [5] number of segments whose FP32-updated residual met the stop test
// (mixed precision: S_PEND set; the first iteration of the next block
// refreshes their residual in FP64 and tests it), [6] FP64 residual refreshes
// performed
enum { F_STOP = 0, F_CONV0, F_KMAX, F_NDONE, F_NEVER = 4, F_NPEND = 5, F_NREF = 6, F_NFLAGS = 8 };

// iteration kinds of k_scalars / k_refresh: FP64 solve; mixed precision without
// a refresh; every live segment refreshed in FP64 (every refresh_every
// iterations); only the segments with S_PEND set refreshed
enum { IT_FP64 = 0, IT_MIXED = 1, IT_REFRESH_ALL = 2, IT_REFRESH_PENDING = 3 };

constexpr int kMB = kMapWG;  // map-domain workgroup size (pixels)

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

// threads of the scalar reductions (one workgroup per segment)
constexpr int kRB = 1024;

// fixed-order reduction of nparts partials by one workgroup of kRB threads
// (red: kRB / 64 doubles of LDS); every thread returns the sum
__device__ __forceinline__ double reduce_parts(const double *__restrict__ p, int nparts, double *red) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kRB) s += p[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __syncthreads();  // a previous call may still read red
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kRB / 64; w++) t += red[w];
    return t;
}

// pixel handled by this thread of a map-domain workgroup, its segment and the segment end
struct SegPix {
    int r, q, qend;
};
__device__ __forceinline__ SegPix seg_pixel(const int *__restrict__ wg_seg, const int *__restrict__ wg_q0,
                                            const int *__restrict__ seg_q) {
    SegPix p;
    p.r = wg_seg[blockIdx.x];
    p.q = wg_q0[blockIdx.x] + (int) threadIdx.x;
    p.qend = seg_q[p.r + 1];
    return p;
}

}  // namespace mgpu
