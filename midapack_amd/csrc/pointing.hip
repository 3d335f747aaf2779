// pointing.hip -- local index tables and the pointing products P, P^T (gfx950).
//
// Reference semantics (one rank):
//   MatLocalShape  mapmat.c:496-521 with counting_sort (csort.c:111-135) and
//                  dichotomy (cindex.c:102-115): lindices = sorted unique pix,
//                  indices = position in lindices, trash flag if lindices[0]<0.
//   MatVecProd     mapmat.c:704-729: y[t] = ((0 + w0 x0) + w1 x1) + w2 x2.
//   TrMatVecProd   mapmat.c:819-851: x[i] += w y[t] in increasing t.
//
// Layout choices for HBM:
//   * P runs sample-major (one lane per sample, 4 samples per lane strided by
//     the workgroup size, so every load instruction reads a contiguous span):
//     q (4 B), the interleaved (Q, U) weights (one 16-byte load) and the y
//     store are fully coalesced; the map gather x[3q+k] is L2-resident
//     (map << TOD).  The per-block sums needed by the mean removal of N^-1 are
//     fused into the epilogue (one partial per (chunk, block) pair, fixed-order
//     tree, deterministic).
//   * P^T runs pixel-major over a sliced-ELL layout with slice height 64 (one
//     wavefront): slice s stores hit i of lane l at base_s + 64 i + l, hits of a
//     pixel in increasing t.  Each lane accumulates its pixel sequentially, so
//     the summation order equals the reference's, the result is bit-exact and
//     deterministic, and no atomics are used.  All loads of the ELL arrays are
//     coalesced; the gather of y[t] is coalesced for scans whose neighbouring
//     pixels are neighbours in time (rasters).
//   * The I weight is stored only when some sample has w0 != 1 (the usual
//     pointing has w0 = 1 everywhere; x0 * 1 == x0 exactly, so skipping it is
//     bit-exact).
//
// Every kernel is compiled with -ffp-contract=off: products and sums are
// separately rounded like the reference's C build.
#include <algorithm>
#include <hipcub/hipcub.hpp>

#include "mgpu_internal.h"

namespace mgpu {

constexpr int kSlice = 64;          // ELL slice height (one wavefront)
constexpr int kPBlock = 256;        // P workgroup size (kChunk / 4 samples per lane)
constexpr int kChunkGroup = 32;     // consecutive P chunks kept together in the chunk order
constexpr int kClassifyGrid = 2048; // grid-stride validation: 8 workgroups per CU

// ---------------------------------------------------------------------------
// validation + pixel extraction
// ---------------------------------------------------------------------------
// status bits: 1 = malformed row, 2 = trash row present, 4 = I weight != 1.
// Grid-stride over the samples, one reduction per workgroup: the three
// device-wide atomics are issued once per workgroup instead of once per wave.
__global__ __launch_bounds__(256) void k_classify(const int *__restrict__ pix, const double *__restrict__ w,
                                                  int64_t m, int *__restrict__ p_of_t, int *__restrict__ status,
                                                  int *__restrict__ pmin, int *__restrict__ pmax,
                                                  int *__restrict__ first_bad) {
    __shared__ int slo[4], shi[4], sst[4];
    int lo = INT32_MAX, hi = INT32_MIN, st = 0;
    const int64_t stride = (int64_t) gridDim.x * blockDim.x;
    for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < m; t += stride) {
        const int a = pix[3 * t], b = pix[3 * t + 1], c = pix[3 * t + 2];
        int p = -1;
        bool bad = false;
        if (a >= 0) {
            if (a % 3 != 0 || b != a + 1 || c != a + 2)
                bad = true;
            else
                p = a / 3;
        } else {
            if (a != -3 || b != -2 || c != -1)
                bad = true;
            else
                st |= 2;
        }
        if (bad) {
            st |= 1;
            atomicMin(first_bad, (int) t);  // m < 2^31
        }
        if (w[3 * t] != 1.0) st |= 4;
        p_of_t[t] = p;
        if (p >= 0) {
            lo = min(lo, p);
            hi = max(hi, p);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o));
        hi = max(hi, __shfl_xor(hi, o));
        st |= __shfl_xor(st, o);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        slo[wave] = lo;
        shi[wave] = hi;
        sst[wave] = st;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int) (blockDim.x >> 6); i++) {
            lo = min(lo, slo[i]);
            hi = max(hi, shi[i]);
            st |= sst[i];
        }
        if (lo != INT32_MAX) {
            atomicMin(pmin, lo);
            atomicMax(pmax, hi);
        }
        if (st) atomicOr(status, st);
    }
}

__global__ void k_mark(const int *__restrict__ p_of_t, int64_t m, int pmin, unsigned *__restrict__ flags) {
    int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) {
        int p = p_of_t[t];
        if (p >= 0) flags[p - pmin] = 1u;
    }
}

// w: the caller's weights (3 per sample), or nullptr when the Q/U weights were
// uploaded packed (they are then P's w12 already and w12 is not written)
__global__ void k_assign(const int *__restrict__ p_of_t, int64_t m, int pmin, const unsigned *__restrict__ rank,
                         const double *__restrict__ w, int *__restrict__ q, double *__restrict__ w0,
                         double2 *__restrict__ w12, unsigned *__restrict__ key, int *__restrict__ tval,
                         unsigned trash_key) {
    int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) {
        int p = p_of_t[t];
        int qq = p >= 0 ? (int) rank[p - pmin] : -1;
        q[t] = qq;
        if (w0) w0[t] = w[3 * t];
        if (w) w12[t] = make_double2(w[3 * t + 1], w[3 * t + 2]);
        key[t] = qq >= 0 ? (unsigned) qq : trash_key;
        tval[t] = (int) t;
    }
}

__global__ void k_pixg(const unsigned *__restrict__ flags, const unsigned *__restrict__ rank, int64_t span, int pmin,
                       int *__restrict__ pixg) {
    int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < span && flags[i]) pixg[rank[i]] = pmin + (int) i;
}

// CSR offsets from the sorted keys: off[k] = first position of key k
__global__ void k_offsets(const unsigned *__restrict__ skey, int64_t nh, int npix, int *__restrict__ off) {
    int64_t h = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (h < nh) {
        unsigned k = skey[h];
        if (h == 0 || skey[h - 1] != k) off[k] = (int) h;
    }
    if (h == 0) off[npix] = (int) nh;
}

// slice widths: width[s] = max hits over the 64 pixels of slice s
__global__ void k_slice_width(const int *__restrict__ off, int npix, int nslice, int64_t *__restrict__ width) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nslice) {
        int wmax = 0;
        for (int l = 0; l < kSlice; l++) {
            int qq = s * kSlice + l;
            if (qq < npix) wmax = max(wmax, off[qq + 1] - off[qq]);
        }
        width[s] = (int64_t) wmax;
    }
}

// Scatter the CSR into sliced-ELL (pixel q = 64 s + lane; hit i at base_s + 64 i + lane)
template <bool W0>
__global__ void k_fill_ell(const int *__restrict__ off, const int *__restrict__ st, const int64_t *__restrict__ base,
                           int npix, int nslice, const double *__restrict__ w0, const double2 *__restrict__ w12,
                           int *__restrict__ ct, double *__restrict__ cw0, double2 *__restrict__ cw12) {
    int qq = blockIdx.x * blockDim.x + threadIdx.x;
    int s = qq / kSlice, lane = qq % kSlice;
    if (s >= nslice) return;
    int64_t b0 = base[s], width = (base[s + 1] - base[s]) / kSlice;
    int lo = 0, hi = 0;
    if (qq < npix) {
        lo = off[qq];
        hi = off[qq + 1];
    }
    for (int64_t i = 0; i < width; i++) {
        int64_t dst = b0 + i * kSlice + lane;
        int h = lo + (int) i;
        if (h < hi) {
            int t = st[h];
            ct[dst] = t;
            if (W0) cw0[dst] = w0[t];
            cw12[dst] = w12[t];
        } else {
            ct[dst] = -1;
            if (W0) cw0[dst] = 0.0;
            cw12[dst] = make_double2(0.0, 0.0);
        }
    }
}

__global__ void k_scale64(int64_t *a, int64_t n, int64_t f) {
    int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] *= f;
}

// number of valid samples: keys < npix are sorted before the trash keys
__global__ void k_count_valid(const unsigned *__restrict__ skey, int64_t m, unsigned npix, int64_t *__restrict__ cnt) {
    int64_t h = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (h < m && skey[h] < npix && (h + 1 == m || skey[h + 1] >= npix)) *cnt = h + 1;
}

__global__ void k_to_f32(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
    int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float) a[i];
}

__global__ void k_to_f32x2(const double2 *__restrict__ a, float2 *__restrict__ b, int64_t n) {
    int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double2 v = a[i];
        b[i] = make_float2((float) v.x, (float) v.y);
    }
}

// compaction after the pixel cut
__global__ void k_kept_counts(const unsigned *__restrict__ kept, const int *__restrict__ off, int npix,
                              unsigned *__restrict__ keep_u, int *__restrict__ cnt) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < npix) {
        keep_u[q] = kept[q] ? 1u : 0u;
        cnt[q] = kept[q] ? off[q + 1] - off[q] : 0;
    } else if (q == npix) {
        keep_u[q] = 0u;
        cnt[q] = 0;
    }
}

__global__ void k_compact_pixels(const unsigned *__restrict__ kept, const unsigned *__restrict__ newidx,
                                 const int *__restrict__ hscan, const int *__restrict__ pixg_old, int npix,
                                 int *__restrict__ pixg_new, int *__restrict__ off_new, int npix_new) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < npix && kept[q]) {
        pixg_new[newidx[q]] = pixg_old[q];
        off_new[newidx[q]] = hscan[q];
    }
    if (q == npix) off_new[npix_new] = hscan[npix];
}

__global__ void k_compact_hits(const unsigned *__restrict__ kept, const unsigned *__restrict__ newidx,
                               const int *__restrict__ off_old, const int *__restrict__ cst_old,
                               const int *__restrict__ off_new, int npix, int *__restrict__ cst_new) {
    int q = blockIdx.x;
    if (q >= npix || !kept[q]) return;
    const int lo = off_old[q], hi = off_old[q + 1], dst = off_new[newidx[q]];
    for (int h = lo + threadIdx.x; h < hi; h += blockDim.x) cst_new[dst + (h - lo)] = cst_old[h];
}

__global__ void k_remap_q(int *__restrict__ q, int64_t m, const unsigned *__restrict__ kept,
                          const unsigned *__restrict__ newidx) {
    int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) {
        int qq = q[t];
        if (qq >= 0) q[t] = kept[qq] ? (int) newidx[qq] : -1;
    }
}

// ---------------------------------------------------------------------------
// build
// ---------------------------------------------------------------------------
static void release_fp32(Problem &pb) {
    Pointing &P = pb.pt;
    for (void *p : {(void *) P.d_cw12f, (void *) P.d_cw0f, (void *) P.d_w12f, (void *) P.d_w0f})
        if (p) pb.arena.release(p);
    P.d_cw12f = nullptr;
    P.d_cw0f = nullptr;
    P.d_w12f = nullptr;
    P.d_w0f = nullptr;
}

// ---------------------------------------------------------------------------
// P's workgroup order.  Every chunk's outputs and block sums are the same in
// any order; the order only decides where the map gathers hit.  Chunks are
// sorted by the smallest pixel they touch and dealt to the XCDs in contiguous
// runs (workgroup b runs on XCD b % 8), so chunks that gather the same map
// rows (other detectors' passes over them) run on one XCD close together in
// time and find those rows in that XCD's L2 instead of the MALL.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_chunk_key(const int *__restrict__ q, int64_t m, int *__restrict__ key) {
    __shared__ int red[4];
    const int64_t c0 = (int64_t) blockIdx.x * kChunk;
    int v = INT32_MAX;
    for (int i = threadIdx.x; i < kChunk; i += 256) {
        const int64_t t = c0 + i;
        if (t < m) {
            const int qq = q[t];
            if (qq >= 0) v = min(v, qq);
        }
    }
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) key[blockIdx.x] = min(min(red[0], red[1]), min(red[2], red[3]));
}

int pointing_chunk_keys(Problem &pb, std::vector<int> &key) {
    Pointing &P = pb.pt;
    hipStream_t s = pb.stream;
    const int nch = (int) ((P.m + kChunk - 1) / kChunk);
    int *d_key;
    MGPU_RET(pb.arena.alloc(&d_key, (size_t) nch));
    k_chunk_key<<<nch, 256, 0, s>>>(P.d_q, P.m, d_key);
    MGPU_LAUNCHED();
    key.assign((size_t) nch, 0);
    MGPU_HIP(hipMemcpyAsync(key.data(), d_key, sizeof(int) * nch, hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipStreamSynchronize(s));
    pb.arena.release(d_key);
    return MGPU_OK;
}

static int build_chunk_order(Problem &pb) {
    Pointing &P = pb.pt;
    hipStream_t s = pb.stream;
    const int nch = (int) ((P.m + kChunk - 1) / kChunk);
    std::vector<int> key, order((size_t) nch), perm((size_t) nch);
    MGPU_RET(pointing_chunk_keys(pb, key));
    // Groups of kChunkGroup consecutive chunks, sorted by their smallest pixel:
    // the map rows a group gathers stay in L2 while the detectors' passes over
    // them run back to back, and each group streams its pointing from HBM as
    // one contiguous run.  On the 1024-sample sweeps of the 64 x 2^20 raster,
    // P takes 0.388 ms with single chunks, 0.361 ms with groups of 16, 0.353 ms
    // with 32 or 64, 0.417 ms with 128 (the gathers start to miss L2).
    {
        const int ng = (nch + kChunkGroup - 1) / kChunkGroup;
        std::vector<int> gkey((size_t) ng, INT32_MAX), gord((size_t) ng);
        for (int c = 0; c < nch; c++) gkey[c / kChunkGroup] = std::min(gkey[c / kChunkGroup], key[c]);
        for (int g = 0; g < ng; g++) gord[g] = g;
        std::stable_sort(gord.begin(), gord.end(), [&](int a, int b) { return gkey[a] < gkey[b]; });
        int k = 0;
        for (int g : gord)
            for (int c = g * kChunkGroup; c < std::min(nch, (g + 1) * kChunkGroup); c++) order[k++] = c;
    }
    // each XCD gets a contiguous run of the sorted chunks
    perm = deal_to_xcds(order);
    P.h_chunk_order = order;
    if (!P.d_chunk_perm) MGPU_RET(pb.arena.alloc(&P.d_chunk_perm, (size_t) nch));
    MGPU_HIP(hipMemcpyAsync(P.d_chunk_perm, perm.data(), sizeof(int) * nch, hipMemcpyHostToDevice, s));
    MGPU_HIP(hipStreamSynchronize(s));
    return MGPU_OK;
}

static int build_ell(Problem &pb) {
    Pointing &P = pb.pt;
    hipStream_t s = pb.stream;
    const int npix = P.npix;
    const int nslice = (npix + kSlice - 1) / kSlice;
    for (void *p : {(void *) P.d_ell_base, (void *) P.d_ct, (void *) P.d_cw0, (void *) P.d_cw12})
        if (p) pb.arena.release(p);
    P.d_ell_base = nullptr;
    P.d_ct = nullptr;
    P.d_cw0 = nullptr;
    P.d_cw12 = nullptr;
    int64_t *d_w, *d_base;
    MGPU_RET(pb.arena.alloc(&d_w, (size_t) nslice + 1));
    MGPU_RET(pb.arena.alloc(&d_base, (size_t) nslice + 1));
    MGPU_HIP(hipMemsetAsync(d_w, 0, sizeof(int64_t) * (nslice + 1), s));
    if (nslice > 0) {
        k_slice_width<<<grid_for(nslice, 256), 256, 0, s>>>(P.d_off, npix, nslice, d_w);
        MGPU_LAUNCHED();
    }
    size_t tb = 0;
    MGPU_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, d_w, d_base, nslice + 1, s));
    MGPU_RET(pb.cub_tmp.ensure(tb));
    MGPU_HIP(hipcub::DeviceScan::ExclusiveSum(pb.cub_tmp.p, tb, d_w, d_base, nslice + 1, s));
    k_scale64<<<grid_for(nslice + 1, 256), 256, 0, s>>>(d_base, nslice + 1, kSlice);
    MGPU_LAUNCHED();
    int64_t nell = 0;
    MGPU_HIP(hipMemcpyAsync(&nell, d_base + nslice, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipStreamSynchronize(s));
    pb.arena.release(d_w);
    MGPU_RET(pb.arena.alloc(&P.d_ct, (size_t) nell));
    if (!P.unit_w0) MGPU_RET(pb.arena.alloc(&P.d_cw0, (size_t) nell));
    MGPU_RET(pb.arena.alloc(&P.d_cw12, (size_t) nell));
    P.nh = nell;
    P.d_ell_base = d_base;
    const int nthr = nslice * kSlice;
    if (nthr > 0) {
        if (P.unit_w0)
            k_fill_ell<false><<<grid_for(nthr, 256), 256, 0, s>>>(P.d_off, P.d_cst, d_base, npix, nslice, nullptr,
                                                                  P.d_w12, P.d_ct, nullptr, P.d_cw12);
        else
            k_fill_ell<true><<<grid_for(nthr, 256), 256, 0, s>>>(P.d_off, P.d_cst, d_base, npix, nslice, P.d_w0,
                                                                 P.d_w12, P.d_ct, P.d_cw0, P.d_cw12);
        MGPU_LAUNCHED();
    }
    release_fp32(pb);  // the FP32 copies are stale now
    return build_chunk_order(pb);
}

int pointing_build(Problem &pb, const int *d_pix, const double *d_wts, int64_t m, int nnz) {
    Pointing &P = pb.pt;
    MGPU_ARG(nnz == 3, "nnz=%d: the GPU path supports the I,Q,U pointing (nnz=3) only", nnz);
    MGPU_ARG(m > 0 && m < (int64_t) INT32_MAX, "m=%lld out of range", (long long) m);
    P.m = m;
    P.nnz = nnz;
    hipStream_t s = pb.stream;

    int *d_p, *d_stat;
    MGPU_RET(pb.arena.alloc(&d_p, (size_t) m));
    MGPU_RET(pb.arena.alloc(&d_stat, 4));
    int init[4] = {0, INT32_MAX, INT32_MIN, INT32_MAX};
    MGPU_HIP(hipMemcpyAsync(d_stat, init, sizeof(init), hipMemcpyHostToDevice, s));
    const unsigned gcl = (unsigned) std::min<int64_t>(kClassifyGrid, (m + 255) / 256);
    k_classify<<<gcl, 256, 0, s>>>(d_pix, d_wts, m, d_p, d_stat, d_stat + 1, d_stat + 2, d_stat + 3);
    MGPU_LAUNCHED();
    int hst[4];
    MGPU_HIP(hipMemcpyAsync(hst, d_stat, sizeof(hst), hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipStreamSynchronize(s));
    pb.arena.release(d_stat);
    if (hst[0] & 1) {
        int bad[3] = {0, 0, 0};
        MGPU_HIP(hipMemcpy(bad, d_pix + 3 * (int64_t) hst[3], sizeof(bad), hipMemcpyDeviceToHost));
        set_error("pointing is not encoded 3p+k (mappraiser.py:1214-1216) with flagged samples as -3,-2,-1 "
                  "(utils.py:148-150): sample %d has (%d, %d, %d)", hst[3], bad[0], bad[1], bad[2]);
        return MGPU_EINVAL;
    }
    MGPU_ARG(hst[1] != INT32_MAX, "no valid sample: every sample is flagged");
    P.unit_w0 = !(hst[0] & 4);
    MGPU_RET(pointing_build_from_p(pb, d_p, hst[1], hst[2], (hst[0] & 2) != 0, d_wts, nullptr));
    pb.arena.release(d_p);
    return MGPU_OK;
}

int pointing_build_packed(Problem &pb, const int *d_p, double2 *d_w12, int64_t m, int pmin, int pmax,
                          bool any_flagged) {
    Pointing &P = pb.pt;
    MGPU_ARG(m > 0 && m < (int64_t) INT32_MAX, "m=%lld out of range", (long long) m);
    MGPU_ARG(pmin <= pmax, "no valid sample: every sample is flagged");
    P.m = m;
    P.nnz = 3;
    P.unit_w0 = true;
    return pointing_build_from_p(pb, d_p, pmin, pmax, any_flagged, nullptr, d_w12);
}

// the pointing tables from each sample's global pixel p (-1: flagged): local
// pixel ids, Q/U weights (from the caller's weights d_wts, or adopted from the
// packed upload d_w12), the sample order by pixel and the sliced ELL of P^T
int pointing_build_from_p(Problem &pb, const int *d_p, int pmin, int pmax, bool any_flagged, const double *d_wts,
                          double2 *d_w12) {
    Pointing &P = pb.pt;
    const int64_t m = P.m;
    hipStream_t s = pb.stream;
    P.trash0 = P.trash = any_flagged ? 1 : 0;
    const int64_t span = (int64_t) pmax - pmin + 1;

    unsigned *d_flags, *d_rank;
    MGPU_RET(pb.arena.alloc(&d_flags, (size_t) span + 1));
    MGPU_RET(pb.arena.alloc(&d_rank, (size_t) span + 1));
    MGPU_HIP(hipMemsetAsync(d_flags, 0, sizeof(unsigned) * (span + 1), s));
    k_mark<<<grid_for(m, 256), 256, 0, s>>>(d_p, m, pmin, d_flags);
    MGPU_LAUNCHED();
    size_t tb = 0;
    MGPU_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, d_flags, d_rank, span + 1, s));
    MGPU_RET(pb.cub_tmp.ensure(tb));
    MGPU_HIP(hipcub::DeviceScan::ExclusiveSum(pb.cub_tmp.p, tb, d_flags, d_rank, span + 1, s));
    unsigned npix_u = 0;
    MGPU_HIP(hipMemcpyAsync(&npix_u, d_rank + span, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipStreamSynchronize(s));
    P.npix0 = P.npix = (int) npix_u;

    MGPU_RET(pb.arena.alloc(&P.d_q, (size_t) m));
    if (!P.unit_w0) MGPU_RET(pb.arena.alloc(&P.d_w0, (size_t) m));
    if (d_w12)
        P.d_w12 = d_w12;  // the packed upload is P's Q/U weights (sample order)
    else
        MGPU_RET(pb.arena.alloc(&P.d_w12, (size_t) m));
    MGPU_RET(pb.arena.alloc(&P.d_pixg, (size_t) P.npix0));
    unsigned *d_key, *d_skey;
    int *d_t;
    MGPU_RET(pb.arena.alloc(&d_key, (size_t) m));
    MGPU_RET(pb.arena.alloc(&d_skey, (size_t) m));
    MGPU_RET(pb.arena.alloc(&d_t, (size_t) m));
    MGPU_RET(pb.arena.alloc(&P.d_cst, (size_t) m));
    k_assign<<<grid_for(m, 256), 256, 0, s>>>(d_p, m, pmin, d_rank, d_wts, P.d_q, P.d_w0, P.d_w12, d_key, d_t,
                                              (unsigned) P.npix0);
    MGPU_LAUNCHED();
    k_pixg<<<grid_for(span, 256), 256, 0, s>>>(d_flags, d_rank, span, pmin, P.d_pixg);
    MGPU_LAUNCHED();

    // stable radix sort of (pixel, t): t stays increasing within each pixel
    int end_bit = 1;
    while (end_bit < 32 && ((unsigned) P.npix0 >> end_bit) != 0) end_bit++;
    tb = 0;
    MGPU_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, d_key, d_skey, d_t, P.d_cst, (int) m, 0, end_bit, s));
    MGPU_RET(pb.cub_tmp.ensure(tb));
    MGPU_HIP(hipcub::DeviceRadixSort::SortPairs(pb.cub_tmp.p, tb, d_key, d_skey, d_t, P.d_cst, (int) m, 0, end_bit,
                                                s));
    int64_t *d_cnt;
    MGPU_RET(pb.arena.alloc(&d_cnt, 1));
    MGPU_HIP(hipMemsetAsync(d_cnt, 0, sizeof(int64_t), s));
    k_count_valid<<<grid_for(m, 256), 256, 0, s>>>(d_skey, m, (unsigned) P.npix0, d_cnt);
    MGPU_LAUNCHED();
    int64_t nvalid = 0;
    MGPU_HIP(hipMemcpyAsync(&nvalid, d_cnt, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipStreamSynchronize(s));
    P.nvalid = nvalid;
    MGPU_RET(pb.arena.alloc(&P.d_off, (size_t) P.npix0 + 1));
    k_offsets<<<grid_for(nvalid > 0 ? nvalid : 1, 256), 256, 0, s>>>(d_skey, nvalid, P.npix0, P.d_off);
    MGPU_LAUNCHED();
    MGPU_RET(build_ell(pb));
    MGPU_HIP(hipStreamSynchronize(s));
    for (void *p : {(void *) d_cnt, (void *) d_key, (void *) d_skey, (void *) d_t, (void *) d_flags, (void *) d_rank})
        pb.arena.release(p);
    return MGPU_OK;
}

int pointing_compact(Problem &pb, const unsigned *d_kept) {
    Pointing &P = pb.pt;
    hipStream_t s = pb.stream;
    const int npix = P.npix;
    unsigned *d_keep, *d_new;
    int *d_cnt, *d_hscan;
    MGPU_RET(pb.arena.alloc(&d_keep, (size_t) npix + 1));
    MGPU_RET(pb.arena.alloc(&d_new, (size_t) npix + 1));
    MGPU_RET(pb.arena.alloc(&d_cnt, (size_t) npix + 1));
    MGPU_RET(pb.arena.alloc(&d_hscan, (size_t) npix + 1));
    k_kept_counts<<<grid_for(npix + 1, 256), 256, 0, s>>>(d_kept, P.d_off, npix, d_keep, d_cnt);
    MGPU_LAUNCHED();
    size_t tb = 0, tb2 = 0;
    MGPU_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, d_keep, d_new, npix + 1, s));
    MGPU_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, d_cnt, d_hscan, npix + 1, s));
    MGPU_RET(pb.cub_tmp.ensure(tb > tb2 ? tb : tb2));
    MGPU_HIP(hipcub::DeviceScan::ExclusiveSum(pb.cub_tmp.p, tb, d_keep, d_new, npix + 1, s));
    MGPU_HIP(hipcub::DeviceScan::ExclusiveSum(pb.cub_tmp.p, tb2, d_cnt, d_hscan, npix + 1, s));
    unsigned nnew = 0;
    int nvalid_new = 0;
    MGPU_HIP(hipMemcpyAsync(&nnew, d_new + npix, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipMemcpyAsync(&nvalid_new, d_hscan + npix, sizeof(int), hipMemcpyDeviceToHost, s));
    MGPU_HIP(hipStreamSynchronize(s));
    int *d_pixg2, *d_off2, *d_cst2;
    MGPU_RET(pb.arena.alloc(&d_pixg2, (size_t) nnew));
    MGPU_RET(pb.arena.alloc(&d_off2, (size_t) nnew + 1));
    MGPU_RET(pb.arena.alloc(&d_cst2, (size_t) nvalid_new));
    k_compact_pixels<<<grid_for(npix + 1, 256), 256, 0, s>>>(d_kept, d_new, d_hscan, P.d_pixg, npix, d_pixg2, d_off2,
                                                             (int) nnew);
    MGPU_LAUNCHED();
    if (npix > 0) {
        k_compact_hits<<<npix, 64, 0, s>>>(d_kept, d_new, P.d_off, P.d_cst, d_off2, npix, d_cst2);
        MGPU_LAUNCHED();
    }
    k_remap_q<<<grid_for(P.m, 256), 256, 0, s>>>(P.d_q, P.m, d_kept, d_new);
    MGPU_LAUNCHED();
    MGPU_HIP(hipStreamSynchronize(s));
    pb.arena.release(P.d_pixg);
    pb.arena.release(P.d_off);
    pb.arena.release(P.d_cst);
    P.d_pixg = d_pixg2;
    P.d_off = d_off2;
    P.d_cst = d_cst2;
    P.nvalid = nvalid_new;
    P.npix = (int) nnew;
    MGPU_RET(build_ell(pb));
    MGPU_HIP(hipStreamSynchronize(s));
    for (void *p : {(void *) d_keep, (void *) d_new, (void *) d_cnt, (void *) d_hscan}) pb.arena.release(p);
    return MGPU_OK;
}

int pointing_make_fp32(Problem &pb) {
    Pointing &P = pb.pt;
    if (P.d_w12f) return MGPU_OK;
    MGPU_RET(pb.arena.alloc(&P.d_w12f, (size_t) P.m));
    MGPU_RET(pb.arena.alloc(&P.d_cw12f, (size_t) P.nh));
    k_to_f32x2<<<grid_for(P.m, 256), 256, 0, pb.stream>>>(P.d_w12, P.d_w12f, P.m);
    if (P.nh > 0) k_to_f32x2<<<grid_for(P.nh, 256), 256, 0, pb.stream>>>(P.d_cw12, P.d_cw12f, P.nh);
    if (!P.unit_w0) {
        MGPU_RET(pb.arena.alloc(&P.d_w0f, (size_t) P.m));
        MGPU_RET(pb.arena.alloc(&P.d_cw0f, (size_t) P.nh));
        k_to_f32<<<grid_for(P.m, 256), 256, 0, pb.stream>>>(P.d_w0, P.d_w0f, P.m);
        if (P.nh > 0) k_to_f32<<<grid_for(P.nh, 256), 256, 0, pb.stream>>>(P.d_cw0, P.d_cw0f, P.nh);
    }
    MGPU_LAUNCHED();
    return MGPU_OK;
}

// local indices of MatLocalShape for every (t,k): valid rows 3(q+trash)+k,
// trash rows k (lindices starts with -3,-2,-1)
__global__ void k_local_indices(const int *__restrict__ q, int64_t m, int trash, int *__restrict__ out) {
    int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) {
        int qq = q[t];
        for (int k = 0; k < 3; k++) out[3 * t + k] = qq >= 0 ? 3 * (qq + trash) + k : k;
    }
}

__global__ void k_lindices(const int *__restrict__ pixg, int npix, int trash, int *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3 * trash) out[i] = i - 3;
    if (i < npix)
        for (int k = 0; k < 3; k++) out[3 * trash + 3 * i + k] = 3 * pixg[i] + k;
}

int pointing_local_indices(Problem &pb, int *d_out) {
    k_local_indices<<<grid_for(pb.pt.m, 256), 256, 0, pb.stream>>>(pb.pt.d_q, pb.pt.m, pb.pt.trash, d_out);
    MGPU_LAUNCHED();
    return MGPU_OK;
}

int pointing_lindices(Problem &pb, int *d_out) {
    int n = pb.pt.npix > 3 ? pb.pt.npix : 3;
    k_lindices<<<grid_for(n, 256), 256, 0, pb.stream>>>(pb.pt.d_pixg, pb.pt.npix, pb.pt.trash, d_out);
    MGPU_LAUNCHED();
    return MGPU_OK;
}

// ---------------------------------------------------------------------------
// P: y = P x, sample-major, with fused per-(chunk,block) partial sums
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wg_sum_256(double v, double *red) {
    // deterministic: fixed xor-butterfly inside each wave, then waves in order
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        for (int i = 0; i < nw; i++) s += red[i];
    }
    return s;  // valid in thread 0
}

// One workgroup per CPW chunks of the chunk list perm[0..nperm): every
// streaming load of its chunks is issued before the map gathers.  (Two chunks
// per workgroup in FP32, for more bytes in flight per lane, measured slower on
// the 64 x 2^20 raster: 0.248 against 0.232 ms per launch.)
template <typename TX, typename TW, typename TW2, bool W0, bool PARTIALS, int CPW>
__global__ __launch_bounds__(kPBlock) void k_matvec(const int *__restrict__ q, const TW *__restrict__ w0,
                                                    const TW2 *__restrict__ w12, const TX *__restrict__ x,
                                                    TX *__restrict__ y, int64_t m, const int64_t *__restrict__ start,
                                                    const int *__restrict__ chunk_block, int nb,
                                                    double *__restrict__ partials, const int *__restrict__ perm,
                                                    int nperm, const int *__restrict__ chunk_seg,
                                                    const int *__restrict__ state) {
    __shared__ double red[kPBlock / 64];
    constexpr int PER = kChunk / kPBlock;
    int64_t cs[CPW];
    bool live[CPW];  // uniform over the workgroup
#pragma unroll
    for (int j = 0; j < CPW; j++) {
        const int k = CPW * (int) blockIdx.x + j;
        cs[j] = k < nperm ? perm[k] : 0;
        live[j] = k < nperm;
        if (live[j] && state) {  // the chunk's segment has stopped: nothing of it is used any more
            const int sg = chunk_seg[cs[j]];
            if (sg >= 0 && state[sg]) live[j] = false;
        }
    }
    int qq[CPW][PER];
    TW2 wv[CPW][PER];
#pragma unroll
    for (int j = 0; j < CPW; j++)
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int64_t t = cs[j] * kChunk + threadIdx.x + (int64_t) i * kPBlock;
            qq[j][i] = -1;
            if (live[j] && t < m) {
                qq[j][i] = q[t];
                wv[j][i] = w12[t];
            }
        }
    TX yv[CPW][PER];
#pragma unroll
    for (int j = 0; j < CPW; j++)
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int64_t t = cs[j] * kChunk + threadIdx.x + (int64_t) i * kPBlock;
            TX v = (TX) 0;
            if (qq[j][i] >= 0) {
                const TX *xp = x + 3 * (int64_t) qq[j][i];
                const TX x0 = xp[0], x1 = xp[1], x2 = xp[2];
                TX acc = (TX) 0;
                acc += (W0 ? (TX) w0[t] * x0 : x0);
                acc += (TX) wv[j][i].x * x1;
                acc += (TX) wv[j][i].y * x2;
                v = acc;
            }
            if (live[j] && t < m) y[t] = v;
            yv[j][i] = v;
        }
    if (PARTIALS) {
#pragma unroll
        for (int j = 0; j < CPW; j++) {
            if (!live[j]) continue;
            const int64_t c = cs[j], c0 = c * kChunk;
            const int64_t c1 = min(c0 + (int64_t) kChunk, m);
            int b = chunk_block[c];
            for (; b < nb && start[b] < c1; b++) {
                const int64_t lo = start[b], hi = start[b + 1];
                double s = 0.0;
#pragma unroll
                for (int i = 0; i < PER; i++) {
                    const int64_t t = c0 + threadIdx.x + (int64_t) i * kPBlock;
                    if (t >= lo && t < hi) s += (double) yv[j][i];
                }
                s = wg_sum_256(s, red);
                if (threadIdx.x == 0) partials[c + b] = s;
            }
        }
    }
}

template <typename TX, typename TW, typename TW2>
static int launch_matvec_t(Problem &pb, const TX *d_x, TX *d_y, bool partials, const TW *w0, const TW2 *w12) {
    constexpr int CPW = 1;  // chunks per workgroup (both precisions)
    const Pointing &P = pb.pt;
    const ToeplitzPlan &T = pb.tp;
    int nperm = (int) grid_for(P.m, kChunk);
    partials = partials && T.nb > 0 && T.d_partials && T.d_chunk_block && T.lambda > 1 &&
               T.offset_mode != MGPU_OFFSET_NONE;
    const int *q = P.d_q;
    const int64_t *st = T.d_start;
    const int *cb = T.d_chunk_block;
    const int *pm = P.d_chunk_perm;
    MGPU_ARG(pm, "pointing has no chunk order");
    if (pb.live_state && P.nch_run >= 0) {  // segmented PCG: the running segments' chunks only
        pm = P.d_perm_run;
        nperm = P.nch_run;
        if (nperm == 0) return MGPU_OK;
    }
    const unsigned g = (unsigned) ((nperm + CPW - 1) / CPW);
    const int *cs = pb.sg.d_chunk_seg, *ls = cs ? pb.live_state : nullptr;
    if (P.unit_w0) {
        if (partials)
            k_matvec<TX, TW, TW2, false, true, CPW><<<g, kPBlock, 0, pb.stream>>>(
                    q, nullptr, w12, d_x, d_y, P.m, st, cb, T.nb, T.d_partials, pm, nperm, cs, ls);
        else
            k_matvec<TX, TW, TW2, false, false, CPW><<<g, kPBlock, 0, pb.stream>>>(
                    q, nullptr, w12, d_x, d_y, P.m, st, cb, T.nb, nullptr, pm, nperm, cs, ls);
    } else {
        if (partials)
            k_matvec<TX, TW, TW2, true, true, CPW><<<g, kPBlock, 0, pb.stream>>>(
                    q, w0, w12, d_x, d_y, P.m, st, cb, T.nb, T.d_partials, pm, nperm, cs, ls);
        else
            k_matvec<TX, TW, TW2, true, false, CPW><<<g, kPBlock, 0, pb.stream>>>(
                    q, w0, w12, d_x, d_y, P.m, st, cb, T.nb, nullptr, pm, nperm, cs, ls);
    }
    MGPU_LAUNCHED();
    return MGPU_OK;
}

int launch_matvec(Problem &pb, const double *d_x, double *d_y, bool partials) {
    return launch_matvec_t<double, double, double2>(pb, d_x, d_y, partials, pb.pt.d_w0, pb.pt.d_w12);
}

int launch_matvec_f32(Problem &pb, const float *d_x, float *d_y, bool partials) {
    return launch_matvec_t<float, float, float2>(pb, d_x, d_y, partials, pb.pt.d_w0f, pb.pt.d_w12f);
}

// ---------------------------------------------------------------------------
// P^T: x = P^T y over sliced ELL; optional fused dot <x, h>
// ---------------------------------------------------------------------------
constexpr int kPtBlock = 256;

// Workgroups follow the map segments (one segment per workgroup, first pixel
// wg_q0[wg]), so the fused dot partials are those of the PCG map kernels.
template <typename TY, typename TW, typename TW2, bool W0, bool DOT>
__global__ __launch_bounds__(kPtBlock) void k_trmatvec(const int64_t *__restrict__ base, const int *__restrict__ ct,
                                                       const TW *__restrict__ cw0, const TW2 *__restrict__ cw12,
                                                       const TY *__restrict__ y, double *__restrict__ x,
                                                       const int *__restrict__ wg_seg, const int *__restrict__ wg_q0,
                                                       const int *__restrict__ seg_q, const double *__restrict__ hvec,
                                                       double *__restrict__ dotpart, const int *__restrict__ state) {
    if (state && state[wg_seg[blockIdx.x]]) return;  // stopped segment (uniform over the workgroup)
    __shared__ double red[kPtBlock / 64];
    const int qq = wg_q0[blockIdx.x] + (int) threadIdx.x;
    const int qend = seg_q[wg_seg[blockIdx.x] + 1];
    const int s = qq / kSlice, lane = qq % kSlice;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    if (qq < qend) {
        // The slice width is uniform over the wave: every lane walks all of it,
        // its padding entries (t = -1, weights 0) adding +0.0, which leaves the
        // sums unchanged (they start at +0.0 and so are never -0.0).  U hits
        // per step keep 2U independent loads in flight per lane (on the
        // 64 x 2^20 raster, FP64: 4 hits 0.395 ms, 8 hits 0.386 ms, 16 hits
        // 0.380 ms; FP32: 4 hits 0.222 ms, 16 hits 0.243 ms).
        const int64_t b0 = base[s];
        const int width = (int) ((base[s + 1] - b0) / kSlice);
        const int *__restrict__ cte = ct + b0 + lane;
        const TW2 *__restrict__ cwe = cw12 + b0 + lane;
        int i = 0;
        constexpr int U = sizeof(TY) == 8 ? 16 : 4;
        for (; i + U <= width; i += U) {
            int t[U];
            TW2 w[U];
            double yt[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                t[u] = cte[(i + u) * kSlice];
                w[u] = cwe[(i + u) * kSlice];
            }
#pragma unroll
            for (int u = 0; u < U; u++) yt[u] = t[u] >= 0 ? (double) y[t[u]] : 0.0;
#pragma unroll
            for (int u = 0; u < U; u++) {
                a0 += (W0 ? (double) cw0[b0 + lane + (int64_t) (i + u) * kSlice] * yt[u] : yt[u]);
                a1 += (double) w[u].x * yt[u];
                a2 += (double) w[u].y * yt[u];
            }
        }
        for (; i < width; i++) {
            const int t = cte[i * kSlice];
            const TW2 w = cwe[i * kSlice];
            const double yt = t >= 0 ? (double) y[t] : 0.0;
            a0 += (W0 ? (double) cw0[b0 + lane + (int64_t) i * kSlice] * yt : yt);
            a1 += (double) w.x * yt;
            a2 += (double) w.y * yt;
        }
        x[3 * qq] = a0;
        x[3 * qq + 1] = a1;
        x[3 * qq + 2] = a2;
    }
    if (DOT) {
        double d = 0.0;
        if (qq < qend) {
            d += hvec[3 * qq] * a0;
            d += hvec[3 * qq + 1] * a1;
            d += hvec[3 * qq + 2] * a2;
        }
        d = wg_sum_256(d, red);
        if (threadIdx.x == 0) dotpart[blockIdx.x] = d;
    }
}

int trmatvec_nparts(const Problem &pb) { return pb.sg.nwg; }

template <typename TY, typename TW, typename TW2>
static int launch_trmatvec_t(Problem &pb, const TY *d_y, double *d_x, const double *d_h, double *d_dot,
                             const TW *cw0, const TW2 *cw12, int *nparts) {
    static_assert(kPtBlock == kMapWG, "P^T workgroups are the map-domain workgroups");
    MGPU_RET(segments_ensure(pb));
    const Pointing &P = pb.pt;
    const Segments &G = pb.sg;
    const unsigned g = (unsigned) G.nwg;
    if (nparts) *nparts = (int) g;
    if (g == 0) return MGPU_OK;
    const bool dot = d_h && d_dot;
    const int64_t *b = P.d_ell_base;
    const int *ws = G.d_wg_seg, *wq = G.d_wg_q0, *sq = G.d_q;
    if (P.unit_w0) {
        if (dot)
            k_trmatvec<TY, TW, TW2, false, true>
                    <<<g, kPtBlock, 0, pb.stream>>>(b, P.d_ct, nullptr, cw12, d_y, d_x, ws, wq, sq, d_h, d_dot, pb.live_state);
        else
            k_trmatvec<TY, TW, TW2, false, false>
                    <<<g, kPtBlock, 0, pb.stream>>>(b, P.d_ct, nullptr, cw12, d_y, d_x, ws, wq, sq, nullptr, nullptr, pb.live_state);
    } else {
        if (dot)
            k_trmatvec<TY, TW, TW2, true, true>
                    <<<g, kPtBlock, 0, pb.stream>>>(b, P.d_ct, cw0, cw12, d_y, d_x, ws, wq, sq, d_h, d_dot, pb.live_state);
        else
            k_trmatvec<TY, TW, TW2, true, false>
                    <<<g, kPtBlock, 0, pb.stream>>>(b, P.d_ct, cw0, cw12, d_y, d_x, ws, wq, sq, nullptr, nullptr, pb.live_state);
    }
    MGPU_LAUNCHED();
    return MGPU_OK;
}

int launch_trmatvec(Problem &pb, const double *d_y, double *d_x, const double *d_h, double *d_dot, int *nparts) {
    return launch_trmatvec_t<double, double, double2>(pb, d_y, d_x, d_h, d_dot, pb.pt.d_cw0, pb.pt.d_cw12, nparts);
}

int launch_trmatvec_f32(Problem &pb, const float *d_y, double *d_x, const double *d_h, double *d_dot,
                        int *nparts) {
    return launch_trmatvec_t<float, float, float2>(pb, d_y, d_x, d_h, d_dot, pb.pt.d_cw0f, pb.pt.d_cw12f, nparts);
}

}  // namespace mgpu
