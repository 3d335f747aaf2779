"""The allocation-free N^-1 entry point (mappraiser_gpu_apply_weights_noalloc)
against mappraiser_gpu_apply_weights, through the C-ABI: two calls in a row
on two problems (the second smaller), so the second call runs on the scratch
array the first one left behind.  Both must be bit-identical to the allocating
entry point.  The test's device vectors come from the HIP runtime the library
itself runs on (hipMalloc through ctypes)."""
import ctypes as ct

import numpy as np
import pytest

from midapack_amd import workloads as W

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    hip = ct.CDLL("libamdhip64.so.7")  # the runtime already loaded by the library
    hip.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    hip.hipFree.argtypes = [ct.c_void_p]
    hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    return hip


class DeviceArray:
    def __init__(self, hip, host):
        self.hip, self.n = hip, host.nbytes
        self.ptr = ct.c_void_p()
        assert hip.hipMalloc(ct.byref(self.ptr), self.n) == 0
        assert hip.hipMemcpy(self.ptr, host.ctypes.data, self.n, H2D) == 0

    def get(self, like):
        out = np.empty_like(like)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, self.n, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def test_noalloc_apply_weights_matches_allocating_entry(gpu):
    hip = _hip()
    cases = [
        (W.make_workload("noalloc-a", 4, 1 << 17, 128, 128, 8192, 31, block_len=1 << 16,
                         fknee_range=(0.05, 0.5), alpha_range=(1.0, 2.0)), gpu.MGPU_OFFSET_SYMMETRIC),
        (W.make_workload("noalloc-b", 2, 40000, 64, 64, 1024, 32, block_len=15000), gpu.MGPU_OFFSET_REFERENCE),
    ]
    for i, (wl, mode) in enumerate(cases):
        print(f"case: {wl.name}, m {wl.m}, lambda {wl.lam}, offset mode {mode}", flush=True)
        v = np.random.default_rng(40 + i).standard_normal(wl.m) + 0.3
        with gpu.Problem(wl.pix, wl.weights, wl.block_sizes, wl.lam, wl.invtt,
                         opts=gpu.options(offset_mode=mode)) as pb:
            ref = pb.apply_weights(v)
            d_in = DeviceArray(hip, v)
            d_out = DeviceArray(hip, np.full_like(v, np.nan))
            try:
                pb.apply_weights_noalloc(d_in.ptr.value, d_out.ptr.value)
                out = d_out.get(v)
                back = d_in.get(v)
            finally:
                d_in.free()
                d_out.free()
        print(f"{wl.name}: max |noalloc - apply_weights| = {np.max(np.abs(out - ref)):.3e}", flush=True)
        assert np.array_equal(back, v), "input modified"
        assert np.array_equal(out, ref)
