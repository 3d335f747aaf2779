"""Noise-model builder on the device (mappraiser_gpu_noise_invtt) against the
oracle's restatement of noise2invtt (scipy periodogram + curve_fit).

The device fit (bounded Levenberg-Marquardt) and the reference's (scipy TRF)
minimise the same objective; parameters and rows agree to the fit tolerance
stated below.  Block lengths cover the power-of-two FFT path and Bluestein
(odd and even lengths)."""
from __future__ import annotations

import numpy as np
import pytest

from midapack_amd import workloads as W
from oracle import noise as N

pytestmark = pytest.mark.gpu

FS = 10.0
PARAM_RTOL = 1e-4   # fitted (a, alpha, fknee): relative
ROW_RTOL = 1e-4     # invtt rows: max |diff| / max |row|


def _noise(lens, seed):
    rng = np.random.default_rng(seed)
    parts = []
    for i, n in enumerate(lens):
        fk, al = rng.uniform(0.05, 0.5), rng.uniform(1.0, 2.0)
        parts.append(W.one_over_f_noise(np.random.default_rng([seed, i]), n, 1.0 + 0.2 * i, fk, al))
    return np.concatenate(parts)


@pytest.mark.parametrize("lens,lam,kind", [
    ([65536, 50001, 40000], 1024, "chebwin"),
    ([131072, 131072], 8192, "chebwin"),
    ([30000, 32768], 512, "gaussian"),
], ids=["pow2+bluestein-odd+even-l1024", "c2-blocks-l8192", "gaussian-l512"])
def test_noise_invtt_matches_oracle(gpu, lens, lam, kind):
    print(f"case: blocks {lens}, lambda {lam}, {kind}", flush=True)
    x = _noise(lens, sum(lens) % 1000)
    ref, rpar = N.compute_invtt(x, lens, lam, FS, kind)
    out, par = gpu.noise_invtt(x, np.array(lens, np.int32), lam, FS,
                               gpu.MGPU_APOD_CHEBWIN if kind == "chebwin" else gpu.MGPU_APOD_GAUSSIAN)
    for b in range(len(lens)):
        dp = np.abs(par[b, :3] - rpar[b, :3]) / np.abs(rpar[b, :3])
        r, o = ref[b * lam:(b + 1) * lam], out[b * lam:(b + 1) * lam]
        er = float(np.max(np.abs(o - r)) / np.max(np.abs(r)))
        print(f"block {b} (n={lens[b]}): params {par[b]} vs {rpar[b]}, rel {dp}, row error {er:.2e}", flush=True)
        assert np.all(dp <= PARAM_RTOL)
        assert er <= ROW_RTOL
