"""GPU parity at the full one-rank workload: 64 detectors x 2^20 samples on a
1024^2 grid, lambda = 8192, 2^17-sample Toeplitz blocks (512 blocks sharing 64
filters, 4096 FFT tiles per product, 67M samples, about a million pixels).

Checked against the CPU oracle on the same seeded inputs:
  * lindices / local indices (MatLocalShape, mapmat.c:496-521) bit-exact;
  * P and P^T (mapmat.c:704-729, 819-859) bit-exact;
  * N^-1 (stbmmProd, toeplitz_wizard.c:80-111) within 1e-12 of the oracle's FFT
    product, with the reference's offsets (T C) and the symmetric ones (C T C);
  * cond, hits, the cut set and the cut signal (precond.c:456-702) bit-exact;
  * PCG_GLS_true (pcg_true.c:20-241): x_3 within 4e-12 and the residual norms
    of iterations 0..3 within 1e-10 relative;
  * full FP64 and mixed-precision solves to 1e-8 converge, the mixed map within
    1e-6 of the FP64 one and its FP64 true residual at or below the tolerance.
The oracle's serial C build takes about a minute for its part on one core.
"""
import time

import numpy as np
import pytest

from midapack_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _log(msg):
    print(f"[c2] {msg}", flush=True)


@pytest.fixture(scope="module")
def c2():
    t = time.time()
    wl = W.config_c2()
    _log(f"generated {wl.name}: m={wl.m} in {time.time() - t:.1f}s")
    return wl


@pytest.fixture(scope="module")
def c2_oracle_pcg(c2):
    """The oracle's MLmap setup and 3 PCG iterations (reference offsets)."""
    wl = c2
    t = time.time()
    with O.ninv_mode(O.NINV_REAL_SPECTRUM):
        r = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-30,
                        maxiter=20, stop_after=3)
    _log(f"oracle setup + 3 iterations in {time.time() - t:.1f}s ({r.n_cut} pixels cut)")
    return r


def test_c2_local_shape_and_products(gpu, c2):
    wl = c2
    _log("case: index tables, P and P^T at full size")
    t = time.time()
    lind, loc, tr = O.local_shape(wl.pix)
    _log(f"oracle local shape in {time.time() - t:.1f}s: lcount {lind.size}")
    rng = np.random.default_rng(202)
    x = rng.standard_normal(lind.size)
    y = rng.standard_normal(wl.m)
    with gpu.Problem(wl.pix, wl.weights, wl.block_sizes, wl.lam, wl.invtt) as pb:
        glind, gidx, gtr = pb.local_shape()
        assert gtr == tr
        assert np.array_equal(glind, lind)
        assert np.array_equal(gidx, loc)
        px = pb.matvec(x)
        pty = pb.trmatvec(y)
    assert np.array_equal(px, O.matvec(loc, wl.weights, tr, 3, x)), "P x differs from the oracle"
    assert np.array_equal(pty, O.trmatvec(loc, wl.weights, tr, 3, lind.size, y)), "P^T y differs from the oracle"


@pytest.mark.parametrize("mode", ["reference", "symmetric"])
def test_c2_ninv(gpu, c2, mode):
    wl = c2
    _log(f"case: N^-1 on 512 blocks, {mode} offsets")
    gmode = gpu.MGPU_OFFSET_REFERENCE if mode == "reference" else gpu.MGPU_OFFSET_SYMMETRIC
    omode = O.OFFSET_REFERENCE if mode == "reference" else O.OFFSET_SYMMETRIC
    v = np.random.default_rng(203).standard_normal(wl.m) + 0.25
    with gpu.Problem(wl.pix, wl.weights, wl.block_sizes, wl.lam, wl.invtt,
                     opts=gpu.options(offset_mode=gmode)) as pb:
        out = pb.apply_weights(v)
    t = time.time()
    with O.ninv_mode(O.NINV_REAL_SPECTRUM), O.offset_mode(omode):
        ref = O.apply_weights(v, wl.block_sizes, wl.lam, wl.invtt)
    err = _rel(out, ref)
    _log(f"N^-1 {mode}: oracle {time.time() - t:.1f}s, rel. diff {err:.2e}")
    assert err <= 1e-12


def test_c2_precond_bitexact(gpu, c2, c2_oracle_pcg):
    wl, r = c2, c2_oracle_pcg
    _log("case: block-Jacobi preconditioner with the cut at full size")
    with gpu.Problem(wl.pix, wl.weights, wl.block_sizes, wl.lam, wl.invtt) as pb:
        sig, cond, hits = pb.precond_build(wl.signal)
        lind, idx, tr = pb.local_shape()
    _log(f"{r.n_cut} pixels cut, {cond.size} kept")
    assert tr == r.trash_pix
    assert np.array_equal(lind, r.lindices) and np.array_equal(idx, r.pix_local)
    assert np.array_equal(hits, r.hits)
    assert np.array_equal(cond, r.cond)
    assert np.array_equal(sig, r.signal)


def test_c2_pcg_iterates(gpu, c2, c2_oracle_pcg):
    wl, r = c2, c2_oracle_pcg
    _log("case: PCG iterates x_3 (reference offsets)")
    with gpu.Problem(wl.pix, wl.weights, wl.block_sizes, wl.lam, wl.invtt) as pb:
        sig, _, _ = pb.precond_build(wl.signal)
        x3, hist, rep = pb.pcg(sig, wl.noise, tol=1e-30, maxiter=20, stop_after=3)
    err = _rel(x3, r.x)
    _log(f"x_3 rel. diff {err:.2e}; residuals GPU {hist} oracle {r.res_hist}")
    assert rep.iterations == r.iterations == 3
    assert err <= 4e-12
    np.testing.assert_allclose(hist, r.res_hist, rtol=1e-10)


def test_c2_full_solves_fp64_and_mixed(gpu, c2):
    wl = c2
    _log("case: full FP64 and mixed solves to 1e-8 (symmetric offsets)")
    sym = gpu.MGPU_OFFSET_SYMMETRIC
    g64 = gpu.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                      maxiter=2000, opts=gpu.options(offset_mode=sym))
    gmx = gpu.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                      maxiter=2000, opts=gpu.options(precision=gpu.MGPU_MIXED, offset_mode=sym))
    err = _rel(gmx.x, g64.x)
    _log(f"FP64 {g64.report['iterations']} it (converged {g64.report['converged']}), mixed "
         f"{gmx.report['iterations']} it (converged {gmx.report['converged']}, true residual "
         f"{gmx.report['true_res_rel']:.2e}, {gmx.report['refreshes']} refreshes), map error {err:.2e}")
    assert g64.report["converged"] and gmx.report["converged"]
    assert gmx.report["true_res_rel"] <= 1e-8
    assert err <= 1e-6
