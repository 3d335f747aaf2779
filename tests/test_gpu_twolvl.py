"""GPU parity of the two-level preconditioners (MLmap precond = 1 a priori,
2 a posteriori; build_precond precond.c:1310-1398, apply_precond 1412-1422)
against the pinned CPU oracle on the same seeded inputs.

Bars: PCG iterates x_k within 1e-10 (1 + k) relative for small k (the
deflation space and pinv(E) add rounding that E's condition number, about
3e2 here, amplifies); full solves converge with iteration counts within 2 of
the oracle's and maps within 1e-7 relative; two runs bit-identical; the
HIP-graph replay bit-identical to the eager launches; mixed precision within
1e-6 of the FP64 map.
"""
import numpy as np
import pytest

from midapack_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [(1, 1), (1, 4), (2, 4), (2, 8)]


@pytest.mark.parametrize("precond,zn", CASES)
@pytest.mark.parametrize("k", [1, 3])
def test_twolvl_iterates(gpu, precond, zn, k):
    wl = W.config_small(seed=11, ndet=3, nsamp=8192, nx=24, ny=24, lam=128)
    print(f"case: precond {precond} Zn {zn} x_{k}", flush=True)
    with O.ninv_mode(O.NINV_REAL_SPECTRUM):
        r = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-30,
                        maxiter=50, stop_after=k, precond=precond, z_2lvl=zn)
    o = gpu.options(precond=precond, z_2lvl=zn)
    with gpu.Problem(wl.pix, wl.weights, wl.block_sizes, wl.lam, wl.invtt, opts=o) as pb:
        sig, _, _ = pb.precond_build(wl.signal)
        x, hist, rep = pb.pcg(sig, wl.noise, tol=1e-30, maxiter=50, stop_after=k)
    assert rep.iterations == r.iterations == k
    print(f"x_{k} rel diff {_rel(x, r.x):.2e}; build {rep.t_2lvl_ms:.2f} ms", flush=True)
    assert _rel(x, r.x) <= 1e-10 * (1 + k)
    np.testing.assert_allclose(hist, r.res_hist, rtol=1e-9)


@pytest.mark.parametrize("precond,zn", [(1, 4), (2, 4), (2, 8)])
@pytest.mark.parametrize("mode", ["reference", "symmetric"])
def test_twolvl_full_solve(gpu, precond, zn, mode):
    wl = W.make_workload("c1-shape", 2, 65536, 64, 64, 1024, 3)
    print(f"case: 2 x 65536, lambda 1024, precond {precond} Zn {zn}, {mode} offsets", flush=True)
    om = O.OFFSET_REFERENCE if mode == "reference" else O.OFFSET_SYMMETRIC
    gm = gpu.MGPU_OFFSET_REFERENCE if mode == "reference" else gpu.MGPU_OFFSET_SYMMETRIC
    with O.ninv_mode(O.NINV_REAL_SPECTRUM), O.offset_mode(om):
        r = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                        maxiter=300, precond=precond, z_2lvl=zn)
    o = gpu.options(precond=precond, z_2lvl=zn, offset_mode=gm)
    g1 = gpu.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                     maxiter=300, opts=o)
    g2 = gpu.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                     maxiter=300, opts=o)
    assert np.array_equal(g1.x, g2.x), "two runs differ"
    assert g1.report["converged"] and r.converged
    print(f"iterations GPU {g1.report['iterations']} oracle {r.iterations}; map rel diff {_rel(g1.x, r.x):.2e}",
          flush=True)
    assert abs(g1.report["iterations"] - r.iterations) <= 2
    assert _rel(g1.x, r.x) <= 1e-7


def test_twolvl_graph_and_mixed(gpu):
    wl = W.make_workload("c1-shape", 2, 65536, 64, 64, 1024, 3)
    sym = gpu.MGPU_OFFSET_SYMMETRIC
    print("case: precond 2 Zn 8, eager vs graph (FP64), mixed", flush=True)
    args = (wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt)
    e = gpu.pcg_gls(*args, tol=1e-8, maxiter=300, opts=gpu.options(precond=2, z_2lvl=8, offset_mode=sym, graph=0))
    g = gpu.pcg_gls(*args, tol=1e-8, maxiter=300, opts=gpu.options(precond=2, z_2lvl=8, offset_mode=sym, graph=1))
    assert np.array_equal(e.x, g.x) and e.report["iterations"] == g.report["iterations"]
    mx = gpu.pcg_gls(*args, tol=1e-8, maxiter=300,
                     opts=gpu.options(precision=gpu.MGPU_MIXED, precond=2, z_2lvl=8, offset_mode=sym))
    assert mx.report["converged"]
    print(f"mixed vs FP64 map {_rel(mx.x, e.x):.2e}, iterations {mx.report['iterations']} / "
          f"{e.report['iterations']}, true residual {mx.report['true_res_rel']:.2e}", flush=True)
    # both maps meet the residual test; the a posteriori space leaves this system
    # slow to converge (161 iterations against 44 for block Jacobi), and the two
    # converged maps differ by about 1.5e-6 (the FP64 map itself is only fixed
    # to that level by a 1e-8 residual)
    assert mx.report["true_res_rel"] <= 1e-8
    assert _rel(mx.x, e.x) <= 3e-6


@pytest.mark.parametrize("precond,zn", [(1, 2), (2, 4)])
def test_twolvl_lambda1_iterates(gpu, precond, zn):
    """White noise (lambda 1): the two-level products call stbmmProd, which
    still removes the block means, while the PCG's own N^-1 (apply_weights) does
    not (pcg_true.c:578-600)."""
    wl = W.config_small(seed=11, ndet=2, nsamp=4096, nx=16, ny=16, lam=1)
    print(f"case: lambda 1, precond {precond} Zn {zn}, x_3", flush=True)
    r = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-30,
                    maxiter=50, stop_after=3, precond=precond, z_2lvl=zn)
    with gpu.Problem(wl.pix, wl.weights, wl.block_sizes, wl.lam, wl.invtt,
                     opts=gpu.options(precond=precond, z_2lvl=zn)) as pb:
        sig, _, _ = pb.precond_build(wl.signal)
        x, hist, rep = pb.pcg(sig, wl.noise, tol=1e-30, maxiter=50, stop_after=3)
    assert rep.iterations == r.iterations == 3
    print(f"x_3 rel diff {_rel(x, r.x):.2e}", flush=True)
    assert _rel(x, r.x) <= 1e-10
    # block Jacobi is exact for white noise: the residuals after the first step are rounding
    np.testing.assert_allclose(hist, r.res_hist, rtol=1e-9, atol=1e-12)


def test_twolvl_rejected_where_unsupported(gpu):
    wl = W.config_small(seed=11, ndet=2, nsamp=4096, nx=16, ny=16, lam=64)
    item = dict(pix=wl.pix, weights=wl.weights, signal=wl.signal, noise=wl.noise, block_sizes=wl.block_sizes,
                lam=wl.lam, invtt=wl.invtt)
    with pytest.raises(gpu.GPUError, match="precond"):
        gpu.pcg_batch([item], 16 * 16, opts=gpu.options(precond=2, z_2lvl=2))
    with pytest.raises(gpu.GPUError, match="z_2lvl"):
        gpu.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                    maxiter=20, opts=gpu.options(precond=2, z_2lvl=1000))
