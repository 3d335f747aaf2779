"""Template marginalisation on the GPU (MTmap / PCG_GLS_templates) against the
oracle's restatement, through the C-ABI (mappraiser_gpu_templates_* and MTmap).

Bars: T tau and the kernels B_b bit-exact (same per-sample class order, same
sample order in every kernel entry); T^T d within 1e-13 (fixed-order workgroup
sums instead of one running sum); the pseudo-inverses (rocSOLVER's symmetric
eigensolver against the oracle's Jacobi SVD, both cutting at 1e-12 s_max) with
the same ranks and within 1e-8; the PCG residual history within 1e-9 and the
map within 1e-6 once the I monopole, which the offset-like templates leave
undetermined, is removed.
"""
import dataclasses
import os

import numpy as np
import pytest

from midapack_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _otm(wl, **over):
    md = dict(wl.meta)
    md.update(over)
    return O.Templates(wl.block_sizes, md["nces"], md["ndet"], wl.invtt, npoly=md["npoly"], ground=md["ground"],
                       n_sss_bins=md["n_sss_bins"], nhwp=md["nhwp"], delta_t=md["delta_t"],
                       sampling_freq=md["sampling_freq"], sweeptstamps=md["sweeptstamps"], az=md["az"],
                       az_min=md["az_min"], az_max=md["az_max"], hwp_angle=md["hwp_angle"])


def _problem(gpu, wl, **over):
    md = dict(wl.meta)
    md.update(over)
    pb = gpu.Problem(wl.pix, wl.weights, wl.block_sizes, 1, wl.invtt)
    pb.set_templates(md)
    return pb


CASES = [dict(), dict(npoly=1, ground=0, nhwp=0), dict(npoly=0, ground=1, nhwp=0), dict(npoly=0, ground=0, nhwp=3),
         dict(npoly=5, ground=1, nhwp=1)]


@pytest.mark.parametrize("over", CASES, ids=lambda o: ",".join(f"{k}={v}" for k, v in o.items()) or "default")
def test_products_and_kernels_match_the_oracle(gpu, over):
    wl = W.config_templates_small()
    print(f"case: templates {over}", flush=True)
    tm = _otm(wl, **over)
    with _problem(gpu, wl, **over) as pb:
        ntot, nt, ranks = pb.templates_info()
        assert ntot == tm.ntot and np.array_equal(nt, tm.nt)
        rng = np.random.default_rng(1)
        tau = rng.standard_normal(ntot)
        assert np.array_equal(pb.templates_apply(gpu.MGPU_TMPL_T, tau), tm.tvec(tau))
        d = rng.standard_normal(wl.m)
        got, want = pb.templates_apply(gpu.MGPU_TMPL_TT, d), tm.trtvec(d)
        assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
        B, Bi = pb.templates_kernel()
        oB = tm.kernel()
        assert np.array_equal(B, oB)
        oBi, oranks = tm.invkernel(oB)
        assert np.array_equal(ranks, oranks)
        for g, o in zip(tm.blocks(Bi), tm.blocks(oBi)):
            assert np.abs(g - o).max() <= 1e-8 * np.abs(o).max()
        # the filter on a weighted TOD
        w = np.repeat(wl.invtt, wl.block_sizes)
        y = w * d
        got, want = pb.templates_apply(gpu.MGPU_TMPL_FILTER, y), tm.filter(oBi, y)
        assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()
        # templates are filtered out
        pure = w * tm.tvec(tau)
        assert np.abs(pb.templates_apply(gpu.MGPU_TMPL_FILTER, pure)).max() <= 1e-9 * np.abs(pure).max()


def test_pseudo_inverses_with_detector_weights_of_zero(gpu):
    """The kernels of a CES's detectors differ only by their weight, so one
    eigensolve per CES serves them all (scaled by the weight ratio).  A detector
    of weight 0 has B_b = 0: pseudo-inverse 0, rank 0; a CES whose detectors all
    weigh 0 likewise; the others still match the oracle's per-block SVD."""
    wl = W.config_templates_small(nces=3, ndet=3)
    inv = wl.invtt.copy()
    inv[0] = 0.0          # the first detector of CES 0 (the representative moves to the next)
    inv[3:6] = 0.0        # every detector of CES 1
    wl = dataclasses.replace(wl, invtt=inv)
    print("case: templates, detector weights of zero", flush=True)
    tm = _otm(wl)
    with _problem(gpu, wl) as pb:
        ntot, nt, ranks = pb.templates_info()
        B, Bi = pb.templates_kernel()
    oB = tm.kernel()
    assert np.array_equal(B, oB)
    oBi, oranks = tm.invkernel(oB)
    assert np.array_equal(ranks, oranks)
    assert ranks[0] == 0 and all(r == 0 for r in ranks[3:6]) and all(r > 0 for r in ranks[6:])
    for k, (g, o) in enumerate(zip(tm.blocks(Bi), tm.blocks(oBi))):
        if not np.any(o):
            assert not np.any(g), f"block {k}: nonzero pseudo-inverse of a zero kernel"
        else:
            assert np.abs(g - o).max() <= 1e-8 * np.abs(o).max(), f"block {k}"


@pytest.mark.parametrize("precond,zn", [(0, 0), (1, 2)])
def test_template_pcg_matches_the_oracle(gpu, precond, zn):
    wl = W.config_templates_small(flag_frac=0.01)
    print(f"case: template PCG precond {precond} Zn {zn}", flush=True)
    tm = _otm(wl)
    r = O.gls_solve_tmpl(wl.pix, wl.weights, wl.signal, wl.noise, tm, wl.invtt, tol=1e-10, maxiter=300,
                         precond=precond, z_2lvl=zn)
    opts = gpu.options(precond=precond, z_2lvl=zn)
    pb = gpu.Problem(wl.pix, wl.weights, wl.block_sizes, 1, wl.invtt, opts=opts)
    with pb:
        pb.set_templates(wl.meta)
        sig, cond, hits = pb.precond_build(wl.signal)
        assert np.array_equal(sig, r.signal) and np.array_equal(cond, r.cond) and np.array_equal(hits, r.hits)
        x, hist, rep = pb.pcg(sig, wl.noise, tol=1e-10, maxiter=300)
    assert rep.converged and r.converged
    print(f"iterations gpu {rep.iterations} oracle {r.iterations}", flush=True)
    assert abs(rep.iterations - r.iterations) <= 1
    k = min(hist.size, r.res_hist.size)
    assert np.abs(hist[:k] - r.res_hist[:k]).max() <= 1e-9 * max(1.0, np.abs(r.res_hist[:k]).max())
    # the monopole of I (and modes nearly degenerate with the templates) is set by
    # the rounding of the null space: compare after removing the I mean
    dx = x - r.x
    dx[0::3] -= dx[0::3].mean()
    print(f"map diff (I mean removed) {np.abs(dx).max():.2e}", flush=True)
    assert np.abs(dx).max() <= 1e-6 * np.abs(r.x).max()


def test_template_errors(gpu):
    wl = W.config_templates_small()
    with gpu.Problem(wl.pix, wl.weights, wl.block_sizes, 1, wl.invtt, opts=gpu.options(precision=gpu.MGPU_MIXED)) as pb:
        with pytest.raises(gpu.GPUError, match="FP64"):
            pb.set_templates(wl.meta)
    lam_wl = W.config_small()
    with gpu.Problem(lam_wl.pix, lam_wl.weights, lam_wl.block_sizes, lam_wl.lam, lam_wl.invtt) as pb:
        md = dict(wl.meta)
        md["nces"] = 1
        with pytest.raises(gpu.GPUError, match="lambda"):
            pb.set_templates(md)
    md = dict(wl.meta)
    md["sweeptstamps"] = [s[:-1] for s in md["sweeptstamps"]]  # last stamp short of the CES length
    with gpu.Problem(wl.pix, wl.weights, wl.block_sizes, 1, wl.invtt) as pb:
        with pytest.raises(gpu.GPUError, match="CES length"):
            pb.set_templates(md)


def test_mtmap_writes_the_reference_outputs(gpu, tmp_path):
    wl = W.config_templates_small(nx=64, ny=48, ces_len=8192, flag_frac=0.01, npoly=1)
    nside = 16  # 3072 = 64 x 48 raster pixels, used as HEALPix indices
    print("case: MTmap 2x2x8192 on 64x48, nside 16", flush=True)
    pix, w, sig = wl.pix.copy(), wl.weights.copy(), wl.signal.copy()
    gpu.MTmap(str(tmp_path), "t", 0, 0, 1e-8, 300, nside, wl.meta, wl.block_sizes, 3, pix, w, sig, wl.noise,
              wl.invtt)
    assert gpu.last_error() == "", gpu.last_error()
    tm = _otm(wl)
    r = O.gls_solve_tmpl(wl.pix, wl.weights, wl.signal, wl.noise, tm, wl.invtt, tol=1e-8, maxiter=300)
    # the reference's side effects on the caller's arrays
    assert np.array_equal(pix, r.pix_local) and np.array_equal(w, r.weights) and np.array_equal(sig, r.signal)
    hist = np.fromfile(os.path.join(tmp_path, "pcg_residuals_t.dat"), dtype=np.float64)
    assert hist[0] == 1.0 and abs(hist.size - r.res_hist.size) <= 1
    I, cards = gpu.read_map(os.path.join(tmp_path, "mapI_t.fits"))
    assert int(cards["NSIDE"]) == nside
    hits, _ = gpu.read_map(os.path.join(tmp_path, "Hits_t.fits"))
    pixels = r.lindices[3 * r.trash_pix::3] // 3
    assert np.array_equal(hits[pixels].astype(np.int64), r.hits.astype(np.int64))
    dI = I[pixels] - r.x[0::3]
    dI -= dI.mean()
    print(f"map I vs oracle (mean removed) {np.abs(dI).max():.2e}", flush=True)
    assert np.abs(dI).max() <= 1e-5 * np.abs(r.x[0::3]).max()


def test_irregular_sweeps_and_ces_lengths(gpu):
    """CES of different lengths, sweeps starting after the CES start and a
    one-sample sweep (npoly 1: order 0 only)."""
    iw = W.templates_irregular(W.config_templates_small(flag_frac=0.01))
    print("case: irregular sweeps, CES lengths 2048 / 1500", flush=True)
    tm = _otm(iw, npoly=1)
    md = dict(iw.meta)
    md["npoly"] = 1
    with gpu.Problem(iw.pix, iw.weights, iw.block_sizes, 1, iw.invtt) as pb:
        pb.set_templates(md)
        ntot, nt, ranks = pb.templates_info()
        assert ntot == tm.ntot and np.array_equal(nt, tm.nt)
        rng = np.random.default_rng(4)
        tau = rng.standard_normal(ntot)
        assert np.array_equal(pb.templates_apply(gpu.MGPU_TMPL_T, tau), tm.tvec(tau))
        B, _ = pb.templates_kernel()
        assert np.array_equal(B, tm.kernel())
        sig, _, _ = pb.precond_build(iw.signal)
        x, hist, rep = pb.pcg(sig, iw.noise, tol=1e-10, maxiter=300)
    r = O.gls_solve_tmpl(iw.pix, iw.weights, iw.signal, iw.noise, tm, iw.invtt, tol=1e-10, maxiter=300)
    assert rep.converged and r.converged and abs(rep.iterations - r.iterations) <= 1
    k = min(hist.size, r.res_hist.size)
    assert np.abs(hist[:k] - r.res_hist[:k]).max() <= 1e-9
    dx = x - r.x
    dx[0::3] -= dx[0::3].mean()
    assert np.abs(dx).max() <= 1e-6 * np.abs(r.x).max()


def test_template_pcg_graph_replay_is_bit_identical(gpu):
    """The template kernels inside HIP-graph-captured iteration blocks give the
    eager solve's iterates bit for bit."""
    wl = W.config_templates_small(flag_frac=0.01)
    print("case: template PCG eager vs graph", flush=True)
    out = []
    for graph in (0, 1):
        with gpu.Problem(wl.pix, wl.weights, wl.block_sizes, 1, wl.invtt, opts=gpu.options(graph=graph)) as pb:
            pb.set_templates(wl.meta)
            sig, _, _ = pb.precond_build(wl.signal)
            out.append(pb.pcg(sig, wl.noise, tol=1e-10, maxiter=300))
    (x0, h0, r0), (x1, h1, r1) = out
    assert r0.iterations == r1.iterations and np.array_equal(x0, x1) and np.array_equal(h0, h1)
