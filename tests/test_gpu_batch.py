"""Batched solves and graph-captured PCG iterations against solo solves.

A batch solves independent problems as one segmented problem; with the same
tile length every item must reproduce its solo solve bit for bit (same kernels,
same summation groups).  Graph replay must reproduce the eager launches bit for
bit.  Sizes are small (seconds); every test prints its case before the GPU work.
"""
from __future__ import annotations

import numpy as np
import pytest

from midapack_amd import workloads as W

pytestmark = pytest.mark.gpu


def _item(wl):
    return dict(pix=wl.pix, weights=wl.weights, signal=wl.signal, noise=wl.noise, block_sizes=wl.block_sizes,
                lam=wl.lam, invtt=wl.invtt)


def _solo_pixels(g):
    trash = g.report["trash_pix"]
    return g.lindices[3 * trash::3] // 3


def _batch_workloads():
    # uneven lengths, one item with flagged samples, two lambdas, 64x64 grid
    return [
        W.make_workload("b0", 2, 40000, 64, 64, 256, 11, fknee_range=(0.05, 0.5), alpha_range=(1.0, 2.0)),
        W.make_workload("b1", 3, 21000, 64, 64, 256, 12, block_len=9000, flag_frac=0.01),
        W.make_workload("b2", 1, 70001, 64, 64, 256, 13),
        W.make_workload("b3", 2, 30000, 64, 64, 1024, 14),
    ]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_items_match_solo_bitwise(gpu, precision):
    print(f"case: batch of 4 (lambda 256 x3, 1024 x1), {precision}", flush=True)
    prec = gpu.MGPU_FP64 if precision == "fp64" else gpu.MGPU_MIXED
    wls = _batch_workloads()
    # automatic tile lengths: the block-length profiles give the same choice alone and in the batch
    opts = gpu.options(precision=prec, offset_mode=gpu.MGPU_OFFSET_SYMMETRIC)
    res = gpu.pcg_batch([_item(w) for w in wls], 64 * 64, tol=1e-8, maxiter=300, opts=opts)
    for w, r in zip(wls, res):
        g = gpu.pcg_gls(w.pix, w.weights, w.signal, w.noise, w.block_sizes, w.lam, w.invtt, tol=1e-8,
                        maxiter=300, opts=opts)
        print(f"{w.name}: batch {r.report['iterations']} it, solo {g.report['iterations']} it, "
              f"cut {r.report['n_cut']}/{g.report['n_cut']}", flush=True)
        assert r.report["iterations"] == g.report["iterations"]
        assert r.report["converged"] == g.report["converged"] == 1
        assert r.report["n_cut"] == g.report["n_cut"]
        np.testing.assert_array_equal(r.pixels, _solo_pixels(g))
        np.testing.assert_array_equal(r.hits, g.hits)
        np.testing.assert_array_equal(r.cond, g.cond)
        assert np.array_equal(r.x, g.x), f"{w.name}: max diff {np.max(np.abs(r.x - g.x))}"


def test_batch_rejects_bad_items(gpu):
    wl = W.config_small(ndet=1, nsamp=256, nx=4, ny=4, lam=8)
    it = _item(wl)
    with pytest.raises(gpu.GPUError, match="grid_npix"):
        gpu.pcg_batch([it], 8)  # pixels up to 15
    bad = dict(it, block_sizes=np.array([100], np.int32))
    with pytest.raises(gpu.GPUError, match="sum to"):
        gpu.pcg_batch([bad], 16)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_matches_eager_bitwise(gpu, precision):
    print(f"case: graph-captured PCG blocks vs eager, {precision}", flush=True)
    prec = gpu.MGPU_FP64 if precision == "fp64" else gpu.MGPU_MIXED
    wl = W.make_workload("graph", 4, 1 << 15, 64, 64, 512, 21, fknee_range=(0.05, 0.5), alpha_range=(1.0, 2.0))
    out = {}
    for graph in (0, 1):
        o = gpu.options(precision=prec, graph=graph, offset_mode=gpu.MGPU_OFFSET_SYMMETRIC)
        out[graph] = gpu.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt,
                                 tol=1e-8, maxiter=400, opts=o)
    print(f"eager {out[0].report['iterations']} it, graph {out[1].report['iterations']} it", flush=True)
    assert out[0].report["converged"] and out[1].report["converged"]
    assert out[0].report["iterations"] == out[1].report["iterations"]
    assert out[0].report["refreshes"] == out[1].report["refreshes"]
    assert np.array_equal(out[0].x, out[1].x)


def test_server_results_match_solo_bitwise(gpu):
    """Requests submitted together are admitted as the device frees up, queued
    requests are batched; every result equals its solo solve."""
    print("case: server, 6 requests (2 lambdas), one group at a time, batches of <= 3", flush=True)
    wls = [W.make_workload(f"s{i}", 2 + i % 3, 20000 + 3000 * i, 32, 32, 128 if i % 2 else 256, 40 + i)
           for i in range(6)]
    opts = gpu.options(offset_mode=gpu.MGPU_OFFSET_SYMMETRIC)
    with gpu.Server(hbm_budget_gb=0, max_slots=1, max_batch=3, opts=opts) as srv:
        tickets = [srv.submit(_item(w), 32 * 32, tol=1e-8, maxiter=300) for w in wls]
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    print(f"server stats: {st}", flush=True)
    assert st["admitted"] == 6 and st["groups"] < 6 and st["peak_running"] == 1
    for w, (r, lat, q) in zip(wls, got):
        g = gpu.pcg_gls(w.pix, w.weights, w.signal, w.noise, w.block_sizes, w.lam, w.invtt, tol=1e-8,
                        maxiter=300, opts=opts)
        assert lat >= q >= 0.0
        assert r.report["iterations"] == g.report["iterations"] and r.report["converged"]
        np.testing.assert_array_equal(r.pixels, _solo_pixels(g))
        assert np.array_equal(r.x, g.x), f"{w.name}: max diff {np.max(np.abs(r.x - g.x))}"


def test_server_budget_serializes_solves(gpu):
    # a budget smaller than any request admits one solve at a time (never zero)
    print("case: server with a 1 MB budget", flush=True)
    wls = [W.make_workload(f"t{i}", 2, 16384, 16, 16, 64, 60 + i) for i in range(3)]
    opts = gpu.options(offset_mode=gpu.MGPU_OFFSET_SYMMETRIC)
    with gpu.Server(hbm_budget_gb=1e-3, max_slots=1, max_batch=4, opts=opts) as srv:
        tickets = [srv.submit(_item(w), 256, tol=1e-8, maxiter=300) for w in wls]
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["peak_running"] == 1 and st["groups"] == 3
    assert all(r.report["converged"] for r, _, _ in res)
