"""GPU parity of the serving workload: seeded requests of 4-128 detectors x
2^16-2^20 samples on 256^2-1024^2 grids with lambda 1024-8192 (the first 8
requests of the serving pool, 246M samples), submitted to the request loop
under Poisson arrivals.

  * FP64 server: every map equals the request's solo solve bit for bit, and
    the smallest request, submitted with maxiter 4 (3 PCG iterations), gives
    x_3 within 4e-12 of the CPU oracle's;
  * mixed-precision server (the serving configuration): every request
    converges with its map within 1e-6 of the FP64 solo map.
"""
import threading
import time

import numpy as np
import pytest

from midapack_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NREQ = 8


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _log(msg):
    print(f"[c5] {msg}", flush=True)


def _item(wl):
    return dict(pix=wl.pix, weights=wl.weights, signal=wl.signal, noise=wl.noise, block_sizes=wl.block_sizes,
                lam=wl.lam, invtt=wl.invtt)


@pytest.fixture(scope="module")
def c5():
    t = time.time()
    pool = W.c5_pool(5, NREQ)
    _log(f"generated {len(pool)} requests, {sum(w.m for w in pool)} samples in {time.time() - t:.1f}s")
    return pool


@pytest.fixture(scope="module")
def c5_solo(gpu, c5):
    opts = gpu.options(precision=gpu.MGPU_FP64, offset_mode=gpu.MGPU_OFFSET_SYMMETRIC)
    out = []
    for w in c5:
        out.append(gpu.pcg_batch([_item(w)], w.nx * w.ny, tol=1e-8, maxiter=3000, opts=opts)[0])
        _log(f"solo FP64 {w.name}: {out[-1].report['iterations']} it")
    return out


def _serve(gpu, pool, opts, rate, seed, maxiter):
    """Submit the pool under Poisson arrivals; returns [(result, latency_ms)], stats."""
    arrivals = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, len(pool)))
    tickets = [None] * len(pool)
    with gpu.Server(hbm_budget_gb=0, max_slots=1, max_batch=8, opts=opts) as srv:
        t0 = time.perf_counter()

        def feed():
            for i, w in enumerate(pool):
                dt = arrivals[i] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                tickets[i] = srv.submit(_item(w), w.nx * w.ny, tol=1e-8, maxiter=maxiter[i])

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        got = []
        for i in range(len(pool)):
            r, lat, _ = srv.wait(tickets[i])
            got.append((r, lat))
        st = srv.stats()
    return got, st


def test_c5_fp64_server_equals_solo_and_oracle(gpu, c5, c5_solo):
    small = min(range(len(c5)), key=lambda i: c5[i].m)
    maxiter = [4 if i == small else 3000 for i in range(len(c5))]
    opts = gpu.options(precision=gpu.MGPU_FP64, offset_mode=gpu.MGPU_OFFSET_SYMMETRIC)
    got, st = _serve(gpu, c5, opts, rate=5.0, seed=51, maxiter=maxiter)
    _log(f"FP64 server: {st}")
    for i, ((r, lat), s) in enumerate(zip(got, c5_solo)):
        if i == small:
            continue
        _log(f"req {i}: {lat:.0f} ms, {r.report['iterations']} it (solo {s.report['iterations']})")
        assert r.report["converged"] and r.report["iterations"] == s.report["iterations"]
        assert np.array_equal(r.pixels, s.pixels)
        assert np.array_equal(r.x, s.x), f"req {i}: max diff {np.max(np.abs(r.x - s.x))}"
    wl = c5[small]
    with O.ninv_mode(O.NINV_REAL_SPECTRUM), O.offset_mode(O.OFFSET_SYMMETRIC):
        o = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-30,
                        maxiter=4)
    r = got[small][0]
    err = _rel(r.x, o.x)
    _log(f"req {small} ({wl.name}) served with maxiter 4: x_3 rel. diff vs oracle {err:.2e}")
    assert r.report["iterations"] == o.iterations == 4
    assert np.array_equal(r.pixels, o.lindices[3 * o.trash_pix::3] // 3)
    assert err <= 4e-12


def test_c5_mixed_server_map_error(gpu, c5, c5_solo):
    opts = gpu.options(precision=gpu.MGPU_MIXED, offset_mode=gpu.MGPU_OFFSET_SYMMETRIC)
    got, st = _serve(gpu, c5, opts, rate=10.0, seed=52, maxiter=[3000] * len(c5))
    _log(f"mixed server: {st}")
    errs = []
    for i, ((r, lat), s) in enumerate(zip(got, c5_solo)):
        e = _rel(r.x, s.x)
        errs.append(e)
        _log(f"req {i}: {lat:.0f} ms, {r.report['iterations']} it, map error vs FP64 {e:.2e}, true residual "
             f"{r.report['true_res_rel']:.2e}")
        assert r.report["converged"]
        assert np.array_equal(r.pixels, s.pixels)
    assert max(errs) <= 1e-6
