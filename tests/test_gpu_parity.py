"""GPU parity: the HIP path, called through the C-ABI, against the reference's
own outputs (committed goldens of the compiled mapmat code) and the pinned CPU
oracle on the same seeded inputs.

Bars: integer tables (lindices, local indices, hits, cut set) and the pointing
products bit-exact; FP64 operators within 1e-12 relative of the exact product;
PCG iterates x_k within 1e-12 relative for small k.
"""
import numpy as np
import pytest

from midapack_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _flagged():
    wl = W.config_small(seed=21, ndet=2, nsamp=3000, nx=12, ny=12, lam=32, flag_frac=0.05)
    return wl.pix, wl.weights


# ----------------------------------------------------------------------------
# index tables and P / P^T against the reference's compiled mapmat (goldens)
# ----------------------------------------------------------------------------
def test_local_shape_and_products_match_reference_fixture(gpu, fixture_rank1, ref_rank1):
    f = fixture_rank1
    print("case: reference data_bin rank 1", flush=True)
    with gpu.Problem(f["pix"], f["weights"], f["block_sizes"], f["lam"], f["invtt"]) as pb:
        lind, idx, tr = pb.local_shape()
        assert np.array_equal(lind, ref_rank1["lindices"])
        assert np.array_equal(idx, ref_rank1["indices"])
        assert tr == int(ref_rank1["trash"])
        assert np.array_equal(pb.matvec(ref_rank1["x"]), ref_rank1["Px"])
        assert np.array_equal(pb.trmatvec(ref_rank1["y"]), ref_rank1["Pty"])


def test_local_shape_and_products_match_reference_flags(gpu, ref_flags):
    pix, w = _flagged()
    print("case: flagged pointing (trash rows)", flush=True)
    with gpu.Problem(pix, w, np.array([pix.size // 3], np.int32), 4, np.ones(4)) as pb:
        lind, idx, tr = pb.local_shape()
        assert tr == 1 and np.array_equal(lind, ref_flags["lindices"]) and np.array_equal(idx, ref_flags["indices"])
        assert np.array_equal(pb.matvec(ref_flags["x"]), ref_flags["Px"])
        assert np.array_equal(pb.trmatvec(ref_flags["y"]), ref_flags["Pty"])


@pytest.mark.parametrize("w0", [1.0, 0.5])
def test_products_bitexact_vs_oracle_c1(gpu, w0):
    wl = W.config_c1()
    w = wl.weights.copy()
    w[0::3] = w0  # non-unit I weight exercises the general sample-major/ELL path
    print(f"case: C1 pointing, I weight {w0}", flush=True)
    lind, loc, tr = O.local_shape(wl.pix)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(lind.size)
    y = rng.standard_normal(wl.m)
    with gpu.Problem(wl.pix, w, wl.block_sizes, wl.lam, wl.invtt) as pb:
        assert np.array_equal(pb.matvec(x), O.matvec(loc, w, tr, 3, x))
        assert np.array_equal(pb.trmatvec(y), O.trmatvec(loc, w, tr, 3, lind.size, y))


# ----------------------------------------------------------------------------
# N^-1: mean removal + banded Toeplitz (exact closed form = oracle direct mode)
# ----------------------------------------------------------------------------
NINV_CASES = [
    dict(name="lambda=1 (diagonal), tail", blocks=[3000, 1000], lam=1, tail=17),
    dict(name="lambda=16, short block, tail", blocks=[5000, 10, 4000], lam=16, tail=9),
    dict(name="lambda=100 (odd band)", blocks=[20000, 7001], lam=100, tail=0),
    dict(name="lambda=1024", blocks=[65536, 30000], lam=1024, tail=0),
    dict(name="lambda=4096", blocks=[40000], lam=4096, tail=0),
    dict(name="lambda=8192, 2^17 blocks", blocks=[131072, 131072], lam=8192, tail=0),
    dict(name="lambda=8192, block shorter than 2 lambda", blocks=[12000, 70000], lam=8192, tail=0),
]


@pytest.mark.parametrize("case", NINV_CASES, ids=[c["name"] for c in NINV_CASES])
def test_ninv_matches_closed_form(gpu, case):
    print(f"case: {case['name']}", flush=True)
    rng = np.random.default_rng(case["lam"])
    blocks = np.array(case["blocks"], np.int32)
    lam = case["lam"]
    m = int(blocks.sum()) + case["tail"]
    # realistic bands: the windowed inverse 1/f noise of each block length
    invtt = np.concatenate([W.invtt_block(max(int(n), 2 * lam), lam, 1.0, 0.2, 1.5) for n in blocks])
    invtt[lam:] *= 1.3  # blocks carry different filters
    pix = np.repeat(3 * (np.arange(m) % 97), 3).astype(np.int32) + np.tile(np.arange(3, dtype=np.int32), m)
    w = np.tile([1.0, 0.3, -0.2], m)
    v = rng.standard_normal(m) + 0.5
    with gpu.Problem(pix, w, blocks, lam, invtt) as pb:
        out = pb.apply_weights(v)
    with O.ninv_mode(O.NINV_DIRECT):
        ref = O.apply_weights(v, blocks, lam, invtt)
    assert _rel(out, ref) <= 1e-12


@pytest.mark.parametrize("mode", ["symmetric", "none"])
@pytest.mark.parametrize("case", [NINV_CASES[1], NINV_CASES[3], NINV_CASES[5], NINV_CASES[6]],
                         ids=[NINV_CASES[i]["name"] for i in (1, 3, 5, 6)])
def test_ninv_offset_modes_match_closed_form(gpu, case, mode):
    print(f"case: {case['name']}, offset {mode}", flush=True)
    gmode = gpu.MGPU_OFFSET_SYMMETRIC if mode == "symmetric" else gpu.MGPU_OFFSET_NONE
    omode = O.OFFSET_SYMMETRIC if mode == "symmetric" else O.OFFSET_NONE
    rng = np.random.default_rng(case["lam"] + 1)
    blocks = np.array(case["blocks"], np.int32)
    lam = case["lam"]
    m = int(blocks.sum()) + case["tail"]
    invtt = np.concatenate([W.invtt_block(max(int(n), 2 * lam), lam, 1.0, 0.3, 1.2) for n in blocks])
    pix = np.repeat(3 * (np.arange(m) % 97), 3).astype(np.int32) + np.tile(np.arange(3, dtype=np.int32), m)
    w = np.tile([1.0, 0.3, -0.2], m)
    v = rng.standard_normal(m) + 0.5
    with gpu.Problem(pix, w, blocks, lam, invtt, opts=gpu.options(offset_mode=gmode)) as pb:
        out = pb.apply_weights(v)
    with O.ninv_mode(O.NINV_DIRECT), O.offset_mode(omode):
        ref = O.apply_weights(v, blocks, lam, invtt)
    assert _rel(out, ref) <= 1e-12


def test_symmetric_offset_pcg_iterates_and_convergence(gpu):
    wl = W.make_workload("c2-shape", 8, 1 << 18, 512, 512, 8192, 2, block_len=1 << 17,
                         fknee_range=(0.05, 0.5), alpha_range=(1.0, 2.0))
    print("case: 8 x 2^18 on 512^2, lambda 8192, 1/f alpha 1-2, symmetric offsets", flush=True)
    opts = gpu.options(offset_mode=gpu.MGPU_OFFSET_SYMMETRIC)
    with O.ninv_mode(O.NINV_REAL_SPECTRUM), O.offset_mode(O.OFFSET_SYMMETRIC):
        r = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-30,
                        maxiter=20, stop_after=5)
    with gpu.Problem(wl.pix, wl.weights, wl.block_sizes, wl.lam, wl.invtt, opts=opts) as pb:
        sig, _, _ = pb.precond_build(wl.signal)
        x5, _, _ = pb.pcg(sig, wl.noise, tol=1e-30, maxiter=20, stop_after=5)
    assert _rel(x5, r.x) <= 1e-11
    g = gpu.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                    maxiter=3000, opts=opts)
    print(f"symmetric offsets: converged={g.report['converged']} in {g.report['iterations']} iterations",
          flush=True)
    assert g.report["converged"]


def test_ninv_reference_fixture(gpu, fixture_rank1):
    f = fixture_rank1
    v = np.random.default_rng(2).standard_normal(f["signal"].size)
    with gpu.Problem(f["pix"], f["weights"], f["block_sizes"], f["lam"], f["invtt"]) as pb:
        out = pb.apply_weights(v)
    with O.ninv_mode(O.NINV_DIRECT):
        exact = O.apply_weights(v, f["block_sizes"], f["lam"], f["invtt"])
    quirk = O.apply_weights(v, f["block_sizes"], f["lam"], f["invtt"])
    print(f"N^-1 on reference data: GPU vs exact {_rel(out, exact):.2e}, reference FFT path vs exact "
          f"{_rel(quirk, exact):.2e}", flush=True)
    assert _rel(out, exact) <= 1e-12


# ----------------------------------------------------------------------------
# preconditioner: cond, hits and the cut set bit-exact
# ----------------------------------------------------------------------------
def _cut_workload():
    wl = W.make_workload("cut", 2, 65536, 64, 64, 1024, 3)
    pix = wl.pix.copy()
    for t, p in zip((5, 5000, 70000, 100000), (9000, 9001, 9002, 9003)):
        pix[3 * t:3 * t + 3] = [3 * p, 3 * p + 1, 3 * p + 2]
    return wl, pix


def test_precond_bitexact_with_cut(gpu):
    wl, pix = _cut_workload()
    print("case: forced single-hit pixels (cut)", flush=True)
    with O.ninv_mode(O.NINV_REAL_SPECTRUM):
        r = O.gls_solve(pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                        maxiter=2)
    with gpu.Problem(pix, wl.weights, wl.block_sizes, wl.lam, wl.invtt) as pb:
        sig, cond, hits = pb.precond_build(wl.signal)
        lind, idx, tr = pb.local_shape()
    assert r.n_cut >= 4
    assert np.array_equal(lind, r.lindices) and np.array_equal(idx, r.pix_local) and tr == r.trash_pix
    assert np.array_equal(hits, r.hits)
    assert np.array_equal(cond, r.cond)
    assert np.array_equal(sig, r.signal)


def test_precond_bitexact_with_flags_and_cut(gpu):
    # flagged samples (trash rows before the cut) plus forced single-hit pixels:
    # the cut follows each cut pixel's own sample chain (the reference offsets
    # that chain's index by the trash flag, a quirk kept out of both paths)
    wl = W.make_workload("cut+flags", 2, 65536, 64, 64, 1024, 4, flag_frac=0.02)
    pix = wl.pix.copy()
    for t, p in zip((7, 6000, 71000, 99000), (9100, 9101, 9102, 9103)):
        pix[3 * t:3 * t + 3] = [3 * p, 3 * p + 1, 3 * p + 2]
    print("case: flagged samples + forced single-hit pixels", flush=True)
    with O.ninv_mode(O.NINV_REAL_SPECTRUM):
        r = O.gls_solve(pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                        maxiter=2)
    with gpu.Problem(pix, wl.weights, wl.block_sizes, wl.lam, wl.invtt) as pb:
        sig, cond, hits = pb.precond_build(wl.signal)
        lind, idx, tr = pb.local_shape()
    assert r.n_cut >= 4 and r.trash_pix == 1
    assert np.array_equal(lind, r.lindices) and np.array_equal(idx, r.pix_local) and tr == r.trash_pix
    assert np.array_equal(hits, r.hits) and np.array_equal(cond, r.cond)
    assert np.array_equal(sig, r.signal)


def test_precond_reference_fixture(gpu, fixture_rank1):
    f = fixture_rank1
    r = O.gls_solve(f["pix"], f["weights"], f["signal"], f["noise"], f["block_sizes"], f["lam"], f["invtt"],
                    tol=1e-8, maxiter=2)
    with gpu.Problem(f["pix"], f["weights"], f["block_sizes"], f["lam"], f["invtt"]) as pb:
        sig, cond, hits = pb.precond_build(f["signal"])
        lind, idx, tr = pb.local_shape()
    print(f"reference data: {r.n_cut} of {r.n_cut + cond.size} pixels cut", flush=True)
    assert np.array_equal(lind, r.lindices) and np.array_equal(idx, r.pix_local)
    assert np.array_equal(cond, r.cond) and np.array_equal(hits, r.hits) and np.array_equal(sig, r.signal)


# ----------------------------------------------------------------------------
# PCG
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("k", [1, 2, 5, 10])
def test_pcg_iterates(gpu, k):
    wl, pix = _cut_workload()
    print(f"case: x_{k}", flush=True)
    with O.ninv_mode(O.NINV_REAL_SPECTRUM):
        r = O.gls_solve(pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-30,
                        maxiter=50, stop_after=k)
    with gpu.Problem(pix, wl.weights, wl.block_sizes, wl.lam, wl.invtt) as pb:
        sig, _, _ = pb.precond_build(wl.signal)
        x, hist, rep = pb.pcg(sig, wl.noise, tol=1e-30, maxiter=50, stop_after=k)
    assert rep.iterations == r.iterations == k
    assert _rel(x, r.x) <= 1e-12 * (1 + k)
    np.testing.assert_allclose(hist, r.res_hist, rtol=1e-10)


def test_pcg_full_solve_and_determinism(gpu):
    wl = W.make_workload("c1-shape", 2, 65536, 64, 64, 1024, 3)
    print("case: 2 x 65536, lambda 1024, tol 1e-8", flush=True)
    with O.ninv_mode(O.NINV_REAL_SPECTRUM):
        r = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                        maxiter=300)
    res = []
    for _ in range(2):
        g = gpu.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                        maxiter=300)
        res.append(g)
    g = res[0]
    assert np.array_equal(res[0].x, res[1].x), "two runs differ"
    assert g.report["converged"] and r.converged
    assert abs(g.report["iterations"] - r.iterations) <= 1
    print(f"iterations GPU {g.report['iterations']} oracle {r.iterations}; map rel diff {_rel(g.x, r.x):.2e}",
          flush=True)
    assert _rel(g.x, r.x) <= 1e-7


def test_pcg_tight_solve_matches_oracle(gpu):
    """The map gap at tol 1e-8 is the stopping tolerance's: solved to 1e-12,
    GPU and oracle maps agree to 1e-9, and each 1e-8 map is as close to the
    tight solution as the oracle's own 1e-8 map."""
    wl = W.make_workload("c1-shape", 2, 65536, 64, 64, 1024, 3)
    print("case: 2 x 65536, lambda 1024, tol 1e-12 and 1e-8", flush=True)
    maps = {}
    for tol in (1e-12, 1e-8):
        with O.ninv_mode(O.NINV_REAL_SPECTRUM):
            r = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=tol,
                            maxiter=2000)
        g = gpu.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=tol,
                        maxiter=2000)
        assert g.report["converged"] and r.converged
        maps[tol] = (g.x, r.x, g.report["iterations"], r.iterations)
    gt, ot, gi, oi = maps[1e-12]
    g8, o8, _, _ = maps[1e-8]
    print(f"tol 1e-12: {gi} / {oi} iterations, GPU vs oracle {_rel(gt, ot):.2e}; tol 1e-8 vs tight: GPU "
          f"{_rel(g8, ot):.2e}, oracle {_rel(o8, ot):.2e}", flush=True)
    assert _rel(gt, ot) <= 1e-9
    assert _rel(g8, ot) <= 2 * _rel(o8, ot) + 1e-10


def test_dropin_side_effects_match_oracle(gpu, fixture_rank1, tmp_path):
    f = fixture_rank1
    with O.ninv_mode(O.NINV_REAL_SPECTRUM):
        r = O.gls_solve(f["pix"], f["weights"], f["signal"], f["noise"], f["block_sizes"], f["lam"], f["invtt"],
                        tol=1e-8, maxiter=100)
    g = gpu.pcg_gls(f["pix"], f["weights"], f["signal"], f["noise"], f["block_sizes"], f["lam"], f["invtt"],
                    tol=1e-8, maxiter=100, outpath=str(tmp_path), ref="fixture")
    assert np.array_equal(g.pix_local, r.pix_local)
    assert np.array_equal(g.weights, r.weights)
    assert np.array_equal(g.signal, r.signal)
    assert np.array_equal(g.lindices, r.lindices)
    assert np.array_equal(g.cond, r.cond) and np.array_equal(g.hits, r.hits)
    assert g.report["converged"] and abs(g.report["iterations"] - r.iterations) <= 1
    assert g.res_hist is not None and g.res_hist.size == g.report["iterations"] + 1
    print(f"reference data: {g.report['iterations']} iterations, map rel diff {_rel(g.x, r.x):.2e}", flush=True)
    assert _rel(g.x, r.x) <= 1e-6


def _variant_cases():
    # (precision, real tile length) of the lambda = 1024 and lambda = 8192 workloads
    from midapack_amd import lib as G
    out = []
    for prec, tag in ((G.MGPU_FP64, "fp64"), (G.MGPU_MIXED, "fp32")):
        for l2 in ((12, 15) if prec == G.MGPU_FP64 else (12, 15, 16)):
            for v in range(1, len(G.tile_variants(prec, l2)) + 1):
                out.append((prec, l2, v, f"{tag}-N2^{l2}-v{v}"))
    return out


VARIANT_CASES = _variant_cases()


@pytest.mark.parametrize("case", [c for c in VARIANT_CASES if c[0] == 0], ids=lambda c: c[3])
def test_ninv_tile_variants_match_closed_form(gpu, case):
    # every FP64 tile kernel variant gives the closed-form product (interior and edge tiles)
    _, l2, v, name = case
    print(f"case: {name}", flush=True)
    lam = 1024 if l2 == 12 else 8192
    blocks = np.array([131072, 50001, 3 * lam], np.int32)
    m = int(blocks.sum())
    rng = np.random.default_rng(l2 * 10 + v)
    invtt = np.concatenate([W.invtt_block(int(n), lam, 1.0, 0.2, 1.5) for n in blocks])
    pix = np.repeat(3 * (np.arange(m) % 97), 3).astype(np.int32) + np.tile(np.arange(3, dtype=np.int32), m)
    w = np.tile([1.0, 0.3, -0.2], m)
    x = rng.standard_normal(m) + 0.5
    opts = gpu.options(tile_log2=l2, tile_variant=v, offset_mode=gpu.MGPU_OFFSET_SYMMETRIC)
    with gpu.Problem(pix, w, blocks, lam, invtt, opts=opts) as pb:
        out = pb.apply_weights(x)
    with O.ninv_mode(O.NINV_DIRECT), O.offset_mode(O.OFFSET_SYMMETRIC):
        ref = O.apply_weights(x, blocks, lam, invtt)
    assert _rel(out, ref) <= 1e-12


@pytest.mark.parametrize("case", [c for c in VARIANT_CASES if c[0] == 1], ids=lambda c: c[3])
def test_mixed_tile_variants_map_error(gpu, case):
    # every FP32 tile kernel variant: mixed-precision map within 1e-6 of the FP64 one
    _, l2, v, name = case
    print(f"case: {name}", flush=True)
    lam = 1024 if l2 == 12 else 8192
    # 2^16-sample tiles need blocks longer than one tile's useful part
    wl = W.make_workload("variant", 4, 1 << 17, 128, 128, lam, 7, block_len=(1 << 17) if l2 == 16 else (1 << 16))
    g64 = gpu.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                      maxiter=400, opts=gpu.options(tile_log2=l2, offset_mode=gpu.MGPU_OFFSET_SYMMETRIC))
    gm = gpu.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                     maxiter=400, opts=gpu.options(precision=gpu.MGPU_MIXED, tile_log2=l2, tile_variant=v,
                                                   offset_mode=gpu.MGPU_OFFSET_SYMMETRIC))
    err = _rel(gm.x, g64.x)
    print(f"{name}: {gm.report['iterations']} iterations (FP64 {g64.report['iterations']}), error {err:.2e}",
          flush=True)
    assert g64.report["converged"] and gm.report["converged"]
    assert gm.report["tile_variant"] == v
    assert err <= 1e-6


def test_tile_variant_default_is_fixed(gpu):
    """Without tile_variant the products use the fixed default of the tile length
    (FP64 results depend on the inputs only); MGPU_TILE_AUTOTUNE picks one of the
    variants by timing."""
    wl = W.make_workload("autotune", 2, 1 << 17, 64, 64, 8192, 3, block_len=1 << 17)
    print("case: default and autotuned tile variant, lambda 8192", flush=True)
    fixed = {gpu.MGPU_FP64: 2, gpu.MGPU_MIXED: 5}
    for prec in (gpu.MGPU_FP64, gpu.MGPU_MIXED):
        g = gpu.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                        maxiter=50, opts=gpu.options(precision=prec, offset_mode=gpu.MGPU_OFFSET_SYMMETRIC))
        assert g.report["tile_variant"] == fixed[prec]
        a = gpu.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                        maxiter=50, opts=gpu.options(precision=prec, offset_mode=gpu.MGPU_OFFSET_SYMMETRIC,
                                                     tile_variant=gpu.MGPU_TILE_AUTOTUNE))
        nv = len(gpu.tile_variants(prec, 15))
        assert 1 <= a.report["tile_variant"] <= nv
        if prec == gpu.MGPU_FP64 and a.report["tile_variant"] == fixed[prec]:
            assert np.array_equal(a.x, g.x)


def test_mixed_precision_map_error(gpu):
    wl = W.config_c1()
    print("case: C1 mixed precision", flush=True)
    with O.ninv_mode(O.NINV_REAL_SPECTRUM):
        r = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                        maxiter=300)
    g = gpu.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                    maxiter=300, opts=gpu.options(precision=gpu.MGPU_MIXED, refresh_every=10))
    err = _rel(g.x, r.x)
    print(f"mixed: {g.report['iterations']} iterations (FP64 oracle {r.iterations}), map error vs FP64 "
          f"{err:.2e}, true residual {g.report['true_res_rel']:.2e}, refreshes {g.report['refreshes']}",
          flush=True)
    assert g.report["converged"]
    assert err <= 1e-6


# ----------------------------------------------------------------------------
# error handling (no kernel ever faults: inputs are rejected on the host)
# ----------------------------------------------------------------------------
def test_invalid_inputs_are_rejected(gpu):
    wl = W.config_small(ndet=1, nsamp=256, nx=4, ny=4, lam=8)
    bad = wl.pix.copy()
    bad[4] += 1  # breaks the 3p+k encoding
    with pytest.raises(gpu.GPUError, match="3p\\+k"):
        gpu.Problem(bad, wl.weights, wl.block_sizes, wl.lam, wl.invtt)
    with pytest.raises(gpu.GPUError, match="exceeds m"):
        gpu.Problem(wl.pix, wl.weights, np.array([300], np.int32), wl.lam, np.ones(8))
    with pytest.raises(gpu.GPUError, match="nnz"):
        gpu.Problem(wl.pix[: 2 * 200], wl.weights[: 2 * 200], np.array([200], np.int32), wl.lam, wl.invtt, nnz=2)
    with gpu.Problem(wl.pix, wl.weights, wl.block_sizes, wl.lam, wl.invtt) as pb:
        with pytest.raises(gpu.GPUError, match="before precond_build"):
            pb.precond_apply(np.zeros(pb.map_size()))


@pytest.mark.parametrize("w0", ["unit", "nonunit"])
def test_dropin_packed_upload_equals_raw_upload(gpu, w0):
    """The drop-in solve uploads the pointing packed (pixel p, Q/U weights)
    when every I weight is 1, and as given otherwise: either way its map is
    bit-identical to the handle path, which always uploads the raw arrays."""
    wl, pix = _cut_workload()
    for t in (17, 4096, 65537, 131000):  # flagged samples
        pix[3 * t:3 * t + 3] = [-3, -2, -1]
    w = wl.weights.copy()
    if w0 == "nonunit":
        w[0::3] = 0.5 + np.random.default_rng(8).random(w.size // 3)
    print(f"case: drop-in vs handle path, I weights {w0}, 4 flagged samples, forced cut", flush=True)
    g = gpu.pcg_gls(pix, w, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8, maxiter=300)
    with gpu.Problem(pix, w, wl.block_sizes, wl.lam, wl.invtt) as pb:
        sig, cond, hits = pb.precond_build(wl.signal)
        x, hist, rep = pb.pcg(sig, wl.noise, tol=1e-8, maxiter=300)
    assert g.report["iterations"] == rep.iterations and g.report["n_cut"] > 0
    assert np.array_equal(g.x, x) and np.array_equal(g.cond, cond) and np.array_equal(g.hits, hits)


def test_dropin_inplace_equals_copies(gpu, fixture_rank1):
    f = fixture_rank1
    print("case: reference data, drop-in in place vs on copies", flush=True)
    g = gpu.pcg_gls(f["pix"], f["weights"], f["signal"], f["noise"], f["block_sizes"], f["lam"], f["invtt"],
                    tol=1e-8, maxiter=100)
    pix, w, sig = f["pix"].copy(), f["weights"].copy(), f["signal"].copy()
    h = gpu.pcg_gls(pix, w, sig, f["noise"], f["block_sizes"], f["lam"], f["invtt"], tol=1e-8, maxiter=100,
                    inplace=True)
    assert np.array_equal(h.x, g.x) and h.report["iterations"] == g.report["iterations"]
    assert np.array_equal(pix, g.pix_local) and np.array_equal(w, g.weights) and np.array_equal(sig, g.signal)
    with pytest.raises(gpu.GPUError, match="inplace"):
        gpu.pcg_gls(f["pix"].astype(np.int64), w, sig, f["noise"], f["block_sizes"], f["lam"], f["invtt"],
                    inplace=True)
