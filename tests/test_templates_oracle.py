"""Pins of the oracle's template marginalisation (MTmap, mappraiser.c:387-700;
PCG_GLS_templates, pcg_true.c:243-576; algebra/src/templates/templates.c).

The reference ships no test for this path and its templates.c needs LAPACKE
(InvKernel), which this image lacks, so it is not built.  The oracle is pinned
against what the definitions fix, computed independently in numpy:
  * T as a dense matrix: Legendre polynomials from numpy.polynomial.legendre on
    each sweep, one-hot azimuth bins from floor((az - min) / width), HWP
    harmonics from cos(k theta) / sin(k theta) directly (not the recurrence);
  * TVecProd = T tau, TrTVecProd = T^T d, BuildKernel = w T^T T per block;
  * InvKernel: the Penrose conditions and numpy's pinv; the rank equals
    numpy's matrix_rank at the same threshold (the azimuth bins and the sweep
    offsets span a common vector, so the kernels are rank deficient);
  * the filter W - W T B^+ T^T W annihilates every template, and the PCG map
    equals the dense solution of P^T F P x = P^T F d and does not move when
    templates are added to the data.
"""
import numpy as np
import pytest

from midapack_amd import workloads as W
from oracle import oracle as O


def _tm(wl, **over):
    md = dict(wl.meta)
    md.update(over)
    return O.Templates(wl.block_sizes, md["nces"], md["ndet"], wl.invtt, npoly=md["npoly"], ground=md["ground"],
                       n_sss_bins=md["n_sss_bins"], nhwp=md["nhwp"], delta_t=md["delta_t"],
                       sampling_freq=md["sampling_freq"], sweeptstamps=md["sweeptstamps"], az=md["az"],
                       az_min=md["az_min"], az_max=md["az_max"], hwp_angle=md["hwp_angle"])


def dense_T(wl, npoly, ground, nsss, nhwp, delta_t):
    """T (m x ntot) from the definitions, computed without the oracle's code."""
    md = wl.meta
    fs = md["sampling_freq"]
    cols, blocks = [], []
    ntot = 0
    nces, ndet = md["nces"], md["ndet"]
    m = wl.m
    rows = []
    for c in range(nces):
        L = int(wl.block_sizes[c * ndet])
        st = md["sweeptstamps"][c]
        ns = len(st) - 1
        nh = int(np.ceil(L / (delta_t * fs))) if nhwp else 0
        i = np.arange(L)
        for j in range(ndet):
            Tb = np.zeros((L, npoly * ns + ground * nsss + 2 * nhwp * nh))
            for s in range(ns):
                a, b = st[s], st[s + 1]
                if npoly:
                    Tb[a:b, s] = 1.0  # P_0 (its argument is undefined on a one-sample sweep)
                if npoly > 1:
                    u = (2 * (np.arange(a, b) / fs) - a / fs - (b - 1) / fs) / ((b - 1) / fs - a / fs)
                    for o in range(1, npoly):
                        Tb[a:b, o * ns + s] = np.polynomial.legendre.legval(u, [0] * o + [1])
            k0 = npoly * ns
            if ground:
                width = (md["az_max"][c] - md["az_min"][c]) / nsss
                ab = np.minimum(nsss - 1, np.floor((md["az"][c] - md["az_min"][c]) / width).astype(int))
                Tb[i, k0 + ab] = 1.0
                k0 += nsss
            th = md["hwp_angle"][c]
            tb = np.floor(i / (fs * delta_t)).astype(int)
            for h in range(1, nhwp + 1):
                Tb[i, k0 + 2 * (h - 1) * nh + tb] = np.cos(h * th)
                Tb[i, k0 + (2 * (h - 1) + 1) * nh + tb] = np.sin(h * th)
            blocks.append(Tb)
    tot = sum(b.shape[1] for b in blocks)
    T = np.zeros((m, tot))
    r = cc = 0
    for b in blocks:
        T[r: r + b.shape[0], cc: cc + b.shape[1]] = b
        r += b.shape[0]
        cc += b.shape[1]
    return T, blocks


@pytest.fixture(scope="module")
def wl():
    return W.config_templates_small()


@pytest.mark.parametrize("npoly,ground,nhwp", [(3, 1, 2), (1, 0, 0), (0, 1, 0), (0, 0, 3), (4, 1, 1)])
def test_products_and_kernel_match_the_dense_definition(wl, npoly, ground, nhwp):
    md = wl.meta
    tm = _tm(wl, npoly=npoly, ground=ground, nhwp=nhwp)
    T, blocks = dense_T(wl, npoly, ground, md["n_sss_bins"], nhwp, md["delta_t"])
    assert tm.ntot == T.shape[1] and list(tm.nt) == [b.shape[1] for b in blocks]
    rng = np.random.default_rng(5)
    tau = rng.standard_normal(tm.ntot)
    d = rng.standard_normal(wl.m)
    assert np.allclose(tm.tvec(tau), T @ tau, rtol=0, atol=1e-12 * np.abs(T @ tau).max())
    ref = T.T @ d
    assert np.allclose(tm.trtvec(d), ref, rtol=0, atol=1e-11 * np.abs(ref).max())
    Bs = tm.blocks(tm.kernel())
    for b, (Bb, Tb) in enumerate(zip(Bs, blocks)):
        want = wl.invtt[b] * Tb.T @ Tb
        assert np.abs(Bb - want).max() <= 1e-12 * np.abs(want).max()


def test_kernel_entries_accumulate_in_sample_order(wl):
    """B entries are the reference's ordered sums: (w * wt_k) * wt_l added in
    increasing sample order (so a plain numpy sum of the same products in that
    order reproduces them bit for bit)."""
    md = wl.meta
    tm = _tm(wl, npoly=2, ground=1, nhwp=1)
    B0 = tm.blocks(tm.kernel())[0]
    # entry (poly order 1, sweep 3) x (poly order 1, sweep 3): sequential sum
    st = md["sweeptstamps"][0]
    ns = len(st) - 1
    fs = md["sampling_freq"]
    a, b = int(st[3]), int(st[4])
    w = wl.invtt[0]
    acc = 0.0
    for i in range(a, b):
        x, lo, hi = i / fs, a / fs, (b - 1) / fs
        t = (2 * x - lo - hi) / (hi - lo)
        acc += (w * t) * t
    assert B0[1 * ns + 3, 1 * ns + 3] == acc


def test_invkernel_is_the_pseudo_inverse_with_the_rank_deficiency(wl):
    tm = _tm(wl)
    B = tm.kernel()
    Bi, ranks = tm.invkernel(B)
    for Bb, Bib, r in zip(tm.blocks(B), tm.blocks(Bi), ranks):
        n = Bb.shape[0]
        s = np.linalg.svd(Bb, compute_uv=False)
        assert r == np.sum(s > 1e-12 * s[0]) and r < n  # azimuth bins sum to the sweep offsets
        nrm = np.abs(Bb).max()
        assert np.abs(Bb @ Bib @ Bb - Bb).max() <= 1e-9 * nrm
        assert np.abs(Bib @ Bb @ Bib - Bib).max() <= 1e-9 * np.abs(Bib).max()
        assert np.abs(Bib - np.linalg.pinv(Bb, rcond=1e-12)).max() <= 1e-8 * np.abs(Bib).max()


def test_filter_annihilates_templates_and_equals_the_projector(wl):
    tm = _tm(wl)
    B = tm.kernel()
    Bi, _ = tm.invkernel(B)
    T, _ = dense_T(wl, wl.meta["npoly"], wl.meta["ground"], wl.meta["n_sss_bins"], wl.meta["nhwp"],
                   wl.meta["delta_t"])
    w = np.repeat(wl.invtt, wl.block_sizes)
    rng = np.random.default_rng(9)
    tau = rng.standard_normal(tm.ntot)
    pure = w * (T @ tau)
    out = tm.filter(Bi, pure)
    assert np.abs(out).max() <= 1e-9 * np.abs(pure).max()
    d = rng.standard_normal(wl.m)
    F = np.diag(w) - (w[:, None] * T) @ np.linalg.pinv(T.T @ (w[:, None] * T), rcond=1e-12) @ (T.T * w[None, :])
    got = tm.filter(Bi, w * d)
    want = F @ d
    assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()


def test_bin_az_and_hwp_mod():
    az = np.array([0.0, 0.49, 0.5, 3.99, 4.0, 4.0 + 1e-9])
    out = np.zeros(az.size, np.int32)
    assert O.lib().orc_bin_az(az.size, az, 0.0, 4.0, 8, out) == 0
    assert list(out) == [0, 0, 1, 7, 7, 7]  # the top edge folds into the last bin
    assert O.lib().orc_bin_az(1, np.array([-0.1]), 0.0, 4.0, 8, out) != 0


def test_template_solve_matches_dense_and_ignores_templates(wl):
    md = wl.meta
    tm = _tm(wl)
    r = O.gls_solve_tmpl(wl.pix, wl.weights, wl.signal, wl.noise, tm, wl.invtt, tol=1e-12, maxiter=400)
    assert r.converged and r.n_cut > 0
    # dense GLS with the filter, on the matrix and data left by the pixel cut
    # (cut samples keep their place in the time domain with a zero P row)
    T, _ = dense_T(wl, md["npoly"], md["ground"], md["n_sss_bins"], md["nhwp"], md["delta_t"])
    w = np.repeat(wl.invtt, wl.block_sizes)
    F = np.diag(w) - (w[:, None] * T) @ np.linalg.pinv(T.T @ (w[:, None] * T), rcond=1e-12) @ (T.T * w[None, :])
    off = 3 * r.trash_pix
    P = np.zeros((wl.m, r.x.size))
    for t in range(wl.m):
        for k in range(3):
            li = int(r.pix_local[3 * t + k])
            if li >= off:
                P[t, li - off] = r.weights[3 * t + k]
    d = r.signal + wl.noise
    A = P.T @ F @ P
    x = np.linalg.lstsq(A, P.T @ F @ d, rcond=None)[0]
    # the map solves the normal equations (it is defined up to the null space
    # of A: sky modes the templates absorb)
    assert np.abs(A @ r.x - P.T @ F @ d).max() <= 1e-8 * np.abs(P.T @ F @ d).max()
    # adding templates to the data leaves the map unchanged (added to the noise
    # stream: the pixel cut zeroes the signal of cut samples, templates included)
    rng = np.random.default_rng(3)
    tau = rng.standard_normal(tm.ntot)
    r2 = O.gls_solve_tmpl(wl.pix, wl.weights, wl.signal, wl.noise + tm.tvec(tau), tm, wl.invtt, tol=1e-12,
                          maxiter=400)
    # ... along every well-determined map mode (the offset-like templates leave
    # the monopole and nearly degenerate modes to the starting point and rounding)
    lam, V = np.linalg.eigh(0.5 * (A + A.T))
    keep = lam > 1e-3 * lam.max()
    dx = V[:, keep].T @ (r2.x - r.x)
    assert np.abs(dx).max() <= 1e-7 * np.abs(V[:, keep].T @ r.x).max()
    assert np.abs(V[:, keep].T @ (x - r.x)).max() <= 1e-7 * np.abs(V[:, keep].T @ r.x).max()
    print("iterations", r.iterations, "ranks", list(r.ranks), "of", list(tm.nt), "modes kept", keep.sum(),
          "of", keep.size)


def test_irregular_sweeps_and_ces_lengths(wl):
    iw = W.templates_irregular(wl)
    md = iw.meta
    npoly = 1  # order 0 only: a one-sample sweep has no Legendre argument (b == a)
    tm = _tm(iw, npoly=npoly)
    T, blocks = dense_T(iw, npoly, md["ground"], md["n_sss_bins"], md["nhwp"], md["delta_t"])
    assert T[: md["sweeptstamps"][0][0], : md["sweeptstamps"][0].size - 1].sum() == 0
    rng = np.random.default_rng(2)
    tau = rng.standard_normal(tm.ntot)
    d = rng.standard_normal(iw.m)
    assert np.allclose(tm.tvec(tau), T @ tau, rtol=0, atol=1e-12 * np.abs(T @ tau).max())
    ref = T.T @ d
    assert np.allclose(tm.trtvec(d), ref, rtol=0, atol=1e-11 * np.abs(ref).max())
    for b, (Bb, Tb) in enumerate(zip(tm.blocks(tm.kernel()), blocks)):
        want = iw.invtt[b] * Tb.T @ Tb
        assert np.abs(Bb - want).max() <= 1e-12 * np.abs(want).max()
