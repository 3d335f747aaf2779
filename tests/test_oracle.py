"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle function is checked against something it was not derived from:
the reference's own mapmat code compiled from source (committed goldens, and
live when oracle/_ref is built), numpy brute force, closed forms of the
banded Toeplitz product, numpy's FFT, scipy's LAPACK, the adjugate inverse,
and textbook conjugate gradients on a symmetric system.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from midapack_amd import workloads as W
from oracle import oracle as O


# ----------------------------------------------------------------------------
# index tables (mapmat.c:496-521, csort.c:111-135, cindex.c:34-41,102-115)
# ----------------------------------------------------------------------------
def test_local_shape_matches_reference_goldens(fixture_rank1, ref_rank1):
    lind, loc, tr = O.local_shape(fixture_rank1["pix"])
    assert np.array_equal(lind, ref_rank1["lindices"])
    assert np.array_equal(loc, ref_rank1["indices"])
    assert tr == int(ref_rank1["trash"])


def _flagged():
    wl = W.config_small(seed=21, ndet=2, nsamp=3000, nx=12, ny=12, lam=32, flag_frac=0.05)
    return wl.pix, wl.weights


def test_local_shape_flags_matches_reference_goldens(ref_flags):
    pix, _ = _flagged()
    lind, loc, tr = O.local_shape(pix)
    assert tr == 1 and int(ref_flags["trash"]) == 1
    assert np.array_equal(lind, ref_flags["lindices"])
    assert np.array_equal(loc, ref_flags["indices"])
    assert list(lind[:3]) == [-3, -2, -1]


@pytest.mark.parametrize("seed", range(6))
def test_local_shape_bruteforce(seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(-5, 400, size=rng.integers(1, 2000)).astype(np.int32)
    lind, loc, tr = O.local_shape(a)
    u = np.unique(a)
    assert np.array_equal(lind, u)
    assert np.array_equal(loc, np.searchsorted(u, a))
    assert tr == int(u[0] < 0)


# ----------------------------------------------------------------------------
# pointing products (mapmat.c:704-729, 819-851)
# ----------------------------------------------------------------------------
def test_products_match_reference_goldens(fixture_rank1, ref_rank1):
    lind, loc, tr = O.local_shape(fixture_rank1["pix"])
    y = O.matvec(loc, fixture_rank1["weights"], tr, 3, ref_rank1["x"])
    assert np.array_equal(y, ref_rank1["Px"])
    x = O.trmatvec(loc, fixture_rank1["weights"], tr, 3, lind.size, ref_rank1["y"])
    assert np.array_equal(x, ref_rank1["Pty"])


def test_products_flags_match_reference_goldens(ref_flags):
    pix, w = _flagged()
    lind, loc, tr = O.local_shape(pix)
    assert np.array_equal(O.matvec(loc, w, tr, 3, ref_flags["x"]), ref_flags["Px"])
    assert np.array_equal(O.trmatvec(loc, w, tr, 3, lind.size, ref_flags["y"]), ref_flags["Pty"])


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs /root/reference)")
@pytest.mark.parametrize("flag_frac", [0.0, 0.1])
def test_products_vs_live_reference(flag_frac):
    wl = W.config_small(seed=5, ndet=3, nsamp=2000, nx=10, ny=10, lam=16, flag_frac=flag_frac)
    lind, loc, tr = O.local_shape(wl.pix)
    rl, rloc, rtr = O.ref_local_shape(wl.pix, wl.weights, 3)
    assert np.array_equal(lind, rl) and np.array_equal(loc, rloc) and tr == rtr
    n = lind.size - 3 * tr
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal(n), rng.standard_normal(wl.m)
    ry, _ = O.ref_products(wl.pix, wl.weights, 3, x=x)
    _, rx = O.ref_products(wl.pix, wl.weights, 3, y=y)
    assert np.array_equal(O.matvec(loc, wl.weights, tr, 3, x), ry)
    assert np.array_equal(O.trmatvec(loc, wl.weights, tr, 3, lind.size, y), rx)


def _dense_P(loc, w, tr, nnz, n):
    m = loc.size // nnz
    P = np.zeros((m, n))
    for t in range(m):
        if tr and loc[t * nnz] == 0:
            continue
        for k in range(nnz):
            P[t, loc[t * nnz + k] - nnz * tr] += w[t * nnz + k]
    return P


@pytest.mark.parametrize("flag_frac", [0.0, 0.2])
def test_products_dense_bruteforce_and_adjoint(flag_frac):
    wl = W.config_small(seed=9, ndet=2, nsamp=500, nx=6, ny=6, lam=8, flag_frac=flag_frac)
    lind, loc, tr = O.local_shape(wl.pix)
    n = lind.size - 3 * tr
    P = _dense_P(loc, wl.weights, tr, 3, n)
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(n), rng.standard_normal(wl.m)
    Px = O.matvec(loc, wl.weights, tr, 3, x)
    Pty = O.trmatvec(loc, wl.weights, tr, 3, lind.size, y)
    np.testing.assert_allclose(Px, P @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(Pty, P.T @ y, rtol=1e-12, atol=1e-12)
    assert abs(Px @ y - x @ Pty) <= 1e-12 * (np.abs(Px) @ np.abs(y))


# ----------------------------------------------------------------------------
# FFT used by the Toeplitz restatement
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 64, 100, 243, 1000, 4096, 4000])
def test_fft_vs_numpy(n):
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = O.fft(z, -1)
    ref = np.fft.fft(z)
    assert np.max(np.abs(f - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    b = O.fft(z, +1)
    assert np.max(np.abs(b - n * np.fft.ifft(z))) <= 1e-12 * max(1.0, np.max(np.abs(b)))


# ----------------------------------------------------------------------------
# Toeplitz: N^-1 = per-block mean removal + banded product with zero boundaries
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("n,lam,bs", [(65536, 1024, 4096), (131072, 8192, 32768), (24000, 8192, 32768),
                                      (3000, 1000, 4000), (100, 3, 16), (40, 16, 64)])
def test_define_blocksize(n, lam, bs):
    assert O.define_blocksize(n, lam) == bs


def _dense_ninv(block_sizes, lam, invtt, m):
    """Closed form: per block T_b (I - 11^T/n) with T_b the banded symmetric
    Toeplitz matrix of the block (identity product for n < lam-1), identity
    on samples after the last block."""
    N = np.eye(m)
    t0 = 0
    for b, n in enumerate(block_sizes):
        T = invtt[b * lam:(b + 1) * lam]
        C = np.eye(n) - np.ones((n, n)) / n
        if lam == 1:
            Tb, C = T[0] * np.eye(n), np.eye(n)
        elif n < lam - 1:
            Tb = np.eye(n)
        else:
            col = np.zeros(n)
            col[:min(lam, n)] = T[:min(lam, n)]
            Tb = sla.toeplitz(col)
        N[t0:t0 + n, t0:t0 + n] = Tb @ C
        t0 += n
    return N


@pytest.mark.parametrize("case", [
    dict(blocks=[300, 137, 64], lam=16, tail=5),
    dict(blocks=[500], lam=100, tail=0),            # bs = 512 > 2(lam-1), several tiles
    dict(blocks=[250, 90], lam=64, tail=3),         # 4 lam > n: bs = 4 lam (non power of two)
    dict(blocks=[200, 20, 180], lam=40, tail=0),    # middle block shorter than lam-1
    dict(blocks=[128, 128], lam=1, tail=7),         # white noise: diagonal, no mean removal
])
def test_ninv_fft_equals_closed_form(case):
    rng = np.random.default_rng(len(case["blocks"]) * 7 + case["lam"])
    blocks = np.array(case["blocks"], np.int32)
    lam = case["lam"]
    m = int(blocks.sum()) + case["tail"]
    invtt = rng.standard_normal(blocks.size * lam)
    v = rng.standard_normal(m) + 3.0
    ref = _dense_ninv(blocks, lam, invtt, m) @ v
    for use_fft in (True, False):
        out = O.apply_weights(v, blocks, lam, invtt, use_fft=use_fft)
        assert np.max(np.abs(out - ref)) <= 1e-12 * np.max(np.abs(ref)), use_fft


@pytest.mark.parametrize("mode", [O.OFFSET_SYMMETRIC, O.OFFSET_NONE])
def test_ninv_offset_modes_closed_form(mode):
    rng = np.random.default_rng(17)
    blocks = np.array([400, 130, 12, 300], np.int32)
    lam, m = 40, 850
    invtt = rng.standard_normal(blocks.size * lam)
    v = rng.standard_normal(m) + 2.0
    N = np.eye(m)
    t0 = 0
    for b, n in enumerate(blocks):
        C = np.eye(n) - np.ones((n, n)) / n
        if n < lam - 1:
            Tb = np.eye(n)
        else:
            col = np.zeros(n)
            col[:lam] = invtt[b * lam:(b + 1) * lam]
            Tb = sla.toeplitz(col)
        N[t0:t0 + n, t0:t0 + n] = C @ Tb @ C if mode == O.OFFSET_SYMMETRIC else Tb
        t0 += n
    ref = N @ v
    with O.offset_mode(mode):
        out = O.apply_weights(v, blocks, lam, invtt)
    assert np.max(np.abs(out - ref)) <= 1e-12 * np.max(np.abs(ref))
    if mode == O.OFFSET_SYMMETRIC:
        assert np.allclose(N, N.T, atol=1e-12)


def test_symmetric_offset_converges_where_reference_stalls():
    """The one-sided mean removal T(I - 11^T/n) of the reference is not symmetric
    and leaves PCG at a residual floor on a short-band workload; the symmetric
    projection C T C is symmetric positive semi-definite and consistent."""
    wl = W.config_small(seed=7, ndet=2, nsamp=8192, nx=16, ny=16, lam=512)
    ref = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                      maxiter=200)
    with O.offset_mode(O.OFFSET_SYMMETRIC):
        sym = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                          maxiter=200)
    assert not ref.converged and ref.res_hist.min() > 1e-6
    assert sym.converged


def test_ninv_lambda1_is_exact_diagonal():
    rng = np.random.default_rng(0)
    blocks = np.array([10, 20], np.int32)
    inv = np.array([2.5, -0.5])
    v = rng.standard_normal(33)
    out = O.apply_weights(v, blocks, 1, inv)
    assert np.array_equal(out[:10], 2.5 * v[:10]) and np.array_equal(out[10:30], -0.5 * v[10:30])
    assert np.array_equal(out[30:], v[30:])


def test_ninv_delta_and_constant():
    rng = np.random.default_rng(1)
    lam, n = 50, 1000
    T = np.zeros(lam)
    T[0] = 3.0
    v = rng.standard_normal(n)
    out = O.apply_weights(v, np.array([n], np.int32), lam, T)
    np.testing.assert_allclose(out, 3.0 * (v - v.mean()), rtol=0, atol=1e-13)
    c = O.apply_weights(np.full(n, 7.0), np.array([n], np.int32), lam, rng.standard_normal(lam))
    assert np.max(np.abs(c)) <= 1e-11


def test_ninv_fixture_fft_vs_direct(fixture_rank1):
    f = fixture_rank1
    v = np.random.default_rng(2).standard_normal(f["signal"].size)
    b = O.apply_weights(v, f["block_sizes"], f["lam"], f["invtt"], use_fft=False)
    with O.ninv_mode(O.NINV_REAL_SPECTRUM):
        a = O.apply_weights(v, f["block_sizes"], f["lam"], f["invtt"], use_fft=True)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
    # the reference's in-place complex multiply (toeplitz.c:587-588) feeds the
    # rounding noise of Im FFT(T_circ) through Re FFT(T_circ): with T0 ~ 1e7 it
    # moves the FFT path off the exact product by ~1e-10..1e-9 relative
    q = O.apply_weights(v, f["block_sizes"], f["lam"], f["invtt"], use_fft=True)
    dev = np.max(np.abs(q - b)) / np.max(np.abs(b))
    assert 1e-12 < dev < 1e-8


# ----------------------------------------------------------------------------
# 3x3 LAPACK restatement vs scipy's LAPACK
# ----------------------------------------------------------------------------
def test_dlange_exact(lapack_golden):
    for A, an in zip(lapack_golden["A"], lapack_golden["anorm"]):
        assert O.dlange1(A) == an


def test_dgetrf_vs_lapack(lapack_golden):
    for A, lu, piv, info, rc in zip(lapack_golden["A"], lapack_golden["lu"], lapack_golden["ipiv"],
                                    lapack_golden["info"], lapack_golden["rcond"]):
        olu, opiv, oinfo = O.dgetrf(A)
        if rc < 1e-8:
            # numerically singular (rank-deficient map blocks): whether the last
            # pivot rounds to exactly 0 depends on the BLAS (scipy's uses FMA);
            # either way rcond is ~0 and the pixel is cut
            assert O.dgecon1(olu, O.dlange1(A)) < 1e-6
            continue
        assert oinfo == info
        col = np.abs(A)
        ties = any(np.sum(col[:, j] == col[:, j].max()) > 1 for j in range(3))
        if not ties:
            assert np.array_equal(opiv, piv)
            scale = max(np.max(np.abs(lu)), 1e-300)
            assert np.max(np.abs(olu - lu)) <= 1e-12 * scale * max(1.0, np.linalg.cond(A) if info == 0 else 1.0)


def test_dgecon_vs_lapack(lapack_golden):
    rel = []
    for A, an, rc in zip(lapack_golden["A"], lapack_golden["anorm"], lapack_golden["rcond"]):
        lu, _, _ = O.dgetrf(A)
        orc = O.dgecon1(lu, an)
        if rc > 1e-8:
            rel.append(abs(orc - rc) / rc)
            # the cut decision at 0.1 (precond.c:560) agrees away from the threshold
            if abs(rc - 0.1) > 1e-9:
                assert (orc > 0.1) == (rc > 0.1)
        else:
            assert orc <= 1e-6
    assert max(rel) <= 1e-9


def test_dgetri_vs_lapack_and_adjugate(lapack_golden):
    for A, inv, info, rc in zip(lapack_golden["A"], lapack_golden["inv"], lapack_golden["info"],
                                lapack_golden["rcond"]):
        if info != 0 or rc < 1e-6:
            continue
        lu, piv, _ = O.dgetrf(A)
        oinv, oinfo = O.dgetri(lu, piv)
        assert oinfo == 0
        assert np.max(np.abs(oinv - inv)) <= 1e-12 / rc * np.max(np.abs(inv))
        adj = np.array([[np.linalg.det(np.delete(np.delete(A, j, 0), i, 1)) * (-1) ** (i + j) for j in range(3)]
                        for i in range(3)])
        closed = adj / np.linalg.det(A)
        assert np.max(np.abs(oinv - closed)) <= 1e-10 / rc * np.max(np.abs(closed))


def test_dgecon_bounds_true_condition():
    rng = np.random.default_rng(4)
    for _ in range(200):
        A = rng.standard_normal((3, 3))
        lu, _, _ = O.dgetrf(A)
        est = O.dgecon1(lu, O.dlange1(A))
        true = 1.0 / (np.linalg.norm(A, 1) * np.linalg.norm(np.linalg.inv(A), 1))
        # Hager/Higham estimates ||A^-1||_1 from below: rcond is never below the true value
        assert est >= true * (1 - 1e-12) and est <= 3.0 * true


# ----------------------------------------------------------------------------
# getlocalW: closed form B_q = sum_t w w^T T_b(t)[0]
# ----------------------------------------------------------------------------
def test_getlocalW_closed_form():
    wl = W.config_small(seed=3, ndet=3, nsamp=700, nx=8, ny=8, lam=4, block_len=300, flag_frac=0.05)
    m = wl.m
    blocks = wl.block_sizes[:-1]  # leave the last block's samples outside any block
    lind, loc, tr = O.local_shape(wl.pix)
    B, hits = O.getlocalW(loc, wl.weights, tr, 3, lind.size, blocks, wl.lam, wl.invtt)
    npix = (lind.size - 3 * tr) // 3
    ref = np.zeros((npix, 3, 3))
    cnt = np.zeros(npix, int)
    starts = np.concatenate([[0], np.cumsum(blocks)])
    for t in range(m):
        if tr and loc[3 * t] == 0:
            continue
        q = loc[3 * t] // 3 - tr
        w = wl.weights[3 * t:3 * t + 3]
        b = np.searchsorted(starts, t, side="right") - 1
        T0 = wl.invtt[b * wl.lam] if b < blocks.size else 1.0
        ref[q] += np.outer(w, w) * T0
        cnt[q] += 1
    np.testing.assert_allclose(B, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    assert np.array_equal(hits, cnt)


# ----------------------------------------------------------------------------
# PCG (pcg_true.c:20-241)
# ----------------------------------------------------------------------------
def test_pcg_white_noise_reaches_dense_gls_solution():
    wl = W.config_small(seed=4, ndet=2, nsamp=1024, nx=8, ny=8, lam=1)
    r = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, 1, wl.invtt, tol=1e-12, maxiter=200)
    assert r.converged
    n = r.x.size
    P = _dense_P(r.pix_local, r.weights, r.trash_pix, 3, n)
    Ninv = np.diag(np.repeat(wl.invtt, wl.block_sizes))
    d = r.signal + wl.noise
    x = np.linalg.solve(P.T @ Ninv @ P, P.T @ Ninv @ d)
    assert np.max(np.abs(r.x - x)) <= 1e-9 * np.max(np.abs(x))


def test_pcg_iterates_match_textbook_cg_on_symmetric_system():
    """For lambda=1 the GLS operator is symmetric positive definite; Polak-Ribiere
    and Fletcher-Reeves PCG coincide in exact arithmetic, so x_k equals the
    textbook PCG iterate built from dense matrices."""
    wl = W.config_small(seed=8, ndet=2, nsamp=600, nx=6, ny=6, lam=1)
    # a non-uniform diagonal so that the preconditioner is not exact
    inv = np.array([1.0, 3.0])
    r_all = [O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, 1, inv, tol=1e-300, maxiter=50,
                         stop_after=k) for k in (1, 2, 3, 4)]
    r = r_all[0]
    n = r.x.size
    P = _dense_P(r.pix_local, r.weights, r.trash_pix, 3, n)
    Nd = np.diag(np.repeat(inv, wl.block_sizes))
    A = P.T @ Nd @ P
    b = P.T @ Nd @ (r.signal + wl.noise)
    M = np.zeros_like(A)
    for q in range(n // 3):
        s = slice(3 * q, 3 * q + 3)
        M[s, s] = np.linalg.inv(A[s, s])
    x = np.zeros(n)
    g = b.copy()
    z = M @ g
    p = z.copy()
    for k, rk in enumerate(r_all, start=1):
        Ap = A @ p
        a = (g @ z) / (p @ Ap)
        x = x + a * p
        g_new = g - a * Ap
        z_new = M @ g_new
        beta = (g_new @ z_new) / (g @ z)
        p = z_new + beta * p
        g, z = g_new, z_new
        assert np.max(np.abs(rk.x - x)) <= 1e-9 * np.max(np.abs(x)), k


def test_pcg_loop_runs_even_when_absolute_test_passes():
    wl = W.config_small(seed=4, ndet=1, nsamp=512, nx=6, ny=6, lam=1)
    r = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, 1, wl.invtt, tol=1e300, maxiter=10)
    assert r.iterations == 1 and r.converged


def test_pixel_cut_and_rebuild():
    wl = W.config_small(seed=6, ndet=1, nsamp=600, nx=10, ny=10, lam=1)
    pix = wl.pix.copy()
    # three samples pointing once each at new pixels: rank-1 blocks, rcond ~ 0
    for t, p in zip((5, 50, 500), (1000, 1001, 1002)):
        pix[3 * t:3 * t + 3] = [3 * p, 3 * p + 1, 3 * p + 2]
    r = O.gls_solve(pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, 1, wl.invtt, tol=1e-10, maxiter=50)
    assert r.n_cut >= 3 and r.trash_pix == 1
    assert list(r.lindices[:3]) == [-3, -2, -1]
    assert not np.isin(np.array([3000, 3003, 3006]), r.lindices).any()
    for t in (5, 50, 500):
        assert list(r.pix_local[3 * t:3 * t + 3]) == [0, 1, 2]
        assert r.signal[t] == 0.0 and np.all(r.weights[3 * t:3 * t + 3] == 0.0)
    assert np.all(r.cond > 0.1)


def test_c1_shape_converges():
    wl = W.make_workload("t", 2, 65536, 64, 64, 1024, 3)
    r = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-8,
                    maxiter=300)
    assert r.converged and r.iterations < 150
    assert r.res_hist[-1] <= 1e-8


def test_openmp_build_matches_sequential():
    """The all-cores timing build computes the same solve (P^T's atomic adds
    only reorder its sums)."""
    from midapack_amd import workloads as W

    wl = W.config_small(seed=3, ndet=3, nsamp=8192, nx=24, ny=24, lam=128)
    with O.offset_mode(O.OFFSET_SYMMETRIC):
        r = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-10,
                        maxiter=300)
        O.set_threads(4)
        q = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=1e-10,
                        maxiter=300, omp=True)
    assert r.converged and q.converged
    assert np.array_equal(r.lindices, q.lindices) and np.array_equal(r.hits, q.hits)
    assert abs(r.iterations - q.iterations) <= 1
    assert np.max(np.abs(r.x - q.x)) <= 1e-8 * np.max(np.abs(r.x))
