"""ctypes bindings for the CPU oracle (liboracle.so) and the reference mapmat
build (oracle/_ref/libref_mapmat.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` -- as the checker or the timed CPU
baseline, never by the product path.  Every function restates one reference
routine on one rank; see oracle.h for the citations and what pins each one.
"""
from __future__ import annotations

import ctypes as ct
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_REF = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class OracleReport(ct.Structure):
    _fields_ = [
        ("iterations", ct.c_int),
        ("converged", ct.c_int),
        ("res0", ct.c_double),
        ("res", ct.c_double),
        ("n_map", ct.c_int),
        ("lcount", ct.c_int),
        ("trash_pix", ct.c_int),
        ("n_cut", ct.c_int),
    ]


class _Tmpl(ct.Structure):
    _fields_ = [
        ("npoly", ct.c_int),
        ("ground", ct.c_int),
        ("n_sss_bins", ct.c_int),
        ("nhwp", ct.c_int),
        ("delta_t", ct.c_double),
        ("sampling_freq", ct.c_double),
        ("nces", ct.c_int),
        ("ndet", ct.c_int),
        ("block_sizes", ct.c_void_p),
        ("nsweeps", ct.c_void_p),
        ("sweeptstamps", ct.c_void_p),
        ("az_binned", ct.c_void_p),
        ("hwp_cos", ct.c_void_p),
        ("hwp_sin", ct.c_void_p),
        ("w", ct.c_void_p),
    ]


_LIBS = {}


def lib(omp: bool = False):
    """Load liboracle.so (built by ``make -C oracle``), or with omp=True its
    OpenMP build liboracle_omp.so (all-cores CPU timing; same arithmetic except
    P^T's summation order, which its atomic adds leave to the thread schedule)."""
    name = "liboracle_omp.so" if omp else "liboracle.so"
    if name not in _LIBS:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run __graft_entry__.build())")
        L = ct.CDLL(path)
        L.orc_local_shape.argtypes = [ct.c_int64, _i32p, _i32p, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]
        L.orc_matvec.argtypes = [ct.c_int, ct.c_int, _i32p, _f64p, ct.c_int, _f64p, _f64p]
        L.orc_matvec.restype = None
        L.orc_trmatvec.argtypes = [ct.c_int, ct.c_int, _i32p, _f64p, ct.c_int, ct.c_int, _f64p, _f64p]
        L.orc_trmatvec.restype = None
        L.orc_pixel_lists.argtypes = [ct.c_int, ct.c_int, _i32p, ct.c_int, ct.c_int, _i32p, _i32p]
        L.orc_pixel_lists.restype = None
        L.orc_fft.argtypes = [ct.c_int, _f64p, _f64p, ct.c_int]
        L.orc_define_blocksize.argtypes = [ct.c_int, ct.c_int]
        L.orc_set_ninv_mode.argtypes = [ct.c_int]
        L.orc_set_offset_mode.argtypes = [ct.c_int]
        L.orc_stmm_block.argtypes = [_f64p, ct.c_int, _f64p, ct.c_int, ct.c_int]
        L.orc_stbmm.argtypes = [ct.c_int, _i32p, ct.c_int, _f64p, ct.c_int64, _f64p, ct.c_int]
        L.orc_apply_weights.argtypes = [ct.c_int, _i32p, ct.c_int, _f64p, ct.c_int64, _f64p, ct.c_int]
        L.orc_dlange1.argtypes = [ct.c_int, _f64p]
        L.orc_dlange1.restype = ct.c_double
        L.orc_dgetrf.argtypes = [ct.c_int, _f64p, _i32p]
        L.orc_dgecon1.argtypes = [ct.c_int, _f64p, ct.c_double]
        L.orc_dgecon1.restype = ct.c_double
        L.orc_dgetri.argtypes = [ct.c_int, _f64p, _i32p]
        L.orc_getlocalW.argtypes = [ct.c_int, ct.c_int, _i32p, _f64p, ct.c_int, ct.c_int, ct.c_int, _i32p,
                                    ct.c_int, _f64p, _f64p, _i32p]
        L.orc_getlocalW.restype = None
        L.orc_gls_solve.argtypes = [ct.c_int, ct.c_int, _i32p, _f64p, _f64p, _f64p, ct.c_int, _i32p, ct.c_int,
                                    _f64p, ct.c_double, ct.c_int, ct.c_int, _f64p, _i32p, _f64p, _i32p, _f64p,
                                    ct.POINTER(OracleReport)]
        L.orc_gls_solve2.argtypes = [ct.c_int, ct.c_int, _i32p, _f64p, _f64p, _f64p, ct.c_int, _i32p, ct.c_int,
                                     _f64p, ct.c_double, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _f64p, _i32p,
                                     _f64p, _i32p, _f64p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                     ct.POINTER(OracleReport)]
        L.orc_pinv.argtypes = [ct.c_int, _f64p, ct.c_double, _f64p]
        L.orc_bin_az.argtypes = [ct.c_int, _f64p, ct.c_double, ct.c_double, ct.c_int, _i32p]
        L.orc_hwp_mod.argtypes = [ct.c_int, _f64p, _f64p, _f64p]
        L.orc_hwp_mod.restype = None
        _tp = ct.POINTER(_Tmpl)
        L.orc_tmpl_sizes.argtypes = [_tp, ct.c_void_p]
        L.orc_tmpl_tvec.argtypes = [_tp, _f64p, _f64p]
        L.orc_tmpl_trtvec.argtypes = [_tp, _f64p, _f64p]
        L.orc_tmpl_kernel.argtypes = [_tp, _f64p]
        L.orc_tmpl_invkernel.argtypes = [_tp, _f64p, _f64p, _i32p]
        L.orc_tmpl_filter.argtypes = [_tp, _f64p, _f64p]
        L.orc_gls_solve_tmpl.argtypes = [ct.c_int, ct.c_int, _i32p, _f64p, _f64p, _f64p, _i32p, _f64p, ct.c_double,
                                         ct.c_int, ct.c_int, ct.c_int, ct.c_int, _tp, _f64p, _i32p, _f64p, _i32p,
                                         _f64p, _i32p, ct.POINTER(OracleReport)]
        L.orc_symeig.argtypes = [ct.c_int, _f64p, _f64p, _f64p]
        L.orc_set_threads.argtypes = [ct.c_int]
        # a newly loaded build follows the modes set on the others
        if _LIBS:
            other = next(iter(_LIBS.values()))
            L.orc_set_ninv_mode(other.orc_set_ninv_mode(-1))
            L.orc_set_offset_mode(other.orc_set_offset_mode(-1))
        _LIBS[name] = L
    return _LIBS[name]


def set_threads(n: int) -> int:
    """Threads of the OpenMP build; returns the count in use."""
    return int(lib(omp=True).orc_set_threads(int(n)))


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libref_mapmat.so"))


def ref():
    """Load the reference mapmat build (only where /root/reference was present at build time)."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_mapmat.so")
        if not os.path.exists(path):
            raise RuntimeError(f"reference mapmat build missing: {path}")
        R = ct.CDLL(path)
        R.ref_local_shape.argtypes = [ct.c_int, ct.c_int, _i32p, _f64p, _i32p, ct.POINTER(ct.c_int),
                                      ct.POINTER(ct.c_int)]
        _vp = ct.c_void_p
        R.ref_products.argtypes = [ct.c_int, ct.c_int, _i32p, _f64p, _vp, _vp, _vp, _vp]
        _REF = R
    return _REF


def ref_iofiles_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libref_iofiles.so"))


def ref_write_map(data, nside, filename, nest=True, coordsys="C"):
    """The reference's write_map (iofiles.c:19-53, cfitsio) compiled from
    source into oracle/_ref; data float64 -> TDOUBLE, int32 -> TINT."""
    R = ct.CDLL(os.path.join(_HERE, "_ref", "libref_iofiles.so"))
    R.write_map.argtypes = [ct.c_void_p, ct.c_int, ct.c_long, ct.c_char_p, ct.c_char, ct.c_char_p]
    R.write_map.restype = None
    a = np.ascontiguousarray(data)
    typ = 82 if a.dtype == np.float64 else 31
    assert a.dtype in (np.float64, np.int32) and a.size == 12 * nside * nside
    if os.path.exists(filename):
        os.remove(filename)
    R.write_map(a.ctypes.data, typ, int(nside), filename.encode(), bytes([1 if nest else 0]), coordsys.encode())


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ----------------------------------------------------------------------------
# index tables and pointing
# ----------------------------------------------------------------------------
def local_shape(pix):
    """MatLocalShape: returns (lindices, local indices, trash_pix)."""
    idx = _c(pix, np.int32).copy()
    lind = np.empty(max(idx.size, 1), np.int32)
    lc, tr = ct.c_int(0), ct.c_int(0)
    rc = lib().orc_local_shape(idx.size, idx, lind, ct.byref(lc), ct.byref(tr))
    if rc:
        raise RuntimeError(f"orc_local_shape rc={rc}")
    return lind[: lc.value].copy(), idx, int(tr.value)


def matvec(indices, values, trash, nnz, x):
    m = indices.size // nnz
    y = np.empty(m, np.float64)
    lib().orc_matvec(m, nnz, _c(indices, np.int32), _c(values, np.float64), int(trash), _c(x, np.float64), y)
    return y


def trmatvec(indices, values, trash, nnz, lcount, y):
    m = indices.size // nnz
    x = np.empty(max(lcount, 1), np.float64)
    lib().orc_trmatvec(m, nnz, _c(indices, np.int32), _c(values, np.float64), int(trash), int(lcount),
                       _c(y, np.float64), x)
    n = lcount - (nnz if trash else 0)
    return x[:n].copy()


def ref_local_shape(pix, weights, nnz):
    idx = _c(pix, np.int32).copy()
    w = _c(weights, np.float64).copy()
    lind = np.empty(max(idx.size, 1), np.int32)
    lc, tr = ct.c_int(0), ct.c_int(0)
    ref().ref_local_shape(idx.size // nnz, nnz, idx, w, lind, ct.byref(lc), ct.byref(tr))
    return lind[: lc.value].copy(), idx, int(tr.value)


def ref_products(pix, weights, nnz, x=None, y=None):
    """Reference MatVecProd(x) and/or TrMatVecProd(y) on global indices."""
    idx = _c(pix, np.int32)
    w = _c(weights, np.float64).copy()
    m = idx.size // nnz
    lind, _, trash = ref_local_shape(pix, weights, nnz)
    yo = np.empty(m, np.float64) if x is not None else None
    xo = np.empty(max(lind.size, 1), np.float64) if y is not None else None
    xin = _c(x, np.float64) if x is not None else None
    yin = _c(y, np.float64) if y is not None else None
    ref().ref_products(m, nnz, idx, w,
                       xin.ctypes.data if xin is not None else None,
                       yo.ctypes.data if yo is not None else None,
                       yin.ctypes.data if yin is not None else None,
                       xo.ctypes.data if xo is not None else None)
    if xo is not None:
        xo = xo[: lind.size - (nnz if trash else 0)]
    return yo, xo


def pixel_lists(indices, trash, nnz, npix):
    m = indices.size // nnz
    idl = np.empty(max(npix, 1), np.int32)
    ll = np.empty(max(m, 1), np.int32)
    lib().orc_pixel_lists(m, nnz, _c(indices, np.int32), int(trash), int(npix), idl, ll)
    return idl[:npix], ll[:m]


# ----------------------------------------------------------------------------
# FFT / Toeplitz
# ----------------------------------------------------------------------------
def fft(z, sign=-1):
    z = np.ascontiguousarray(z, dtype=np.complex128)
    a = z.view(np.float64).copy()
    out = np.empty_like(a)
    rc = lib().orc_fft(z.size, a, out, int(sign))
    if rc:
        raise RuntimeError(rc)
    return out.view(np.complex128)


NINV_REFERENCE, NINV_REAL_SPECTRUM, NINV_DIRECT = 0, 1, 2


def set_ninv_mode(mode: int) -> int:
    """Select how the oracle computes the Toeplitz tiles (see oracle.h); returns the old mode."""
    old = int(lib().orc_set_ninv_mode(int(mode)))
    for L in _LIBS.values():
        L.orc_set_ninv_mode(int(mode))
    return old


class ninv_mode:
    """Context manager: ``with ninv_mode(NINV_DIRECT): ...``"""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        self.old = set_ninv_mode(self.mode)

    def __exit__(self, *a):
        set_ninv_mode(self.old)


OFFSET_REFERENCE, OFFSET_SYMMETRIC, OFFSET_NONE = 0, 1, 2


class offset_mode:
    """Context manager selecting N^-1's per-block offset handling (see oracle.h)."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        self.old = int(lib().orc_set_offset_mode(int(self.mode)))
        for L in _LIBS.values():
            L.orc_set_offset_mode(int(self.mode))

    def __exit__(self, *a):
        for L in _LIBS.values():
            L.orc_set_offset_mode(self.old)


def define_blocksize(n, lam):
    return int(lib().orc_define_blocksize(int(n), int(lam)))


def stmm_block(v, T, lam, use_fft=True):
    V = _c(v, np.float64).copy()
    rc = lib().orc_stmm_block(V, V.size, _c(T, np.float64), int(lam), int(bool(use_fft)))
    if rc < 0:
        raise RuntimeError(rc)
    return V


def apply_weights(tod, block_sizes, lam, invtt, use_fft=True):
    """apply_weights (pcg_true.c:578-600): N^-1 applied to a TOD vector (copy)."""
    V = _c(tod, np.float64).copy()
    bs = _c(block_sizes, np.int32)
    rc = lib().orc_apply_weights(bs.size, bs, int(lam), _c(invtt, np.float64), V.size, V, int(bool(use_fft)))
    if rc < 0:
        raise RuntimeError(f"orc_apply_weights rc={rc}")
    return V


# ----------------------------------------------------------------------------
# LAPACK restatement / preconditioner
# ----------------------------------------------------------------------------
def dlange1(a):
    a = _c(a, np.float64)
    return lib().orc_dlange1(a.shape[0], a)


def dgetrf(a):
    lu = _c(a, np.float64).copy()
    ipiv = np.zeros(lu.shape[0], np.int32)
    info = lib().orc_dgetrf(lu.shape[0], lu, ipiv)
    return lu, ipiv, info


def dgecon1(lu, anorm):
    return lib().orc_dgecon1(lu.shape[0], _c(lu, np.float64), float(anorm))


def dgetri(lu, ipiv):
    a = _c(lu, np.float64).copy()
    info = lib().orc_dgetri(a.shape[0], a, _c(ipiv, np.int32))
    return a, info


def getlocalW(indices, values, trash, nnz, lcount, block_sizes, lam, invtt):
    m = indices.size // nnz
    n_valid = lcount - (nnz if trash else 0)
    W = np.empty(max(n_valid * nnz, 1), np.float64)
    hits = np.zeros(max(n_valid // nnz, 1), np.int32)
    bs = _c(block_sizes, np.int32)
    lib().orc_getlocalW(m, nnz, _c(indices, np.int32), _c(values, np.float64), int(trash), int(lcount), bs.size,
                        bs, int(lam), _c(invtt, np.float64), W, hits)
    return W[: n_valid * nnz].reshape(-1, nnz, nnz), hits[: n_valid // nnz]


@dataclass
class SolveResult:
    x: np.ndarray
    lindices: np.ndarray
    cond: np.ndarray
    hits: np.ndarray
    res_hist: np.ndarray
    iterations: int
    converged: bool
    res0: float
    res: float
    n_cut: int
    trash_pix: int
    pix_local: np.ndarray
    weights: np.ndarray
    signal: np.ndarray


def pinv(a, rcond=1e-15):
    """inverse_svd_bis (precond.c:777-796): pseudo-inverse, singular values
    <= rcond * s_max dropped."""
    a = _c(a, np.float64)
    n = a.shape[0]
    out = np.zeros((n, n), np.float64)
    if lib().orc_pinv(n, a.ravel().copy(), float(rcond), out.ravel()):
        raise RuntimeError("orc_pinv failed")
    return out


def symeig(a):
    """LAPACKE_dsyev row-major 'V' (precond.c:1240): ascending eigenvalues and
    the eigenvectors as columns."""
    a = _c(a, np.float64)
    n = a.shape[0]
    w = np.zeros(n, np.float64)
    v = np.zeros((n, n), np.float64)
    if lib().orc_symeig(n, a.ravel().copy(), w, v.ravel()):
        raise RuntimeError("orc_symeig failed")
    return w, v


def gls_solve(pix, weights, signal, noise, block_sizes, lam, invtt, nnz=3, tol=1e-8, maxiter=500,
              stop_after=0, precond=0, z_2lvl=0, deflation=False, omp=False):
    """MLmap setup + PCG_GLS_true on one rank (copies; inputs are not modified).
    precond / z_2lvl as in MLmap (0 block Jacobi, 1 two-level a priori, 2
    two-level a posteriori).  deflation=True also returns Z, AZ (n_map x Zn)
    and Em1 (Zn x Zn) as result.Z, result.AZ, result.Em1."""
    idx = _c(pix, np.int32).copy()
    w = _c(weights, np.float64).copy()
    s = _c(signal, np.float64).copy()
    nz = _c(noise, np.float64)
    m = idx.size // nnz
    bs = _c(block_sizes, np.int32)
    cap = max(idx.size, 1)
    x = np.zeros(cap, np.float64)
    lind = np.zeros(cap, np.int32)
    cond = np.zeros(max(m, 1) + cap // nnz, np.float64)
    hits = np.zeros(max(m, 1) + cap // nnz, np.int32)
    hist = np.zeros(max(maxiter, 1), np.float64)
    rep = OracleReport()
    zn = z_2lvl if z_2lvl > 0 else 1
    Zb = AZb = Eb = None
    if deflation and precond:
        Zb = np.zeros(zn * cap, np.float64)
        AZb = np.zeros(zn * cap, np.float64)
        Eb = np.zeros(zn * zn, np.float64)
    ptr = lambda a: None if a is None else a.ctypes.data_as(ct.c_void_p)
    rc = lib(omp).orc_gls_solve2(m, nnz, idx, w, s, nz, bs.size, bs, int(lam), _c(invtt, np.float64), float(tol),
                              int(maxiter), int(stop_after), int(precond), int(z_2lvl), x, lind, cond, hits, hist,
                              ptr(Zb), ptr(AZb), ptr(Eb), ct.byref(rep))
    if rc:
        raise RuntimeError(f"orc_gls_solve2 rc={rc}")
    n = rep.n_map
    nh = min(rep.iterations + 1, maxiter) if maxiter > 0 else 0
    res = SolveResult(x=x[:n].copy(), lindices=lind[: rep.lcount].copy(), cond=cond[: n // nnz].copy(),
                       hits=hits[: n // nnz].copy(), res_hist=hist[:nh].copy(), iterations=rep.iterations,
                       converged=bool(rep.converged), res0=rep.res0, res=rep.res, n_cut=rep.n_cut,
                       trash_pix=rep.trash_pix, pix_local=idx, weights=w, signal=s)
    if Zb is not None:
        n = rep.n_map
        res.Z = Zb[: zn * n].reshape(zn, n).T.copy()
        res.AZ = AZb[: zn * n].reshape(zn, n).T.copy()
        res.Em1 = Eb.reshape(zn, zn).copy()
    return res


# ----------------------------------------------------------------------------
# template marginalisation (MTmap / PCG_GLS_templates, templates.c)
# ----------------------------------------------------------------------------
class Templates:
    """Template classes of one rank, as MTmap receives them (mappraiser.c:387-392):
    nces CES of ndet detectors (blocks CES-major, each as long as its CES),
    sweeptstamps / az / hwp_angle as one array per CES.  az is binned with
    bin_az and the HWP angle turned into cos/sin here, as MTmap does.  w is the
    white-noise weight of each block (invtt with lambda 1)."""

    def __init__(self, block_sizes, nces, ndet, w, npoly=0, ground=0, n_sss_bins=0, nhwp=0, delta_t=0.0,
                 sampling_freq=1.0, sweeptstamps=None, az=None, az_min=None, az_max=None, hwp_angle=None):
        L = lib()
        self.nces, self.ndet = int(nces), int(ndet)
        self.block_sizes = _c(block_sizes, np.int32).copy()
        self.w = _c(w, np.float64).copy()
        lens = [int(self.block_sizes[c * self.ndet]) for c in range(self.nces)]
        self.lens = lens
        sw = sweeptstamps if sweeptstamps is not None else [np.array([0, n], np.int32) for n in lens]
        self.nsweeps = np.array([len(x) - 1 for x in sw], np.int32)
        self.stamps = np.concatenate([_c(x, np.int32) for x in sw]).astype(np.int32)
        self.az_binned = None
        if ground:
            parts = []
            for c in range(self.nces):
                out = np.zeros(lens[c], np.int32)
                if L.orc_bin_az(lens[c], _c(az[c], np.float64), float(az_min[c]), float(az_max[c]),
                                int(n_sss_bins), out):
                    raise ValueError("azimuth below az_min")
                parts.append(out)
            self.az_binned = np.concatenate(parts).astype(np.int32)
        self.hwp_cos = self.hwp_sin = None
        if nhwp:
            ang = np.concatenate([_c(a, np.float64)[: lens[c]] for c, a in enumerate(hwp_angle)])
            self.hwp_cos = np.zeros_like(ang)
            self.hwp_sin = np.zeros_like(ang)
            L.orc_hwp_mod(ang.size, ang, self.hwp_cos, self.hwp_sin)
        p = lambda a: None if a is None else a.ctypes.data_as(ct.c_void_p)
        self.st = _Tmpl(int(npoly), int(ground), int(n_sss_bins), int(nhwp), float(delta_t), float(sampling_freq),
                        self.nces, self.ndet, p(self.block_sizes), p(self.nsweeps), p(self.stamps),
                        p(self.az_binned), p(self.hwp_cos), p(self.hwp_sin), p(self.w))
        nb = self.nces * self.ndet
        self.nt = np.zeros(nb, np.int32)
        tot = L.orc_tmpl_sizes(ct.byref(self.st), self.nt.ctypes.data_as(ct.c_void_p))
        if tot < 0:
            raise ValueError(f"orc_tmpl_sizes rc={tot}")
        self.ntot = int(tot)
        self.m = int(self.block_sizes.sum())

    def tvec(self, tau):
        out = np.zeros(self.m, np.float64)
        if lib().orc_tmpl_tvec(ct.byref(self.st), _c(tau, np.float64), out):
            raise RuntimeError("orc_tmpl_tvec failed")
        return out

    def trtvec(self, d):
        out = np.zeros(self.ntot, np.float64)
        if lib().orc_tmpl_trtvec(ct.byref(self.st), _c(d, np.float64), out):
            raise RuntimeError("orc_tmpl_trtvec failed")
        return out

    def kernel(self):
        """BuildKernel of every block: concatenated row-major nt_b x nt_b blocks."""
        out = np.zeros(int((self.nt.astype(np.int64) ** 2).sum()), np.float64)
        if lib().orc_tmpl_kernel(ct.byref(self.st), out):
            raise RuntimeError("orc_tmpl_kernel failed")
        return out

    def invkernel(self, B):
        out = np.zeros_like(B)
        ranks = np.zeros(self.nt.size, np.int32)
        if lib().orc_tmpl_invkernel(ct.byref(self.st), _c(B, np.float64), out, ranks):
            raise RuntimeError("orc_tmpl_invkernel failed")
        return out, ranks

    def blocks(self, flat):
        """split a concatenated kernel array into its nt_b x nt_b blocks"""
        out, o = [], 0
        for n in self.nt:
            n = int(n)
            out.append(flat[o: o + n * n].reshape(n, n))
            o += n * n
        return out

    def filter(self, Binv, tod):
        t = _c(tod, np.float64).copy()
        if lib().orc_tmpl_filter(ct.byref(self.st), _c(Binv, np.float64), t):
            raise RuntimeError("orc_tmpl_filter failed")
        return t


def gls_solve_tmpl(pix, weights, signal, noise, tm: Templates, invtt, nnz=3, tol=1e-8, maxiter=500, stop_after=0,
                   precond=0, z_2lvl=0):
    """MTmap setup + PCG_GLS_templates on one rank (copies; inputs are not
    modified).  invtt holds one white-noise weight per block (lambda 1).
    Returns a SolveResult with .ranks (the kernel ranks InvKernel reports)."""
    idx = _c(pix, np.int32).copy()
    w = _c(weights, np.float64).copy()
    s = _c(signal, np.float64).copy()
    nz = _c(noise, np.float64)
    m = idx.size // nnz
    cap = max(idx.size, 1)
    x = np.zeros(cap, np.float64)
    lind = np.zeros(cap, np.int32)
    cond = np.zeros(max(m, 1) + cap // nnz, np.float64)
    hits = np.zeros(max(m, 1) + cap // nnz, np.int32)
    hist = np.zeros(max(maxiter, 1), np.float64)
    ranks = np.zeros(tm.nt.size, np.int32)
    rep = OracleReport()
    rc = lib().orc_gls_solve_tmpl(m, nnz, idx, w, s, nz, tm.block_sizes, _c(invtt, np.float64), float(tol),
                                  int(maxiter), int(stop_after), int(precond), int(z_2lvl), ct.byref(tm.st), x, lind,
                                  cond, hits, hist, ranks, ct.byref(rep))
    if rc:
        raise RuntimeError(f"orc_gls_solve_tmpl rc={rc}")
    n = rep.n_map
    nh = min(rep.iterations + 1, maxiter) if maxiter > 0 else 0
    res = SolveResult(x=x[:n].copy(), lindices=lind[: rep.lcount].copy(), cond=cond[: n // nnz].copy(),
                      hits=hits[: n // nnz].copy(), res_hist=hist[:nh].copy(), iterations=rep.iterations,
                      converged=bool(rep.converged), res0=rep.res0, res=rep.res, n_cut=rep.n_cut,
                      trash_pix=rep.trash_pix, pix_local=idx, weights=w, signal=s)
    res.ranks = ranks
    return res
