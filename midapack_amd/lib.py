"""Python host mirror of the libmappraiser_gpu.so C-ABI (include/mappraiser_gpu.h).

The library is the product; this module only loads it and marshals numpy
arrays with the dtype contract of the reference's ctypes wrapper
(mappraiser_wrapper/__init__.py:23-28: int32 pixels, float64 everything else).
There is no CPU fallback: if the shared library is missing or fails to load,
every call raises.
"""
from __future__ import annotations

import ctypes as ct
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmappraiser_gpu.so")

MGPU_FP64 = 0
MGPU_MIXED = 1
MGPU_OFFSET_REFERENCE = 0
MGPU_OFFSET_SYMMETRIC = 1
MGPU_OFFSET_NONE = 2
MGPU_TILE_AUTOTUNE = -1
MGPU_REFRESH_ADAPTIVE = -1
MGPU_APOD_CHEBWIN = 0
MGPU_APOD_GAUSSIAN = 1

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = ct.c_void_p

# every symbol declared in include/mappraiser_gpu.h
EXPORTS = [
    "mappraiser_gpu_init", "mappraiser_gpu_finalize", "mappraiser_gpu_last_error",
    "mappraiser_gpu_options_default", "mappraiser_gpu_problem_create",
    "mappraiser_gpu_problem_create_device", "mappraiser_gpu_problem_destroy",
    "mappraiser_gpu_local_shape", "mappraiser_gpu_matvec", "mappraiser_gpu_trmatvec",
    "mappraiser_gpu_apply_weights", "mappraiser_gpu_apply_weights_noalloc", "mappraiser_gpu_precond_build", "mappraiser_gpu_precond_apply",
    "mappraiser_gpu_pcg", "mappraiser_gpu_solve_device", "mappraiser_gpu_kernel_times",
    "mappraiser_gpu_tile_variants", "mappraiser_gpu_pcg_batch", "mappraiser_gpu_server_create",
    "mappraiser_gpu_server_submit", "mappraiser_gpu_server_wait", "mappraiser_gpu_server_stats",
    "mappraiser_gpu_server_destroy", "mappraiser_gpu_noise_invtt", "mappraiser_gpu_pcg_gls",
    "mappraiser_gpu_x2map_pol", "mappraiser_gpu_write_map", "mappraiser_gpu_templates_set",
    "mappraiser_gpu_templates_info", "mappraiser_gpu_templates_apply", "mappraiser_gpu_templates_kernel",
]
# the reference's own entry points, exported with its signatures (map.h:10-18)
REFERENCE_EXPORTS = ["MLmap", "MTmap"]
MGPU_TDOUBLE, MGPU_TINT = 82, 31

# kernel classes of mappraiser_gpu_kernel_times (MGPU_KT_*)
KERNEL_CLASSES = ["P", "mean", "tiles", "Pt", "map", "refresh", "templates"]
MGPU_TMPL_T, MGPU_TMPL_TT, MGPU_TMPL_FILTER = 0, 1, 2


class Options(ct.Structure):
    _fields_ = [("precision", ct.c_int), ("refresh_every", ct.c_int), ("tile_log2", ct.c_int),
                ("poll_every", ct.c_int), ("profile", ct.c_int), ("offset_mode", ct.c_int),
                ("tile_variant", ct.c_int), ("graph", ct.c_int), ("precond", ct.c_int), ("z_2lvl", ct.c_int)]


class Report(ct.Structure):
    _fields_ = [("iterations", ct.c_int), ("converged", ct.c_int), ("res0", ct.c_double),
                ("res", ct.c_double), ("n_map", ct.c_int), ("lcount", ct.c_int), ("trash_pix", ct.c_int),
                ("n_cut", ct.c_int), ("t_setup_ms", ct.c_double), ("t_solve_ms", ct.c_double),
                ("true_res_rel", ct.c_double), ("refreshes", ct.c_int), ("tile_variant", ct.c_int),
                ("t_2lvl_ms", ct.c_double), ("dev_peak_bytes", ct.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchItem(ct.Structure):
    """mgpu_batch_item of include/mappraiser_gpu.h"""
    _fields_ = [("m", ct.c_int), ("pix", _vp), ("pixweights", _vp), ("signal", _vp), ("noise", _vp),
                ("nb", ct.c_int), ("local_blocks_sizes", _vp), ("lam", ct.c_int), ("invtt", _vp),
                ("x", _vp), ("pixels", _vp), ("cond", _vp), ("lhits", _vp), ("x_cap", ct.c_int),
                ("rep", Report)]


class TemplatesSpec(ct.Structure):
    """mgpu_templates (include/mappraiser_gpu.h)"""
    _fields_ = [
        ("npoly", ct.c_int),
        ("ground", ct.c_int),
        ("n_sss_bins", ct.c_int),
        ("nhwp", ct.c_int),
        ("delta_t", ct.c_double),
        ("sampling_freq", ct.c_double),
        ("nces", ct.c_int),
        ("sweeptstamps", ct.c_void_p),
        ("nsweeps", ct.c_void_p),
        ("az", ct.c_void_p),
        ("az_min", ct.c_void_p),
        ("az_max", ct.c_void_p),
        ("hwp_angle", ct.c_void_p),
    ]


def _row_pointers(rows, dt):
    """C-contiguous copies of the rows and an array of their addresses (the
    int** / double** arguments of MTmap, as the reference's wrapper builds them)"""
    keep = [np.ascontiguousarray(r, dtype=dt) for r in rows]
    ptrs = np.array([k.ctypes.data for k in keep], dtype=np.uintp)
    return keep, ptrs


def templates_spec(meta: dict):
    """mgpu_templates from a workload's template metadata; returns (spec, keep),
    keep holding (by name) the arrays the spec points to."""
    keep = {}
    keep["sw_rows"], keep["sweeptstamps"] = _row_pointers(meta["sweeptstamps"], np.int32)
    keep["nsweeps"] = np.array([len(x) - 1 for x in keep["sw_rows"]], np.int32)
    if meta.get("ground"):
        keep["az_rows"], keep["az"] = _row_pointers(meta["az"], np.float64)
        keep["az_min"] = np.ascontiguousarray(meta["az_min"], np.float64)
        keep["az_max"] = np.ascontiguousarray(meta["az_max"], np.float64)
    if meta.get("nhwp"):
        keep["hwp_rows"], keep["hwp_angle"] = _row_pointers(meta["hwp_angle"], np.float64)
    p = lambda k: keep[k].ctypes.data if k in keep else None
    spec = TemplatesSpec(int(meta["npoly"]), int(meta["ground"]), int(meta["n_sss_bins"]), int(meta["nhwp"]),
                         float(meta["delta_t"]), float(meta["sampling_freq"]), int(meta["nces"]), p("sweeptstamps"),
                         p("nsweeps"), p("az"), p("az_min"), p("az_max"), p("hwp_angle"))
    return spec, keep


class GPUError(RuntimeError):
    pass


_LIB = None


def load():
    """Load libmappraiser_gpu.so (raises if it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise GPUError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = ct.CDLL(LIB_PATH)
    L.mappraiser_gpu_init.argtypes = [ct.c_int]
    L.mappraiser_gpu_last_error.restype = ct.c_char_p
    L.mappraiser_gpu_options_default.argtypes = [ct.POINTER(Options)]
    L.mappraiser_gpu_options_default.restype = None
    L.mappraiser_gpu_problem_create.argtypes = [ct.POINTER(_vp), ct.c_int, ct.c_int, _i32p, _f64p, ct.c_int, _i32p,
                                                ct.c_int, _f64p, ct.POINTER(Options)]
    L.mappraiser_gpu_problem_create_device.argtypes = [ct.POINTER(_vp), ct.c_int, ct.c_int, _vp, _vp, ct.c_int,
                                                       _i32p, ct.c_int, _vp, ct.POINTER(Options)]
    L.mappraiser_gpu_problem_destroy.argtypes = [_vp]
    L.mappraiser_gpu_local_shape.argtypes = [_vp, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), _vp, _vp]
    L.mappraiser_gpu_matvec.argtypes = [_vp, _f64p, _f64p]
    L.mappraiser_gpu_trmatvec.argtypes = [_vp, _f64p, _f64p]
    L.mappraiser_gpu_apply_weights.argtypes = [_vp, _f64p]
    L.mappraiser_gpu_apply_weights_noalloc.argtypes = [_vp, _vp, _vp, _vp]
    L.mappraiser_gpu_precond_build.argtypes = [_vp, _vp, _vp, _vp, ct.POINTER(ct.c_int)]
    L.mappraiser_gpu_precond_apply.argtypes = [_vp, _f64p, _f64p]
    L.mappraiser_gpu_pcg.argtypes = [_vp, _f64p, _f64p, ct.c_double, ct.c_int, ct.c_int, _f64p, _vp,
                                     ct.POINTER(Report)]
    L.mappraiser_gpu_solve_device.argtypes = [_vp, _vp, _vp, ct.c_double, ct.c_int, _f64p, _vp, _vp, _vp,
                                              ct.POINTER(Report)]
    L.mappraiser_gpu_kernel_times.argtypes = [_vp, _vp, _vp, ct.c_int]
    L.mappraiser_gpu_tile_variants.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.POINTER(ct.c_int), ct.c_char_p,
                                               ct.c_int]
    L.mappraiser_gpu_pcg_batch.argtypes = [ct.c_int, ct.POINTER(BatchItem), ct.c_int, ct.c_int, ct.c_double,
                                           ct.c_int, ct.POINTER(Options)]
    L.mappraiser_gpu_server_create.argtypes = [ct.POINTER(_vp), ct.c_double, ct.c_int, ct.c_int,
                                               ct.POINTER(Options)]
    L.mappraiser_gpu_server_submit.argtypes = [_vp, ct.POINTER(BatchItem), ct.c_int, ct.c_double, ct.c_int,
                                               ct.POINTER(ct.c_longlong)]
    L.mappraiser_gpu_server_wait.argtypes = [_vp, ct.c_longlong, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]
    L.mappraiser_gpu_server_stats.argtypes = [_vp, ct.POINTER(ct.c_longlong), ct.POINTER(ct.c_longlong),
                                              ct.POINTER(ct.c_double), ct.POINTER(ct.c_int),
                                              ct.POINTER(ct.c_longlong)]
    L.mappraiser_gpu_server_destroy.argtypes = [_vp]
    L.mappraiser_gpu_noise_invtt.argtypes = [_vp, ct.c_int, _i32p, ct.c_int, ct.c_double, ct.c_int, ct.c_int, _vp,
                                             _vp]
    L.mappraiser_gpu_pcg_gls.argtypes = [ct.c_char_p, ct.c_char_p, ct.c_int, ct.c_int, _i32p, _f64p, _f64p, _f64p,
                                         ct.c_int, _i32p, ct.c_int, _f64p, ct.c_double, ct.c_int, _f64p, _f64p,
                                         _i32p, _i32p, ct.POINTER(Options), ct.POINTER(Report)]
    L.mappraiser_gpu_x2map_pol.argtypes = [_f64p, _f64p, _f64p, _f64p, _i32p, _f64p, _i32p, _f64p, _i32p, ct.c_int]
    L.mappraiser_gpu_write_map.argtypes = [_vp, ct.c_int, ct.c_long, ct.c_char_p, ct.c_char, ct.c_char_p]
    # MLmap with the argtypes of the reference's wrapper (mappraiser_wrapper/__init__.py:58-84),
    # the communicator passed as an int handle
    L.mappraiser_gpu_templates_set.argtypes = [_vp, ct.POINTER(TemplatesSpec)]
    L.mappraiser_gpu_templates_info.argtypes = [_vp, ct.POINTER(ct.c_longlong), _vp, _vp]
    L.mappraiser_gpu_templates_apply.argtypes = [_vp, ct.c_int, _f64p, _f64p]
    L.mappraiser_gpu_templates_kernel.argtypes = [_vp, _vp, _vp]
    # MTmap with the argtypes of the reference's wrapper (mappraiser_wrapper/__init__.py:86-122)
    _up = np.ctypeslib.ndpointer(dtype=np.uintp, ndim=1, flags="C_CONTIGUOUS")
    L.MTmap.restype = None
    L.MTmap.argtypes = [ct.c_int, ct.c_char_p, ct.c_char_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                        ct.c_int, ct.c_double, ct.c_int, ct.c_int, ct.c_double, ct.c_int, ct.c_int, ct.c_int,
                        ct.c_int, ct.c_int, _up, _i32p, _vp, _vp, _vp, _vp, ct.c_int, _i32p, ct.c_int, _i32p,
                        ct.c_int, _i32p, _f64p, _f64p, _f64p, ct.c_double, _f64p]
    L.MLmap.restype = None
    L.MLmap.argtypes = [ct.c_int, ct.c_char_p, ct.c_char_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_double,
                        ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _i32p, ct.c_int, _i32p, ct.c_int, _i32p,
                        _f64p, _f64p, _f64p, ct.c_int, _f64p]
    _LIB = L
    return L


def last_error() -> str:
    return load().mappraiser_gpu_last_error().decode(errors="replace")


def _check(rc: int, what: str):
    if rc != 0:
        raise GPUError(f"{what} failed (status {rc}): {last_error()}")


def options(precision=MGPU_FP64, refresh_every=10, tile_log2=0, poll_every=8, profile=False,
            offset_mode=None, tile_variant=0, graph=None, precond=0, z_2lvl=0) -> Options:
    """mgpu_options; precond / z_2lvl are MLmap's precond and Z_2lvl (0 block
    Jacobi, 1 two-level a priori, 2 two-level a posteriori)."""
    o = Options()
    load().mappraiser_gpu_options_default(ct.byref(o))
    o.precision, o.refresh_every, o.tile_log2, o.poll_every = precision, refresh_every, tile_log2, poll_every
    o.profile = int(bool(profile))
    o.tile_variant = int(tile_variant)
    o.precond, o.z_2lvl = int(precond), int(z_2lvl)
    if graph is not None:
        o.graph = int(bool(graph))
    if offset_mode is not None:
        o.offset_mode = int(offset_mode)
    return o


def x2map_pol(x, lstid, cond, lhits, nside):
    """x2map_pol (mappraiser.c:836-853) into zeroed HEALPix maps of 12 nside^2
    pixels: returns mapI, mapQ, mapU, Cond, hits."""
    npix = 12 * nside * nside
    lstid = _c(lstid, np.int32)
    if lstid.size and (lstid.min() < 0 or lstid.max() // 3 >= npix):
        raise GPUError("map index outside the HEALPix map")
    out = [np.zeros(npix) for _ in range(4)]
    hits = np.zeros(npix, np.int32)
    x = _c(x, np.float64)
    _check(load().mappraiser_gpu_x2map_pol(out[0], out[1], out[2], out[3], hits, x, lstid, _c(cond, np.float64),
                                           _c(lhits, np.int32), int(x.size)), "x2map_pol")
    return out[0], out[1], out[2], out[3], hits


def write_map(data, nside, filename, nest=True, coordsys="C"):
    """write_map (iofiles.c:19-53): HEALPix FITS map; float64 -> TDOUBLE, int32 -> TINT."""
    a = np.ascontiguousarray(data)
    if a.dtype not in (np.float64, np.int32) or a.size != 12 * nside * nside:
        raise GPUError("write_map: float64 or int32 data of 12 nside^2 values")
    typ = MGPU_TDOUBLE if a.dtype == np.float64 else MGPU_TINT
    _check(load().mappraiser_gpu_write_map(a.ctypes.data, typ, int(nside), filename.encode(),
                                           bytes([1 if nest else 0]), coordsys.encode()), "write_map")


def MLmap(outpath, ref, precond, Z_2lvl, tol, maxiter, nside, local_blocks_sizes, nnz, pix, pixweights, signal,
          noise, lam, invtt, solver=0):
    """MLmap called exactly as the reference's wrapper calls it (one rank);
    pix, pixweights and signal are modified in place like the reference's."""
    bs = _c(local_blocks_sizes, np.int32)
    m = np.array([_c(signal, np.float64).size], np.int32)
    load().MLmap(0, outpath.encode(), ref.encode(), int(solver), int(precond), int(Z_2lvl), 6, float(tol),
                 int(maxiter), 1, 0, 0, int(nside), m, int(bs.size), bs, int(nnz), pix, pixweights, signal,
                 _c(noise, np.float64), int(lam), _c(invtt, np.float64))


def MTmap(outpath, ref, precond, Z_2lvl, tol, maxiter, nside, meta, local_blocks_sizes, nnz, pix, pixweights,
          signal, noise, invtt, solver=0):
    """MTmap called as the reference's wrapper calls it (one rank); meta holds
    the template inputs (npoly, nhwp, delta_t, ground, n_sss_bins,
    sampling_freq, nces and per-CES sweeptstamps / az / az_min / az_max /
    hwp_angle).  pix, pixweights and signal are modified in place."""
    bs = _c(local_blocks_sizes, np.int32)
    m = np.array([_c(signal, np.float64).size], np.int32)
    _, keep = templates_spec(meta)
    vp = lambda k: keep[k].ctypes.data_as(ct.c_void_p) if k in keep else None
    load().MTmap(0, outpath.encode(), ref.encode(), int(solver), int(precond), int(Z_2lvl), 6, int(meta["npoly"]),
                 int(meta["nhwp"]), float(meta["delta_t"]), int(meta["ground"]), int(meta["n_sss_bins"]), float(tol),
                 int(maxiter), 1, 0, 0, int(nside), keep["sweeptstamps"], keep["nsweeps"], vp("az"), vp("az_min"),
                 vp("az_max"), vp("hwp_angle"),
                 int(meta["nces"]), m, int(bs.size), bs, int(nnz), pix, pixweights, signal, _c(noise, np.float64),
                 float(meta["sampling_freq"]), _c(invtt, np.float64))
    del keep


def read_map(filename):
    """Column SIGNAL of a HEALPix FITS binary table as written by write_map."""
    with open(filename, "rb") as f:
        raw = f.read()
    pos, hdus = 0, []
    while pos < len(raw) and len(hdus) < 2:
        cards = {}
        while True:
            block = raw[pos:pos + 2880]
            pos += 2880
            done = False
            for i in range(0, 2880, 80):
                c = block[i:i + 80].decode("ascii")
                if c.startswith("END"):
                    done = True
                    break
                if c[8:10] == "= ":
                    cards[c[:8].strip()] = c[10:].split("/")[0].strip().strip("'").strip()
            if done:
                break
        hdus.append(cards)
        if len(hdus) == 2:
            n = int(cards["NAXIS2"])
            return np.frombuffer(raw[pos:pos + 8 * n], dtype=">f8").astype(np.float64), cards
    raise GPUError(f"{filename}: no binary table")


def tile_variants(precision: int, tile_log2: int) -> list[str]:
    """Descriptions of the tile kernel variants for a precision and real tile length 2^tile_log2."""
    L = load()
    n = ct.c_int(0)
    _check(L.mappraiser_gpu_tile_variants(int(precision), int(tile_log2), 0, ct.byref(n), None, 0), "tile_variants")
    out = []
    for v in range(1, n.value + 1):
        buf = ct.create_string_buffer(256)
        _check(L.mappraiser_gpu_tile_variants(int(precision), int(tile_log2), v, ct.byref(n), buf, 256),
               "tile_variants")
        out.append(buf.value.decode())
    return out


def finalize():
    """Release the library's cached device blocks and pinned staging buffers."""
    _check(load().mappraiser_gpu_finalize(), "finalize")


def init(device: int = 0):
    _check(load().mappraiser_gpu_init(int(device)), "mappraiser_gpu_init")


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class Problem:
    """One rank's GLS problem resident on the GPU (MatInit + Toeplitz blocks)."""

    def __init__(self, pix, weights, block_sizes, lam, invtt, nnz=3, opts: Options | None = None):
        L = load()
        self._h = _vp()
        pix = _c(pix, np.int32)
        w = _c(weights, np.float64)
        bs = _c(block_sizes, np.int32)
        self.m = pix.size // nnz
        self.opts = opts or options()
        _check(L.mappraiser_gpu_problem_create(ct.byref(self._h), self.m, nnz, pix, w, bs.size, bs, int(lam),
                                               _c(invtt, np.float64), ct.byref(self.opts)),
               "problem_create")
        self.n_map = None
        self._block_count = bs.size

    @classmethod
    def from_device(cls, m, d_pix, d_w, block_sizes, lam, d_invtt, nnz=3, opts: Options | None = None):
        self = cls.__new__(cls)
        L = load()
        self._h = _vp()
        bs = _c(block_sizes, np.int32)
        self.m = int(m)
        self.opts = opts or options()
        _check(L.mappraiser_gpu_problem_create_device(ct.byref(self._h), self.m, nnz, _vp(d_pix), _vp(d_w), bs.size,
                                                      bs, int(lam), _vp(d_invtt), ct.byref(self.opts)),
               "problem_create_device")
        self.n_map = None
        self._block_count = bs.size
        return self

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load().mappraiser_gpu_problem_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- template marginalisation (MTmap) -----------------------------------
    def set_templates(self, meta: dict):
        """Attach the template classes described by meta (see MTmap)."""
        spec, keep = templates_spec(meta)
        _check(load().mappraiser_gpu_templates_set(self._h, ct.byref(spec)), "templates_set")
        del keep
        ntot = ct.c_longlong()
        _check(load().mappraiser_gpu_templates_info(self._h, ct.byref(ntot), None, None), "templates_info")
        self.n_templates = int(ntot.value)

    def templates_info(self):
        nb = self._nb()
        nt = np.zeros(nb, np.int32)
        rk = np.zeros(nb, np.int32)
        ntot = ct.c_longlong()
        _check(load().mappraiser_gpu_templates_info(self._h, ct.byref(ntot), nt.ctypes.data_as(_vp),
                                                    rk.ctypes.data_as(_vp)), "templates_info")
        return int(ntot.value), nt, rk

    def _nb(self):
        return int(self._block_count)

    def templates_apply(self, op, v):
        n_out = self.n_templates if op == MGPU_TMPL_TT else self.m
        out = np.zeros(n_out, np.float64)
        _check(load().mappraiser_gpu_templates_apply(self._h, int(op), _c(v, np.float64), out), "templates_apply")
        return out

    def templates_kernel(self):
        _, nt, _ = self.templates_info()
        n = int((nt.astype(np.int64) ** 2).sum())
        B = np.zeros(max(n, 1), np.float64)
        Bi = np.zeros(max(n, 1), np.float64)
        _check(load().mappraiser_gpu_templates_kernel(self._h, B.ctypes.data_as(_vp), Bi.ctypes.data_as(_vp)),
               "templates_kernel")
        return B[:n], Bi[:n]

    # -- MatLocalShape ------------------------------------------------------
    def local_shape(self, want_indices=True):
        L = load()
        lc, tr = ct.c_int(0), ct.c_int(0)
        _check(L.mappraiser_gpu_local_shape(self._h, ct.byref(lc), ct.byref(tr), None, None), "local_shape")
        lind = np.empty(max(lc.value, 1), np.int32)
        idx = np.empty(self.m * 3, np.int32) if want_indices else None
        _check(L.mappraiser_gpu_local_shape(self._h, ct.byref(lc), ct.byref(tr), lind.ctypes.data,
                                            idx.ctypes.data if idx is not None else None), "local_shape")
        return lind[: lc.value], idx, int(tr.value)

    def map_size(self):
        lind, _, tr = self.local_shape(want_indices=False)
        return lind.size - 3 * tr

    # -- operators ------------------------------------------------------------
    def matvec(self, x):
        y = np.empty(self.m, np.float64)
        _check(load().mappraiser_gpu_matvec(self._h, _c(x, np.float64), y), "matvec")
        return y

    def trmatvec(self, y):
        x = np.empty(self.map_size(), np.float64)
        _check(load().mappraiser_gpu_trmatvec(self._h, _c(y, np.float64), x), "trmatvec")
        return x

    def apply_weights(self, tod):
        v = _c(tod, np.float64).copy()
        _check(load().mappraiser_gpu_apply_weights(self._h, v), "apply_weights")
        return v

    def apply_weights_noalloc(self, d_in, d_out, stream=None):
        """d_out <- N^-1 d_in on device vectors (integer device addresses, or
        objects with data_ptr(); m float64 each) without any allocation; queued
        on `stream` (a raw hipStream_t, None: the handle's stream), not
        synchronised."""
        def ptr(a):
            return a.data_ptr() if hasattr(a, "data_ptr") else int(a)

        _check(load().mappraiser_gpu_apply_weights_noalloc(self._h, _vp(ptr(d_in)), _vp(ptr(d_out)),
                                                           _vp(stream) if stream else None),
               "apply_weights_noalloc")

    def precond_build(self, signal=None):
        """Returns (signal after the cut, cond, lhits)."""
        npix0 = self.map_size() // 3
        cond = np.empty(max(npix0, 1), np.float64)
        hits = np.empty(max(npix0, 1), np.int32)
        n = ct.c_int(0)
        sig = _c(signal, np.float64).copy() if signal is not None else None
        _check(load().mappraiser_gpu_precond_build(self._h, sig.ctypes.data if sig is not None else None,
                                                   cond.ctypes.data, hits.ctypes.data, ct.byref(n)),
               "precond_build")
        self.n_map = n.value
        return sig, cond[: n.value // 3].copy(), hits[: n.value // 3].copy()

    def precond_apply(self, g):
        out = np.empty(self.map_size(), np.float64)
        _check(load().mappraiser_gpu_precond_apply(self._h, _c(g, np.float64), out), "precond_apply")
        return out

    def kernel_times(self, reset=False):
        """{class: (total ms, launches)} measured with HIP events (needs options(profile=True))."""
        ms = np.zeros(len(KERNEL_CLASSES), np.float64)
        cnt = np.zeros(len(KERNEL_CLASSES), np.int64)
        _check(load().mappraiser_gpu_kernel_times(self._h, ms.ctypes.data, cnt.ctypes.data, int(reset)),
               "kernel_times")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_CLASSES)}

    def solve_device(self, d_signal, d_noise, tol=1e-8, maxiter=500):
        """precond_build + PCG on device-resident signal/noise (d_signal is modified on cut samples)."""
        npix0 = self.map_size() // 3
        x = np.empty(max(3 * npix0, 1), np.float64)
        hist = np.zeros(max(maxiter, 1), np.float64)
        rep = Report()
        _check(load().mappraiser_gpu_solve_device(self._h, _vp(d_signal), _vp(d_noise), float(tol), int(maxiter), x,
                                                  hist.ctypes.data, None, None, ct.byref(rep)), "solve_device")
        nh = min(rep.iterations + 1, maxiter)
        return x[: rep.n_map], hist[:nh], rep

    def pcg(self, signal, noise, tol=1e-8, maxiter=500, stop_after=0, history=True):
        n = self.map_size()
        x = np.empty(max(n, 1), np.float64)
        hist = np.zeros(max(maxiter, 1), np.float64) if history else None
        rep = Report()
        _check(load().mappraiser_gpu_pcg(self._h, _c(signal, np.float64), _c(noise, np.float64), float(tol),
                                         int(maxiter), int(stop_after), x,
                                         hist.ctypes.data if hist is not None else None, ct.byref(rep)), "pcg")
        nh = min(rep.iterations + 1, maxiter)
        return x[:n], (hist[:nh] if hist is not None else None), rep


def dropin_kernel_times(reset=False):
    """{class: (total ms, launches)} over the drop-in solves (pcg_gls, MLmap, MTmap)
    and batches run with options(profile=True) in this process."""
    ms = np.zeros(len(KERNEL_CLASSES), np.float64)
    cnt = np.zeros(len(KERNEL_CLASSES), np.int64)
    _check(load().mappraiser_gpu_kernel_times(None, ms.ctypes.data, cnt.ctypes.data, int(reset)), "kernel_times")
    return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_CLASSES)}


@dataclass
class GLSResult:
    x: np.ndarray
    lindices: np.ndarray
    cond: np.ndarray
    hits: np.ndarray
    res_hist: np.ndarray
    report: dict
    pix_local: np.ndarray
    weights: np.ndarray
    signal: np.ndarray


def pcg_gls(pix, weights, signal, noise, block_sizes, lam, invtt, tol=1e-8, maxiter=500, nnz=3,
            opts: Options | None = None, outpath: str | None = None, ref: str | None = None,
            inplace: bool = False) -> GLSResult:
    """PCG_GLS_true drop-in on host arrays.  By default it works on copies (the
    reference's in-place side effects are applied to the copies returned in the
    result); inplace=True modifies pix, weights and signal themselves, like the
    reference (they must then be C-contiguous int32 / float64 arrays)."""
    L = load()
    if inplace:
        for a, dt in ((pix, np.int32), (weights, np.float64), (signal, np.float64)):
            if not (isinstance(a, np.ndarray) and a.dtype == dt and a.flags.c_contiguous and a.flags.writeable):
                raise GPUError("inplace=True needs writable C-contiguous int32 pix and float64 weights/signal")
        p, w, s = pix, weights, signal
    else:
        p = _c(pix, np.int32).copy()
        w = _c(weights, np.float64).copy()
        s = _c(signal, np.float64).copy()
    nz = _c(noise, np.float64)
    m = p.size // nnz
    bs = _c(block_sizes, np.int32)
    x = np.zeros(max(p.size, 1), np.float64)
    cond = np.zeros(max(p.size // nnz, 1), np.float64)
    hits = np.zeros(max(p.size // nnz, 1), np.int32)
    lind = np.zeros(max(p.size, 1), np.int32)
    rep = Report()
    o = opts or options()
    _check(L.mappraiser_gpu_pcg_gls(outpath.encode() if outpath else None, ref.encode() if ref else None, m, nnz, p,
                                    w, s, nz, bs.size, bs, int(lam), _c(invtt, np.float64), float(tol), int(maxiter),
                                    x, cond, hits, lind, ct.byref(o), ct.byref(rep)), "pcg_gls")
    n = rep.n_map
    hist = None
    if outpath and ref:
        fn = os.path.join(outpath, f"pcg_residuals_{ref}.dat")
        if os.path.exists(fn):
            hist = np.fromfile(fn, dtype=np.float64)
    return GLSResult(x=x[:n].copy(), lindices=lind[: rep.lcount].copy(), cond=cond[: n // 3].copy(),
                     hits=hits[: n // 3].copy(), res_hist=hist, report=rep.as_dict(), pix_local=p, weights=w,
                     signal=s)


@dataclass
class BatchResult:
    x: np.ndarray        # (I,Q,U) per kept pixel, pixel order
    pixels: np.ndarray   # global pixel index of each kept pixel
    cond: np.ndarray
    hits: np.ndarray
    report: dict


def pcg_batch(items, grid_npix, tol=1e-8, maxiter=500, opts: Options | None = None) -> list[BatchResult]:
    """Solve independent problems together (mappraiser_gpu_pcg_batch).  Each item
    is a dict with pix, weights, signal, noise, block_sizes, lam, invtt: numpy
    arrays with the dtypes of pcg_gls, or (all items alike) device tensors with
    a data_ptr() (int32 pix, float64 others; block_sizes stays a host array)."""
    L = load()
    n = len(items)
    arr = (BatchItem * n)()
    keep = []
    outs = []
    on_dev = hasattr(items[0]["pix"], "data_ptr")
    for i, it in enumerate(items):
        k, o = _fill_item(arr[i], it, grid_npix)
        keep.append(k)
        outs.append(o)
    o = opts or options()
    _check(L.mappraiser_gpu_pcg_batch(n, arr, int(grid_npix), int(on_dev), float(tol), int(maxiter), ct.byref(o)),
           "pcg_batch")
    return [_item_result(arr[i], outs[i]) for i in range(n)]


def _ptr(a, dt):
    """(pointer, element count, object to keep alive) of a host array or a device tensor"""
    if hasattr(a, "data_ptr"):
        return a.data_ptr(), a.numel(), a
    h = _c(a, dt)
    return h.ctypes.data, h.size, h


def _fill_item(b: BatchItem, it: dict, grid_npix: int):
    """Point a BatchItem at one problem's arrays; returns (inputs to keep alive, outputs)."""
    p, npix3, kp = _ptr(it["pix"], np.int32)
    w, _, kw = _ptr(it["weights"], np.float64)
    s, _, ks = _ptr(it["signal"], np.float64)
    nz, _, kn = _ptr(it["noise"], np.float64)
    inv, _, ki = _ptr(it["invtt"], np.float64)
    bs = _c(it["block_sizes"], np.int32)
    m = npix3 // 3
    cap = 3 * min(int(grid_npix), m)
    x = np.zeros(max(cap, 3), np.float64)
    px = np.zeros(max(cap // 3, 1), np.int32)
    cond = np.zeros(max(cap // 3, 1), np.float64)
    hits = np.zeros(max(cap // 3, 1), np.int32)
    b.m, b.pix, b.pixweights, b.signal, b.noise = m, p, w, s, nz
    b.nb, b.local_blocks_sizes, b.lam, b.invtt = bs.size, bs.ctypes.data, int(it["lam"]), inv
    b.x, b.pixels, b.cond, b.lhits, b.x_cap = x.ctypes.data, px.ctypes.data, cond.ctypes.data, hits.ctypes.data, x.size
    return (kp, kw, ks, kn, bs, ki), (x, px, cond, hits)


def _item_result(b: BatchItem, outs) -> BatchResult:
    x, px, cond, hits = outs
    r = b.rep
    k = r.n_map // 3
    return BatchResult(x=x[: 3 * k].copy(), pixels=px[:k].copy(), cond=cond[:k].copy(), hits=hits[:k].copy(),
                       report=r.as_dict())


class Server:
    """Request loop of mappraiser_gpu_server_*: submit() returns a ticket,
    wait(ticket) returns (BatchResult, latency_ms, queue_ms)."""

    def __init__(self, hbm_budget_gb=0.0, max_slots=1, max_batch=16, opts: Options | None = None):
        L = load()
        self._h = _vp()
        self._opts = opts or options()
        _check(L.mappraiser_gpu_server_create(ct.byref(self._h), float(hbm_budget_gb), int(max_slots),
                                              int(max_batch), ct.byref(self._opts)), "server_create")
        self._live = {}

    def submit(self, item: dict, grid_npix: int, tol=1e-8, maxiter=500) -> int:
        b = BatchItem()
        keep, outs = _fill_item(b, item, grid_npix)
        t = ct.c_longlong(0)
        _check(load().mappraiser_gpu_server_submit(self._h, ct.byref(b), int(grid_npix), float(tol), int(maxiter),
                                                   ct.byref(t)), "server_submit")
        self._live[t.value] = (b, keep, outs)
        return t.value

    def wait(self, ticket: int):
        lat, q = ct.c_double(0), ct.c_double(0)
        rc = load().mappraiser_gpu_server_wait(self._h, int(ticket), ct.byref(lat), ct.byref(q))
        b, keep, outs = self._live.pop(ticket)
        _check(rc, "server_wait")
        return _item_result(b, outs), lat.value, q.value

    def stats(self) -> dict:
        a, g, pk, pr, rq = ct.c_longlong(0), ct.c_longlong(0), ct.c_double(0), ct.c_int(0), ct.c_longlong(0)
        _check(load().mappraiser_gpu_server_stats(self._h, ct.byref(a), ct.byref(g), ct.byref(pk), ct.byref(pr),
                                                  ct.byref(rq)), "server_stats")
        return {"admitted": a.value, "groups": g.value, "peak_hbm_gb": pk.value, "peak_running": pr.value,
                "requeued": rq.value}

    def close(self):
        if self._h:
            load().mappraiser_gpu_server_destroy(self._h)
            self._h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def noise_invtt(noise, block_sizes, lam, fsamp, apod_window=MGPU_APOD_CHEBWIN):
    """compute_invtt on the device: returns (invtt [nb*lam], fit params [nb, 4] = (a, alpha, fknee, fmin))."""
    nz = _c(noise, np.float64)
    bs = _c(block_sizes, np.int32)
    out = np.zeros(bs.size * int(lam), np.float64)
    par = np.zeros((bs.size, 4), np.float64)
    _check(load().mappraiser_gpu_noise_invtt(nz.ctypes.data, bs.size, bs, int(lam), float(fsamp), int(apod_window), 0,
                                             out.ctypes.data, par.ctypes.data), "noise_invtt")
    return out, par
