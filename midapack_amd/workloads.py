"""Synthetic GLS map-making workloads (seeded, reproducible on any host).

Each workload is one rank's share of time-ordered data (TOD) in the exact
layout MAPPRAISER's ``MLmap`` receives from its TOAST operator:

* ``pix``      int32[m*3]   pixel index encoded ``3p+k`` (mappraiser.py:1214-1216),
                            flagged samples ``-3,-2,-1`` (utils.py:148-150)
* ``weights``  float64[m*3] (1, cos 2psi, sin 2psi)
* ``signal``, ``noise``  float64[m]
* ``block_sizes`` int32[nb], ``invtt`` float64[nb*lambda] (utils.py:416-481)

Scan: boustrophedon raster over an Nx x Ny flat grid with per-detector seeded
offsets; psi ~ U[0, pi).  Noise: seeded 1/f timestreams with PSD
sigma^2 (1 + (fknee/f)^alpha).  invtt: the analytic form of ``noise2invtt``
(utils.py:500-601): irfft of the inverse PSD over the block length, apodised
by ifftshift(chebwin(2 lambda, 150 dB))[:lambda].

numpy's PCG64 generator is seeded per (workload, detector, stream), so the
same inputs are regenerated bit-for-bit on every host with this image.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

import numpy as np

FSAMP = 10.0  # Hz


@dataclass
class Workload:
    name: str
    pix: np.ndarray
    weights: np.ndarray
    signal: np.ndarray
    noise: np.ndarray
    block_sizes: np.ndarray
    lam: int
    invtt: np.ndarray
    nx: int
    ny: int
    sky: np.ndarray = field(repr=False)   # true I,Q,U map, float64[npix_grid*3]
    meta: dict = field(default_factory=dict)

    @property
    def m(self) -> int:
        return int(self.signal.size)


def _rng(seed, *keys):
    return np.random.default_rng([int(seed)] + [int(k) for k in keys])


def raster_pixels(nsamp, nx, ny, ox, oy):
    """Boustrophedon raster: row advances every nx samples, odd passes reversed."""
    t = np.arange(nsamp, dtype=np.int64)
    rpass = t // nx
    c0 = t % nx
    c = np.where(rpass % 2 == 0, c0, nx - 1 - c0)
    row = (rpass + oy) % ny
    col = (c + ox) % nx
    return (row * nx + col).astype(np.int64)


def _fast_len(n):
    """Smallest 5-smooth length >= n: the FFT-derived fields of a length with a
    large prime factor are computed on this length (an FFT of a prime length
    costs ~10x more), then truncated."""
    from scipy.fft import next_fast_len

    return int(next_fast_len(int(n), real=True))


def one_over_f_noise(rng, nsamp, sigma, fknee, alpha, fsamp=FSAMP):
    """sigma * irfft(rfft(w) sqrt(1 + (fknee/f)^alpha)), w ~ N(0,1), on the fast
    length >= nsamp, truncated to nsamp."""
    n = _fast_len(nsamp)
    w = rng.standard_normal(n)
    f = np.fft.rfftfreq(n, 1.0 / fsamp)
    amp = np.zeros_like(f)
    amp[1:] = np.sqrt(1.0 + (fknee / f[1:]) ** alpha)
    return sigma * np.fft.irfft(np.fft.rfft(w) * amp, n=n)[:nsamp]


def invtt_block(nb_samp, lam, sigma, fknee, alpha, fsamp=FSAMP):
    """Analytic noise2invtt (utils.py:550-580): first row of the inverse-noise band
    (the inverse PSD sampled on the fast length >= the block length)."""
    if lam > nb_samp // 2:
        raise ValueError("Bandwidth cannot be larger than timestream (utils.py:561).")
    n = _fast_len(nb_samp)  # = nb_samp for the power-of-two blocks of C1/C2
    f = np.fft.rfftfreq(n, 1.0 / fsamp)
    psdm1 = np.zeros_like(f)
    psdm1[1:] = 1.0 / (sigma**2 * (1.0 + (fknee / f[1:]) ** alpha))
    inv_tt = np.fft.irfft(psdm1, n=n)
    if lam == 1:
        return inv_tt[:1].copy()
    return np.multiply(_chebwin_half(lam), inv_tt[:lam], dtype=np.float64)


_CHEB = {}


def _chebwin_half(lam):
    """ifftshift(chebwin(2 lambda, 150 dB))[:lambda] (utils.py:567-570), cached per lambda."""
    from scipy.signal import get_window

    if lam not in _CHEB:
        _CHEB[lam] = np.fft.ifftshift(get_window(("chebwin", 150), 2 * lam))[:lam].copy()
    return _CHEB[lam]


def make_workload(name, ndet, nsamp, nx, ny, lam, seed, block_len=None, fknee=0.1, alpha=1.0,
                  sigma=1.0, fknee_range=None, alpha_range=None, det_lengths=None, flag_frac=0.0,
                  crosslink=False):
    """Build one rank's data.  ``det_lengths`` overrides ``nsamp`` per detector;
    ``block_len`` splits every detector into Toeplitz blocks of that length
    (the last block takes the remainder); all blocks of a detector share its T.
    ``crosslink`` makes every odd detector scan along columns instead of rows."""
    lens = [int(nsamp)] * ndet if det_lengths is None else [int(v) for v in det_lengths]
    m = int(sum(lens))
    npix_grid = nx * ny
    rs = _rng(seed, 0)
    sky = np.empty(npix_grid * 3)
    sky[0::3] = rs.standard_normal(npix_grid)
    sky[1::3] = 0.1 * rs.standard_normal(npix_grid)
    sky[2::3] = 0.1 * rs.standard_normal(npix_grid)

    pix = np.empty(m * 3, np.int32)
    weights = np.empty(m * 3, np.float64)
    signal = np.empty(m, np.float64)
    noise = np.empty(m, np.float64)
    bsizes, invtt = [], []
    t0 = 0
    for d, n_d in enumerate(lens):
        rd = _rng(seed, 1, d)
        ox, oy = int(rd.integers(0, nx)), int(rd.integers(0, ny))
        fk = fknee if fknee_range is None else float(rd.uniform(*fknee_range))
        al = alpha if alpha_range is None else float(rd.uniform(*alpha_range))
        if crosslink and d % 2 == 1:
            pt = raster_pixels(n_d, ny, nx, oy, ox)  # raster of the transposed grid
            p = (pt % ny) * nx + pt // ny
        else:
            p = raster_pixels(n_d, nx, ny, ox, oy)
        psi = _rng(seed, 2, d).random(n_d) * np.pi
        c2, s2 = np.cos(2 * psi), np.sin(2 * psi)
        sl = slice(t0, t0 + n_d)
        base = 3 * p
        pix[3 * t0: 3 * (t0 + n_d): 3] = base
        pix[3 * t0 + 1: 3 * (t0 + n_d): 3] = base + 1
        pix[3 * t0 + 2: 3 * (t0 + n_d): 3] = base + 2
        weights[3 * t0: 3 * (t0 + n_d): 3] = 1.0
        weights[3 * t0 + 1: 3 * (t0 + n_d): 3] = c2
        weights[3 * t0 + 2: 3 * (t0 + n_d): 3] = s2
        signal[sl] = sky[base] + c2 * sky[base + 1] + s2 * sky[base + 2]
        noise[sl] = one_over_f_noise(_rng(seed, 3, d), n_d, sigma, fk, al)
        if flag_frac > 0:
            fl = _rng(seed, 4, d).random(n_d) < flag_frac
            idx = np.nonzero(fl)[0] + t0
            for k in range(3):
                pix[3 * idx + k] = k - 3
        L = n_d if block_len is None else int(block_len)
        nblk = max(1, (n_d + L - 1) // L)
        sizes = [L] * (nblk - 1) + [n_d - L * (nblk - 1)]
        T = invtt_block(L if nblk > 1 else n_d, lam, sigma, fk, al)
        for s in sizes:
            bsizes.append(s)
            invtt.append(T)
        t0 += n_d
    return Workload(name=name, pix=pix, weights=weights, signal=signal, noise=noise,
                    block_sizes=np.asarray(bsizes, np.int32), lam=int(lam),
                    invtt=np.concatenate(invtt).astype(np.float64), nx=nx, ny=ny, sky=sky,
                    meta=dict(ndet=ndet, m=m, seed=seed))


# ----------------------------------------------------------------------------
# Named configurations (one rank's share each)
# ----------------------------------------------------------------------------
def config_c1(seed=1):
    """4 det x 65,536 samples, 128x128, lambda=1024, one block per detector."""
    return make_workload("C1: 4x65536 on 128^2, lambda=1024, FP64", 4, 65536, 128, 128, 1024, seed)


def config_c2(seed=2, ndet=64, nsamp=1 << 20):
    """64 det x 2^20 samples, 1024x1024, lambda=8192, 2^17-sample blocks,
    fknee ~ U[0.05,0.5] Hz, alpha ~ U[1,2] per detector."""
    return make_workload(f"C2: {ndet}x{nsamp} on 1024^2, lambda=8192, 2^17 blocks", ndet, nsamp, 1024, 1024,
                         8192, seed, block_len=1 << 17, fknee_range=(0.05, 0.5), alpha_range=(1.0, 2.0))


def config_small(seed=7, ndet=2, nsamp=4096, nx=16, ny=16, lam=64, block_len=None, flag_frac=0.0,
                 det_lengths=None):
    """Small case for parity tests (oracle runs in well under a second)."""
    return make_workload(f"small {ndet}x{nsamp} {nx}x{ny} lambda={lam}", ndet, nsamp, nx, ny, lam, seed,
                         block_len=block_len, flag_frac=flag_frac, det_lengths=det_lengths)


def config_c4(seed=4, nreal=32, nx=512, ny=512, progress=None):
    """Batched Monte Carlo: nreal independent realizations on a shared 512^2
    grid, each with U{16..64} detectors, per-detector lengths floor(2^U(16,20))
    (one Toeplitz block per detector), lambda from {1024, 4096, 8192},
    fknee ~ U[0.05, 0.5] Hz, alpha ~ U[1, 2]."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    def one(r):
        rr = _rng(seed, 100, r)
        ndet = int(rr.integers(16, 65))
        lens = np.floor(2.0 ** rr.uniform(16.0, 20.0, ndet)).astype(np.int64)
        lam = int(rr.choice([1024, 4096, 8192]))
        wl = make_workload(f"C4 r{r}: {ndet} det, {int(lens.sum())} samples, lambda={lam}", ndet, 0, nx, ny, lam,
                           seed * 1000 + r, det_lengths=lens, fknee_range=(0.05, 0.5), alpha_range=(1.0, 2.0))
        if progress:
            progress(wl.name)
        return wl

    # realizations are independent and numpy releases the GIL in its array kernels
    with ThreadPoolExecutor(max(1, min(8, os.cpu_count() or 1))) as ex:
        return list(ex.map(one, range(nreal)))


def c5_request(seed, idx, max_det=128):
    """One serving request: U{4..max_det} detectors x floor(2^U(16,20)) samples
    (one block per detector), grid side from {256, 512, 1024}, lambda from
    {1024, 4096, 8192}, seeded per request."""
    rr = _rng(seed, 200, idx)
    ndet = int(rr.integers(4, max_det + 1))
    nsamp = int(np.floor(2.0 ** rr.uniform(16.0, 20.0)))
    side = int(rr.choice([256, 512, 1024]))
    lam = int(rr.choice([1024, 4096, 8192]))
    return make_workload(f"C5 req{idx}: {ndet}x{nsamp} on {side}^2, lambda={lam}", ndet, nsamp, side, side, lam,
                         seed * 100000 + idx, fknee_range=(0.05, 0.5), alpha_range=(1.0, 2.0))


def c5_pool(seed, n, max_det=128, progress=None):
    """Requests 0..n-1 of c5_request, generated on a thread pool."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    def one(i):
        wl = c5_request(seed, i, max_det)
        if progress:
            progress(wl.name)
        return wl

    with ThreadPoolExecutor(max(1, min(8, os.cpu_count() or 1))) as ex:
        return list(ex.map(one, range(n)))


# ----------------------------------------------------------------------------
# Template map-making (MTmap) inputs
# ----------------------------------------------------------------------------
def make_templates_workload(name, nces, ndet, ces_len, nx, ny, seed, npoly=3, ground=1, n_sss_bins=16, nhwp=2,
                            delta_t=20.0, fhwp=1.3, sigma=1.0, flag_frac=0.0, amp=1.0):
    """One rank's MTmap inputs (mappraiser.c:387-392): nces constant-elevation
    scans of ndet detectors, blocks CES-major, each as long as its CES.

    The boresight sweeps back and forth along the nx columns (one sweep = nx
    samples, the row advances after every sweep); az is the boresight column
    (+0.5, az range [0, nx]); detector j adds a seeded (row, column) offset.
    The HWP turns at fhwp Hz.  signal = sky + templates with seeded amplitudes
    (sweep offsets and Legendre drifts, a ground profile in az, HWP harmonics
    per delta_t bin) scaled by amp; noise is white (sigma per detector), so
    invtt holds one weight 1/sigma^2 per block (lambda 1)."""
    fs = FSAMP
    m = nces * ndet * ces_len
    npix_grid = nx * ny
    rs = _rng(seed, 0)
    sky = np.empty(npix_grid * 3)
    sky[0::3] = rs.standard_normal(npix_grid)
    sky[1::3] = 0.1 * rs.standard_normal(npix_grid)
    sky[2::3] = 0.1 * rs.standard_normal(npix_grid)
    pix = np.empty(m * 3, np.int32)
    weights = np.empty(m * 3, np.float64)
    signal = np.empty(m, np.float64)
    noise = np.empty(m, np.float64)
    stamps, az, azmin, azmax, hwp = [], [], [], [], []
    bsizes, invtt = [], []
    t = np.arange(ces_len, dtype=np.int64)
    t0 = 0
    for c in range(nces):
        rc = _rng(seed, 10, c)
        oy = int(rc.integers(0, ny))
        sweep = t // nx
        col = np.where(sweep % 2 == 0, t % nx, nx - 1 - t % nx)
        row = (sweep + oy) % ny
        st = np.arange(0, ces_len, nx, dtype=np.int64)
        stamps.append(np.append(st, ces_len).astype(np.int32))
        a = col.astype(np.float64) + 0.5
        az.append(a)
        azmin.append(0.0)
        azmax.append(float(nx))
        ang = np.mod(2 * np.pi * fhwp * t / fs + rc.uniform(0, 2 * np.pi), 2 * np.pi)
        hwp.append(ang)
        for j in range(ndet):
            rd = _rng(seed, 11, c, j)
            dx, dy = int(rd.integers(0, nx)), int(rd.integers(0, ny))
            p = ((row + dy) % ny) * nx + (col + dx) % nx
            psi = rd.random(ces_len) * np.pi
            c2, s2 = np.cos(2 * psi), np.sin(2 * psi)
            sl = slice(t0, t0 + ces_len)
            base = 3 * p
            for k in range(3):
                pix[3 * t0 + k: 3 * (t0 + ces_len): 3] = base + k
            weights[3 * t0: 3 * (t0 + ces_len): 3] = 1.0
            weights[3 * t0 + 1: 3 * (t0 + ces_len): 3] = c2
            weights[3 * t0 + 2: 3 * (t0 + ces_len): 3] = s2
            sig = sky[base] + c2 * sky[base + 1] + s2 * sky[base + 2]
            # template contamination
            ns = len(st)
            offs = rd.standard_normal((max(npoly, 1), ns))
            for o in range(npoly):
                for s_ in range(ns):
                    a0, a1 = int(stamps[-1][s_]), int(stamps[-1][s_ + 1])
                    if a1 - a0 < 2:
                        continue
                    u = (2 * np.arange(a0, a1) / fs - a0 / fs - (a1 - 1) / fs) / ((a1 - 1) / fs - a0 / fs)
                    sig[a0:a1] += amp * offs[o, s_] * np.polynomial.legendre.legval(u, [0] * o + [1])
            if ground:
                prof = rd.standard_normal(n_sss_bins)
                sig += amp * prof[np.minimum(n_sss_bins - 1, np.floor(a / (nx / n_sss_bins)).astype(np.int64))]
            for h in range(1, nhwp + 1):
                sig += amp * (rd.standard_normal() * np.cos(h * ang) + rd.standard_normal() * np.sin(h * ang))
            sd = sigma * (1.0 + 0.2 * rd.random())
            signal[sl] = sig
            noise[sl] = sd * _rng(seed, 12, c, j).standard_normal(ces_len)
            if flag_frac > 0:
                fl = _rng(seed, 13, c, j).random(ces_len) < flag_frac
                idx = np.nonzero(fl)[0] + t0
                for k in range(3):
                    pix[3 * idx + k] = k - 3
            bsizes.append(ces_len)
            invtt.append(1.0 / sd ** 2)
            t0 += ces_len
    meta = dict(nces=nces, ndet=ndet, npoly=npoly, ground=ground, n_sss_bins=n_sss_bins, nhwp=nhwp,
                delta_t=float(delta_t), sampling_freq=fs, sweeptstamps=stamps, az=az, az_min=np.array(azmin),
                az_max=np.array(azmax), hwp_angle=hwp, m=m, seed=seed)
    return Workload(name=name, pix=pix, weights=weights, signal=signal, noise=noise,
                    block_sizes=np.asarray(bsizes, np.int32), lam=1, invtt=np.asarray(invtt, np.float64), nx=nx,
                    ny=ny, sky=sky, meta=meta)


def config_templates_small(seed=21, nces=2, ndet=2, ces_len=2048, nx=32, ny=16, **kw):
    """Small MTmap case for parity tests (the oracle solves it in about a second)."""
    return make_templates_workload(f"templates {nces}x{ndet}x{ces_len} {nx}x{ny}", nces, ndet, ces_len, nx, ny,
                                   seed, **kw)


def config_c6(seed=6):
    """Template map-making (MTmap) at the C2 scale: 8 CES x 8 detectors x 2^20
    samples on 1024^2, white noise (lambda 1), Legendre order 0 on every
    1024-sample sweep, 32 azimuth bins of ground pickup, the first HWP harmonic
    on 128 time bins per CES (1312 templates per detector and CES)."""
    return make_templates_workload("C6: MTmap 8 CES x 8 det x 2^20 on 1024^2, 1312 templates per block", 8, 8,
                                   1 << 20, 1024, 1024, seed, npoly=1, ground=1, n_sss_bins=32, nhwp=1,
                                   delta_t=(1 << 20) / FSAMP / 128)


def templates_irregular(wl, cut=1500, first=5):
    """A template workload with every CES after the first cut to `cut` samples and
    sweeps that start at sample `first` of each CES and include a one-sample
    sweep (samples before the first stamp carry no polynomial template)."""
    md = dict(wl.meta)
    nces, ndet = md["nces"], md["ndet"]
    L0 = int(wl.block_sizes[0])
    lens = [L0 if c == 0 else cut for c in range(nces)]
    keep = np.concatenate([np.arange(b * L0, b * L0 + lens[b // ndet]) for b in range(nces * ndet)])
    pix = wl.pix.reshape(-1, 3)[keep].ravel().copy()
    w = wl.weights.reshape(-1, 3)[keep].ravel().copy()
    stamps = []
    for c in range(nces):
        st = [first] + [int(v) for v in md["sweeptstamps"][c][1:] if v < lens[c]]
        st = sorted(set(st + [st[1] + 1 if len(st) > 1 else first + 1, lens[c]]))
        stamps.append(np.array(st, np.int32))
    md.update(sweeptstamps=stamps, az=[a[:n] for a, n in zip(md["az"], lens)],
              hwp_angle=[h[:n] for h, n in zip(md["hwp_angle"], lens)])
    return dataclasses.replace(wl, pix=pix, weights=w, signal=wl.signal[keep].copy(), noise=wl.noise[keep].copy(),
                               block_sizes=np.repeat(np.array(lens, np.int32), ndet), meta=md)
