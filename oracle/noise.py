"""CPU restatement of MAPPRAISER's noise-model builder -- test infrastructure
only (the oracle of mappraiser_gpu_noise_invtt; the product never calls it).

compute_invtt / noise2invtt (TOAST_interface/utils.py:435-601): for each
Toeplitz block, a one-segment Welch periodogram (utils.py:524-528), a bounded
fit of the log-PSD model (utils.py:531-538, model utils.py:485-486), the
inverse PSD of the fit on rfftfreq(n) (utils.py:550-554), its inverse FFT
(utils.py:557), and the apodisation window's first lambda points
(utils.py:560-584).  The periodogram, the fit and the windows are the scipy
routines the reference calls, with the reference's arguments (scipy is a
third-party dependency of the reference, not vendored; this image has scipy
1.15).
"""
from __future__ import annotations

import numpy as np


def logpsd_model(f, a, alpha, fknee, fmin):
    # utils.py:485-486
    return a + np.log10(1 + ((f + fmin) / fknee) ** alpha)


def inversepsd_model(f, sigma, alpha, fknee, fmin):
    # utils.py:493-494
    return sigma * 1.0 / (1 + ((f + fmin) / fknee) ** alpha)


def apod_window(lam, kind="chebwin"):
    """ifftshift(get_window(kind, 2 lambda))[:lambda] (utils.py:563-584)."""
    from scipy.signal import get_window

    if kind == "chebwin":
        w = get_window(("chebwin", 150), 2 * lam)
    elif kind == "gaussian":
        w = get_window(("general_gaussian", 1, 1 / 3 * lam), 2 * lam)
    else:
        raise ValueError(kind)
    return np.fft.ifftshift(w)[:lam]


def periodogram(x, fsamp):
    """Welch, one segment of len(x) samples, Hann window, linear detrend (utils.py:524-528)."""
    from scipy.signal import welch

    return welch(x, fsamp, window="hann", nperseg=x.size, detrend="linear")


def noise2invtt(x, lam, fsamp, kind="chebwin"):
    """First row (lambda entries) of the inverse noise band of one block, and the fit parameters."""
    from scipy.optimize import curve_fit

    n = x.size
    f, psd = periodogram(x, fsamp)
    popt, _ = curve_fit(logpsd_model, f[1:], np.log10(psd[1:]), p0=np.array([-7, -1.0, 0.1, 1e-6]),
                        bounds=([-20, -10, 0.0, 0.0], [0.0, 0.0, 20, 1]), maxfev=1000)
    f_full = np.fft.rfftfreq(n, 1.0 / fsamp)
    psdm1 = inversepsd_model(f_full, 10 ** (-popt[0]), *popt[1:])
    inv_tt = np.fft.irfft(psdm1, n=n)
    if lam > n // 2:
        raise RuntimeError("Bandwidth cannot be larger than timestream.")
    return np.multiply(apod_window(lam, kind), inv_tt[:lam]), popt


def compute_invtt(noise, block_sizes, lam, fsamp, kind="chebwin"):
    """invtt [nb*lambda] and fit parameters [nb, 4] of every block (utils.py:435-477)."""
    out, par, off = [], [], 0
    for n in block_sizes:
        row, p = noise2invtt(np.asarray(noise[off:off + n], np.float64), lam, fsamp, kind)
        out.append(row)
        par.append(p)
        off += int(n)
    return np.concatenate(out), np.array(par)
