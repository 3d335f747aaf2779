"""Pins of the noise-model oracle (oracle/noise.py) on the CPU.

The fit is pinned by recovery of known PSD parameters from long seeded
timestreams; the inverse-noise row by the closed form for the true PSD; the
windows and the periodogram by their definitions."""
from __future__ import annotations

import numpy as np
import pytest

from midapack_amd import workloads as W
from oracle import noise as N


def test_periodogram_of_white_noise_is_flat():
    x = np.random.default_rng(1).standard_normal(1 << 16)
    f, p = N.periodogram(x, 10.0)
    # one-sided density of unit white noise sampled at 10 Hz: 2 / fs on average
    assert f.size == (1 << 15) + 1 and abs(np.mean(p[1:-1]) - 0.2) < 0.01


@pytest.mark.parametrize("fknee,alpha", [(0.1, 1.0), (0.3, 1.7)])
def test_fit_recovers_known_psd(fknee, alpha):
    n = 1 << 18
    x = W.one_over_f_noise(np.random.default_rng(5), n, 1.0, fknee, alpha)
    _, popt = N.noise2invtt(x, 1024, 10.0)
    a, al, fk, fm = popt
    # one-sided density of sigma^2 = 1 at fs = 10 Hz is 2/fs; the mean of log10 of a
    # chi-square(2)/2 periodogram ordinate is -gamma/ln 10; alpha_ref = -alpha; fmin ~ 0
    a_true = np.log10(2 / 10.0) - np.euler_gamma / np.log(10)
    assert abs(a - a_true) < 0.02 and abs(-al - alpha) < 0.05 and abs(fk - fknee) / fknee < 0.1 and fm < 1e-3


def test_invtt_of_true_model_matches_closed_form():
    # inverse PSD of the exact model: the oracle's steps after the fit equal invtt_block
    n, lam, fs = 1 << 16, 1024, 10.0
    f_full = np.fft.rfftfreq(n, 1 / fs)
    psdm1 = N.inversepsd_model(f_full, 1.0, -1.0, 0.1, 0.0)
    psdm1[0] = 0.0  # the generator's convention at f = 0
    row = N.apod_window(lam) * np.fft.irfft(psdm1, n=n)[:lam]
    np.testing.assert_allclose(row, W.invtt_block(n, lam, 1.0, 0.1, 1.0), rtol=0, atol=1e-15)


def test_windows():
    w = N.apod_window(64)
    assert w.size == 64 and abs(w[0] - 1.0) < 1e-12 and np.all(np.diff(w) <= 1e-12)
    g = N.apod_window(64, "gaussian")
    np.testing.assert_allclose(g, np.exp(-0.5 * (np.arange(64) / (64 / 3)) ** 2), rtol=1e-12)
