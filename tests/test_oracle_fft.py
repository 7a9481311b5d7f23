"""Pins the oracle's FFT (replacement of FFTW 2.1.5 rfftwnd) against numpy
fixtures, a direct DFT, Parseval and inverse(forward) = N * identity."""
import os

import numpy as np
import pytest

import _oracle as O

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden",
                           "fft_fixtures.npz"))


def direct_dft(x, sign=-1):
    n = len(x)
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ x


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 9, 12, 15, 16, 20, 25, 27,
                               30, 32, 48, 60, 64, 96, 125, 128])
def test_fft1d_fixture(n):
    x, X = FIX["x1_%d" % n], FIX["X1_%d" % n]
    assert np.abs(O.fft1d(x) - X).max() <= 1e-13 * np.abs(X).max()
    assert np.abs(O.fft1d(X, +1) / n - x).max() <= 1e-13 * np.abs(x).max()


@pytest.mark.parametrize("n", [7, 11, 14, 21])
def test_fft1d_direct(n):
    x = np.random.default_rng(n).standard_normal(n) * (1 + 0.5j)
    assert np.abs(O.fft1d(x) - direct_dft(x)).max() < 1e-12
    assert np.abs(O.fft1d(x, 1) - direct_dft(x, 1)).max() < 1e-12


@pytest.mark.parametrize("shape", [(8, 8, 8), (12, 12, 12), (20, 20, 20),
                                   (8, 12, 10), (6, 10, 16)])
def test_rfft3d_fixture(shape):
    key = "x".join(map(str, shape))
    r, c = FIX["r_" + key], FIX["c_" + key]
    co = O.rfft3d(r)
    assert np.abs(co - c).max() <= 1e-13 * np.abs(c).max()
    back = O.irfft3d(co, shape[2]) / r.size
    assert np.abs(back - r).max() <= 1e-13 * np.abs(r).max()
    # Parseval on the full spectrum rebuilt from the half spectrum
    full = np.fft.fftn(r)
    assert np.isclose((np.abs(full) ** 2).sum() / r.size, (r ** 2).sum())
