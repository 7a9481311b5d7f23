"""Writes the FFT golden fixtures with numpy.fft (pocketfft, FP64).

The reference ships no test vectors (its FFT is FFTW 2.1.5, not present
here), so the FFT part of the oracle is pinned against numpy's independent
implementation on small grids of every radix the solver uses (2, 3, 5) and
against a direct O(N^2) DFT in the tests.

    python tests/golden/make_fixtures.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for shape in [(8, 8, 8), (12, 12, 12), (20, 20, 20), (8, 12, 10),
                  (6, 10, 16)]:
        r = rng.standard_normal(shape)
        c = np.fft.rfftn(r)
        key = "x".join(map(str, shape))
        out["r_" + key] = r
        out["c_" + key] = c
    for n in [2, 3, 4, 5, 6, 8, 9, 12, 15, 16, 20, 25, 27, 30, 32, 48, 60,
              64, 96, 125, 128]:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out["x1_%d" % n] = x
        out["X1_%d" % n] = np.fft.fft(x)
    np.savez_compressed(os.path.join(HERE, "fft_fixtures.npz"), **out)


if __name__ == "__main__":
    main()
