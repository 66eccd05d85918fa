"""Writes tests/golden/fft_golden.npz: inputs and numpy.fft outputs (complex128)
for small 1-D and 2-D transforms, used to pin the oracle's FFT against a
library routine.  numpy convention: fft = exp(-i), ifft = exp(+i)/n; the
fixtures store the unnormalised backward transform (ifft * n).

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for (ny, nx) in [(4, 4), (8, 8), (12, 10), (7, 7), (30, 30), (11, 13), (16, 48), (25, 9)]:
        a = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
        out[f"x2_{ny}x{nx}"] = a
        out[f"f2_{ny}x{nx}"] = np.fft.fft2(a)
        out[f"b2_{ny}x{nx}"] = np.fft.ifft2(a) * (nx * ny)
    for n in [2, 3, 5, 7, 16, 17, 60, 64, 360]:
        a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"x1_{n}"] = a
        out[f"f1_{n}"] = np.fft.fft(a)
        out[f"b1_{n}"] = np.fft.ifft(a) * n
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


if __name__ == "__main__":
    main()
