"""Pins the oracle's FFT (oracle/jms_oracle.cpp, own Stockham transform) to
the DFT definition (brute force in float64), to numpy.fft fixtures
(tests/golden/fft_golden.npz) and to the transform's invariants, in the
reference's convention: forward exp(-2 pi i nk/N), backward exp(+...),
unscaled (JFFTMKLcore.cpp:117-221, JFFTCUDAcore.cpp:299-336)."""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))


def brute_dft2(a, sign):
    ny, nx = a.shape
    ky = np.exp(sign * 2j * np.pi * np.outer(np.arange(ny), np.arange(ny)) / ny)
    kx = np.exp(sign * 2j * np.pi * np.outer(np.arange(nx), np.arange(nx)) / nx)
    return ky @ a @ kx.T


@pytest.mark.parametrize("ny,nx", [(4, 4), (8, 8), (12, 10), (7, 7), (30, 30), (11, 13), (3, 17), (2, 1)])
def test_fft2_brute_force(ny, nx):
    rng = np.random.default_rng(ny * 100 + nx)
    a = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
    for sign, d in ((-1, -1), (+1, +1)):
        ref = brute_dft2(a, sign)
        r64 = O.fft2(a, d, prec=64)
        r32 = O.fft2(a, d, prec=32)
        s = np.linalg.norm(ref)
        assert np.linalg.norm(r64 - ref) / s < 1e-13
        assert np.linalg.norm(r32 - ref) / s < 5e-7


def test_fft_numpy_golden():
    for k in GOLD.files:
        if k.startswith("x2_"):
            tag = k[3:]
            a = GOLD[k]
            for key, d in (("f2_", -1), ("b2_", +1)):
                ref = GOLD[key + tag]
                e = np.linalg.norm(O.fft2(a, d, prec=64) - ref) / np.linalg.norm(ref)
                assert e < 1e-13, (tag, key, e)
                e32 = np.linalg.norm(O.fft2(a, d, prec=32) - ref) / np.linalg.norm(ref)
                assert e32 < 5e-7, (tag, key, e32)
        if k.startswith("x1_"):
            tag = k[3:]
            a = GOLD[k]
            for key, d in (("f1_", -1), ("b1_", +1)):
                ref = GOLD[key + tag]
                e = np.linalg.norm(O.fft1(a, d, prec=64) - ref) / np.linalg.norm(ref)
                assert e < 1e-13, (tag, key, e)


@pytest.mark.parametrize("ny,nx", [(64, 64), (256, 256), (48, 360), (49, 25)])
def test_fft_invariants(ny, nx):
    rng = np.random.default_rng(7)
    a = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx)))
    n = nx * ny
    F = O.fft2(a, -1, prec=64)
    # Parseval
    assert abs(np.sum(np.abs(F) ** 2) - n * np.sum(np.abs(a) ** 2)) / (n * np.sum(np.abs(a) ** 2)) < 1e-12
    # backward(forward(x)) = n x
    B = O.fft2(F, +1, prec=64)
    assert np.abs(B - n * a).max() / np.abs(n * a).max() < 1e-12
    # delta <-> constant
    d = np.zeros((ny, nx), complex)
    d[0, 0] = 1.0
    assert np.allclose(O.fft2(d, -1, prec=64), 1.0, atol=1e-14)
    # single plane wave -> single bin
    kx0, ky0 = 3 % nx, 5 % ny
    yy, xx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    pw = np.exp(2j * np.pi * (kx0 * xx / nx + ky0 * yy / ny))
    P = O.fft2(pw, -1, prec=64)
    assert abs(P[ky0, kx0] - n) < 1e-9 * n
    P[ky0, kx0] = 0
    assert np.abs(P).max() < 1e-9 * n
    # shift theorem: x(r - s) <-> X(k) exp(-2 pi i k.s / N)
    sx, sy = 5, 2
    S = O.fft2(np.roll(np.roll(a, sy, axis=0), sx, axis=1), -1, prec=64)
    ph = np.exp(-2j * np.pi * (sx * xx / nx + sy * yy / ny))
    assert np.abs(S - F * ph).max() / np.abs(F).max() < 1e-12


def test_fft_mutation_is_detected():
    """A transposed or sign-flipped transform fails the brute-force check."""
    rng = np.random.default_rng(3)
    a = rng.standard_normal((6, 10)) + 1j * rng.standard_normal((6, 10))
    ref = brute_dft2(a, -1)
    good = O.fft2(a, -1, prec=64)
    wrong_sign = O.fft2(a, +1, prec=64)
    assert np.linalg.norm(good - ref) < 1e-9
    assert np.linalg.norm(wrong_sign - ref) > 1.0
