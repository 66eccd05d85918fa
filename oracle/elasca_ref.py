"""CPU restatement of the reference's phase-grating generation from projected
atomic form factors, for checks only (test infrastructure: only tests/ may
use it).

Follows CJEElaSca::CalculatePhasegratingGPU (JEElaSca.cpp:978-1159) for
neutral atoms, with its kernels AddFormFactor2dKernel and PotToPgrKernel
(scapot.cu:29-41,80-89) and the helpers of CJEElaSca::InitHelper
(JEElaSca.cpp:486-633):
  q = fidx(i) / a in FFT order, fidx(i) = ((i + n//2) % n) - n//2 (:529-541)
  V(q) = sum_t sum_{atoms of t} ff_t(q) ap(q) occ exp(-2 pi i (x qx + y qy))
  V(r) = IFT[V(q)] * POT_SCA / vol (unnormalised inverse transform, :1126-1129)
  t(r) = exp(i PGR_SIG cz wl V(r)) (:1141-1142)
  band limit: t = IFT[FT[t] * ap2], ap2 = 1/(nx ny) for |q| <= 2/3 kmax, else 0
  (:1149-1153; aperture :597-604)
The sums run in float64; the frequencies and the aperture edges are formed
in float32 as the reference forms them, so that the same pixels fall inside
the band limits.

Pinned by closed forms in tests/test_elasca_ref.py: a point-like form factor
(ff = 1) gives n * POT_SCA / vol at the atom's pixel and zero elsewhere (shift
theorem, axes, sign, unnormalised transform); a Gaussian form factor
exp(-2 pi^2 s^2 q^2) gives POT_SCA / cz times the periodic sum of the
normalised 2-D Gaussian of width s (scale); the grating phase is
wl / (4 pi) times that sum, which pins POT_SCA * PGR_SIG = 1 / (4 pi); an
imaginary form factor damps |t| by the same closed form.
"""
from __future__ import annotations

import numpy as np

# JEElaSca.h:88,91,94
POT_SCA = 0.03809982
PGR_BWL = 0.6666667
PGR_SIG = 2.08865737


def freqs(n: int, a: float) -> np.ndarray:
    """Spatial frequencies [1/nm] of an n-point grid over a [nm], FFT order,
    float32 (InitHelper, JEElaSca.cpp:529-541)."""
    h = (n - n % 2) >> 1
    i = np.arange(n)
    fi = ((i + h) % n) - h
    return fi.astype(np.float32) * (np.float32(1.0) / np.float32(a))


def helpers(nx: int, ny: int, ax: float, ay: float):
    """|q| (float32, (ny, nx)), kmax and the two band-limit apertures
    (potential: tanh edge; phase grating: 2/3 kmax with weight 1/(nx ny)),
    JEElaSca.cpp:586-607."""
    qx, qy = freqs(nx, ax), freqs(ny, ay)
    g2 = (qx * qx)[None, :] + (qy * qy)[:, None]
    gm = np.sqrt(g2.astype(np.float32))
    itx, ity = np.float32(1.0) / np.float32(ax), np.float32(1.0) / np.float32(ay)
    kmax = min(np.float32(nx // 2) * itx, np.float32(ny // 2) * ity)
    ap_pot = 0.5 - 0.5 * np.tanh((gm.astype(np.float64) / float(kmax) - 0.9) * 30.0)
    kthr = np.float32(kmax) * np.float32(PGR_BWL)
    ap_pgr = np.where(gm > kthr, 0.0, 1.0 / nx / ny)
    return qx, qy, gm, kmax, ap_pot, ap_pgr


def fourier_potential(nx, ny, ax, ay, ff2d, xy, typ, occ=None, bwl_pot=False) -> np.ndarray:
    """V(q) of one atom set, (ny, nx) complex128 (JEElaSca.cpp:1088-1122 with
    AddFormFactor2dKernel): types outer, atoms inner."""
    qx, qy, _, _, ap_pot, _ = helpers(nx, ny, ax, ay)
    ap = ap_pot if bwl_pot else 1.0
    xy = np.asarray(xy, np.float64).reshape(-1, 2)
    typ = np.asarray(typ).reshape(-1)
    occ = np.ones(len(typ)) if occ is None else np.asarray(occ, np.float64)
    pot = np.zeros((ny, nx), np.complex128)
    for t in range(len(ff2d)):
        sel = np.nonzero(typ == t)[0]
        if len(sel) == 0:
            continue
        ex = np.exp(-2j * np.pi * xy[sel, 0][:, None] * qx.astype(np.float64)[None, :])  # (k, nx)
        ey = np.exp(-2j * np.pi * xy[sel, 1][:, None] * qy.astype(np.float64)[None, :])  # (k, ny)
        s = (ey * occ[sel][:, None]).T @ ex  # (ny, nx): sum_k occ ey ex
        pot += np.asarray(ff2d[t], np.complex128) * ap * s
    return pot


def real_potential(nx, ny, ax, ay, cz, ff2d, xy, typ, occ=None, bwl_pot=False) -> np.ndarray:
    """V(r) [V nm] of one atom set: unnormalised inverse transform times
    POT_SCA / vol (JEElaSca.cpp:1126-1129), vol = ax ay cz."""
    vq = fourier_potential(nx, ny, ax, ay, ff2d, xy, typ, occ, bwl_pot)
    return np.fft.ifft2(vq) * (nx * ny) * (POT_SCA / (ax * ay * cz))


def phase_grating(nx, ny, ax, ay, cz, wl, ff2d, xy, typ, occ=None, bwl_pot=False, bwl_pgr=True) -> np.ndarray:
    """Transmission function t(r) of one atom set (JEElaSca.cpp:1136-1156)."""
    v = real_potential(nx, ny, ax, ay, cz, ff2d, xy, typ, occ, bwl_pot)
    t = np.exp(1j * (PGR_SIG * cz * wl) * v)
    if bwl_pgr:
        ap_pgr = helpers(nx, ny, ax, ay)[5]
        t = np.fft.ifft2(np.fft.fft2(t) * ap_pgr) * (nx * ny)
    return t
