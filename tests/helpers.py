"""Shared helpers of the test suite: small setups and error measures."""
import numpy as np

from jmultislicelib_amd import configs
from jmultislicelib_amd.configs import Setup


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def det_errors(g, r) -> np.ndarray:
    """Relative L2 error of every (detection plane, detector) pair: the last
    axis of a readout array indexes detectors, the one before it (if any)
    detection planes, and any axes in front (probe positions) are summed
    over.  A pair whose oracle values are all zero gets the largest absolute
    GPU value instead (it must be exactly zero too)."""
    g = np.asarray(g, np.float64)
    r = np.asarray(r, np.float64)
    assert g.shape == r.shape, (g.shape, r.shape)
    if g.ndim == 1:
        g, r = g[None, None], r[None, None]
    elif g.ndim == 2:
        g, r = g[None], r[None]
    g = g.reshape(-1, g.shape[-2], g.shape[-1])
    r = r.reshape(-1, r.shape[-2], r.shape[-1])
    num = np.sqrt(((g - r) ** 2).sum(0))
    den = np.sqrt((r ** 2).sum(0))
    zero = den == 0
    return np.where(zero, np.abs(g).max(0), num / np.where(zero, 1.0, den))


def assert_dets(g, r, tol, what=""):
    """Every detector at every detection plane within `tol` relative L2 of
    the oracle (see det_errors); returns the error table."""
    e = det_errors(g, r)
    print(f"{what} per-detector rel. L2 (plane x detector): {np.array2string(e, precision=2)}")
    assert np.all(e < tol), f"{what}: detector error {e.max():.3e} >= {tol}"
    return e


def assert_dets_noise_aware(g, r32, fp64, tol, what=""):
    """assert_dets, with one refinement for detectors whose signal sits at
    float32's rounding-noise floor (e.g. a 70-200 mrad ring reading 1e-9 of
    the incident intensity): a (plane, detector) pair that misses `tol`
    against the float32 oracle passes only if the float32 oracle itself is
    more than tol/4 away from its FP64 twin there (the pair is noise-limited
    in float32 on any implementation) and the GPU's error against the FP64
    twin is below 4x the float32 oracle's.  fp64() returns the FP64 twin's
    result (computed only when needed)."""
    e = det_errors(g, r32)
    print(f"{what} per-detector rel. L2 (plane x detector): {np.array2string(e, precision=2)}")
    bad = e >= tol
    if not bad.any():
        return e
    r64 = fp64()
    e_o64, e_g64 = det_errors(r32, r64), det_errors(g, r64)
    print(f"{what} vs FP64: gpu {np.array2string(e_g64[bad], precision=2)}, "
          f"oracle f32 {np.array2string(e_o64[bad], precision=2)}")
    assert np.all(e_o64[bad] > tol / 4), f"{what}: detector error {e.max():.3e} >= {tol} above the float32 noise"
    assert np.all(e_g64[bad] < 4.0 * e_o64[bad]), f"{what}: GPU error vs FP64 above 4x the float32 oracle's"
    return e


def random_setup(nx, ny, nslices=8, nvar=(3, 2), ndetper=0, seed=0, flags=1, ht=200.0,
                 ptype=1, det_ranges=((0, 10), (10, 25), (30, 120))) -> Setup:
    """Generic setup on an nx x ny grid: random weak phase gratings, library
    propagators, ring detectors and probe (used for non-square and
    mixed-radix grids)."""
    from jmultislicelib_amd import JMultiSlice
    rng = np.random.default_rng(seed)
    ax, ay = 0.012 * nx, 0.011 * ny  # ~0.12/0.11 A sampling, in nm
    J = JMultiSlice()
    J.SetHighTension(ht)
    J.SetGridSize(nx, ny)
    J.SetSupercellSize(ax, ay, 0.2)
    wl = J.GetWaveLength()
    gratings = []
    for nv in nvar:
        ph = rng.normal(0, 0.4, (nv, ny, nx))
        gratings.append(np.exp(1j * ph).astype(np.complex64))
    pros = [J.CalculatePropagator(0.19, 0.0, 0.0, ptype), J.CalculatePropagator(0.21, 0.0, 0.0, ptype)]
    dets = [J.CalculateRingDetector(b0, b1) for (b0, b1) in det_ranges]
    probe = J.CalculateProbeWaveFourier(22.0, 0.03)
    objslc = rng.integers(0, len(nvar), nslices).astype(np.int32)
    det_objslc, ndetslc = configs.detection_planes(ndetper, nslices)
    return Setup(nx=nx, ny=ny, a0=(ax, ay, 0.2), ht=ht, wl=wl, nvarslc=list(nvar), gratings=gratings,
                 propagators=pros, proidx=[0, 1][: len(nvar)] + [0] * max(0, len(nvar) - 2),
                 objslc=objslc, det_objslc=det_objslc, ndetslc=ndetslc, detectors=dets, probe=probe,
                 flags=flags, slice_thickness=0.2)
