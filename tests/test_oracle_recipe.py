"""Pins the oracle's restated setup functions and readout arithmetic to
closed forms, library routines and known-answer tests:
  * wavelength (JMultiSlice.cpp:372-378) vs the relativistic formula with
    scipy.constants;
  * propagators (JMultiSlice.cpp:998-1106): modulus 1/n inside the 2/3 band
    limit, 0 outside, phase -pi lambda dz k^2 (Fresnel), planar = Fresnel to
    second order;
  * ring detectors (:1274-1438): half-open rings, disjoint rings add up to a
    disk, segments partition a ring, mask lists exactly the active pixels;
  * probe (JProbeGen.cpp:707-777): unit norm, aperture radius;
  * detection planes (:579-627): enumerated cases;
  * fdncs2m butterfly sum and masked dot products (:38-242, :3279-3330): exact
    on integer data, float-accurate on random data;
  * RNG (rng.cpp): the MT19937 known answer of the C++ standard
    ([rand.predef]: 10000th output of the default-seeded engine is
    4123659995), unirand's closed form, the polar-method normal deviates
    against numpy's normal_distribution-independent moments;
  * variant sequences (:1571-1676): permutation windows, no repeat across a
    window boundary, single-variant slices give 0.
"""
import numpy as np
import pytest
import scipy.constants as sc

from oracle import oracle as O


def test_wavelength_closed_form():
    for ht in (80.0, 200.0, 300.0):
        E = ht * 1e3 * sc.e
        mc2 = sc.m_e * sc.c ** 2
        lam = sc.h * sc.c / np.sqrt(E * (E + 2 * mc2)) * 1e9  # nm
        assert abs(O.wavelength(ht) - lam) / lam < 1e-6
    assert abs(O.wavelength(200.0) - 0.0025079) < 1e-7


def _k2(nx, ny, a0x, a0y):
    fx = ((np.arange(nx) + nx // 2) % nx - nx // 2) / a0x
    fy = ((np.arange(ny) + ny // 2) % ny - ny // 2) / a0y
    return fx[None, :] ** 2 + fy[:, None] ** 2, fx, fy


@pytest.mark.parametrize("nx,ny", [(64, 64), (96, 64), (65, 33)])
def test_propagator_fresnel_closed_form(nx, ny):
    a0x, a0y, wl, dz = 1.7, 1.3, 0.0025079, 0.19525
    P = O.propagator(nx, ny, a0x, a0y, wl, dz, 0.0, 0.0, 1)
    k2, fx, fy = _k2(nx, ny, a0x, a0y)
    gmax = min((nx - nx % 2) / 2 / a0x, (ny - ny % 2) / 2 / a0y) * 0.666666666666667
    inband = k2 <= gmax ** 2
    n = nx * ny
    assert np.allclose(np.abs(P[inband]) * n, 1.0, atol=1e-6)
    assert np.all(P[~inband] == 0)
    ref = np.exp(-1j * np.pi * wl * dz * k2) / n
    assert np.abs(P[inband] - ref[inband]).max() * n < 2e-6


def test_propagator_planar_vs_fresnel_small_angle():
    nx = ny = 64
    a0 = 20.0  # large cell -> small angles
    wl, dz = 0.0025079, 0.2
    Pp = O.propagator(nx, ny, a0, a0, wl, dz, 0.0, 0.0, 0)
    Pf = O.propagator(nx, ny, a0, a0, wl, dz, 0.0, 0.0, 1)
    assert np.abs(Pp - Pf).max() * nx * ny < 1e-4


def test_propagator_tilt_is_shift():
    """A tilted Fresnel propagator = untilted one times a linear phase (slices shift)."""
    nx = ny = 48
    a0, wl, dz = 1.5, 0.0025079, 0.2
    tx = 0.3  # deg
    Pt = O.propagator(nx, ny, a0, a0, wl, dz, tx, 0.0, 1)
    P0 = O.propagator(nx, ny, a0, a0, wl, dz, 0.0, 0.0, 1)
    k2, fx, fy = _k2(nx, ny, a0, a0)
    t = np.deg2rad(tx)
    ph = np.exp(-1j * (dz * np.pi / wl * np.cos(t) * (-2 * wl * fx[None, :] * t)))
    m = P0 != 0
    assert np.abs(Pt[m] - (P0 * ph)[m]).max() * nx * ny < 5e-5


def test_ring_detectors():
    nx = ny = 128
    a0, wl = 3.124, 0.0025079
    k2, fx, fy = _k2(nx, ny, a0, a0)
    g = np.sqrt(k2)
    k0 = 1 / wl
    rings = [(0, 11), (11, 22), (22, 40)]
    tot = np.zeros((ny, nx))
    for b0, b1 in rings:
        det, msk = O.ring_detector(nx, ny, a0, a0, wl, 1, b0, b1)
        ref = ((g >= b0 * 1e-3 * k0) & (g < b1 * 1e-3 * k0)).astype(np.float32)
        assert np.array_equal(det, ref)
        assert np.array_equal(np.sort(msk), np.flatnonzero(ref.ravel()))
        tot += det
    disk, _ = O.ring_detector(nx, ny, a0, a0, wl, 1, 0, 40)
    assert np.array_equal(tot, disk)
    # four 90-degree segments partition the ring
    seg = np.zeros((ny, nx))
    for p0 in (0, 90, 180, 270):
        d, _ = O.ring_detector(nx, ny, a0, a0, wl, 1, 11, 22, p0, p0 + 90)
        seg += d
    full, _ = O.ring_detector(nx, ny, a0, a0, wl, 1, 11, 22)
    assert np.array_equal(seg, full)
    # first moment pattern: sensitivity kx on the ring
    dk, _ = O.ring_detector(nx, ny, a0, a0, wl, 1, 11, 22, kmom1=1)
    assert np.allclose(dk, full * fx[None, :], atol=1e-4)


def test_probe_norm_and_aperture():
    nx = ny = 128
    a0, wl = 3.124, 0.0025079
    for rs in (0.0, 0.03):
        p = O.probe_wave(nx, ny, a0, a0, wl, 25.0, rs)
        assert abs(np.sum(np.abs(p) ** 2) - 1.0) < 1e-5
    p = O.probe_wave(nx, ny, a0, a0, wl, 25.0, 0.0)
    k2, _, _ = _k2(nx, ny, a0, a0)
    qlim = 25e-3 / wl
    inside = np.sqrt(k2) < qlim
    inside[0, 0] = True
    assert np.array_equal(p != 0, inside)
    assert np.all(np.abs(p.imag) == 0)


def _aperture_closed_form(nx, ny, ax, ay, wl, alpha, rs, asym, adir_deg, x0, y0, btx, bty):
    """Probe amplitude in float64 from the geometry: an aperture of radius
    qlim = alpha/wl centred at (x0, y0)/wl, stretched by (1 + asym) along the
    direction adir and by (1 - asym) across it, with the edge profile
    (1 - tanh(pi (|u| - qlim) / (rs qlim))) / 2 in the undistorted radius |u|
    (hard edge for rs = 0), tilted by (btx, bty)/wl, amplitudes <= 1e-6 cut,
    normalised to unit intensity (JProbeGen.cpp:446-505, 707-777)."""
    nx2, ny2 = nx // 2, ny // 2
    qx = ((np.arange(nx) + nx2) % nx - nx2) / ax
    qy = ((np.arange(ny) + ny2) % ny - ny2) / ay
    qlim = alpha * 1e-3 / wl
    dqx = qx[None, :] - btx * 1e-3 / wl - x0 * 1e-3 / wl
    dqy = qy[:, None] - bty * 1e-3 / wl - y0 * 1e-3 / wl
    phi = np.deg2rad(adir_deg)
    along = dqx * np.cos(phi) + dqy * np.sin(phi)
    across = -dqx * np.sin(phi) + dqy * np.cos(phi)
    u = np.hypot(along / (1.0 + asym), across / (1.0 - asym))
    if rs > 0:
        a = 0.5 * (1.0 - np.tanh(np.pi * (u - qlim) / (rs * qlim)))
    else:
        a = (u < qlim).astype(np.float64)
    a[u == 0] = 1.0
    a[a <= 1e-6] = 0.0
    return a / np.sqrt(np.sum(a * a))


@pytest.mark.parametrize("alpha,rs,asym,adir,x0,y0,btx,bty", [
    (25.0, 0.03, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
    (22.0, 0.12, 0.0, 0.0, 1.5, -2.0, 0.0, 0.0),
    (20.0, 0.08, 0.15, 30.0, 0.0, 0.0, 0.0, 0.0),
    (24.0, 0.05, 0.25, -65.0, -1.0, 2.5, 3.0, -1.5),
    (18.0, 0.0, 0.2, 110.0, 0.0, 0.0, 0.0, 0.0),
])
def test_probe_aperture_closed_form(alpha, rs, asym, adir, x0, y0, btx, bty):
    """Smooth edge, asymmetry, aperture shift and beam tilt of the probe
    aperture against the geometric closed form: the amplitude ratio of any
    two pixels is the ratio of the edge profile at their undistorted radii.
    A sigmoid without pi, a smoothing width not scaled by qlim, a transposed
    or sign-flipped back-projection, an undoubled angle or a tilt applied
    twice each fail it."""
    nx, ny, ax, ay, wl = 128, 96, 3.124, 2.343, 0.0025079
    p = O.probe_wave(nx, ny, ax, ay, wl, alpha, rs, asym, adir, x0, y0, btx, bty)
    want = _aperture_closed_form(nx, ny, ax, ay, wl, alpha, rs, asym, adir, x0, y0, btx, bty)
    assert np.all(p.imag == 0)
    got = np.abs(p).astype(np.float64)
    if rs > 0:
        assert np.count_nonzero((want > 0.05 * want.max()) & (want < 0.95 * want.max())) > 20  # edge pixels
        assert np.max(np.abs(got - want)) < 2e-5 * want.max()
    else:  # hard edge: pixels exactly on the rim may round either way in float32
        diff = (got > 0) != (want > 0)
        assert diff.sum() <= 2
        assert np.max(np.abs(got - want)[~diff]) < 1e-5 * want.max()


@pytest.mark.parametrize("nprotype", [1, 0])
def test_radial_profile_sensitivity_closed_form(nprotype):
    """A linear sensitivity profile p[i] = c0 + c1 i is reproduced exactly by
    linear interpolation: on the ring the detector reads
    c0 + c1 min(g (refpix - 1) / g1, len - 1), where g is the pixel's |k| and
    g1 the inner edge in the same units (JMultiSlice.cpp:1247-1272,
    1407-1416)."""
    nx, ny, a0 = 128, 128, 3.124
    wl = 0.0025079
    c0, c1, plen, refpix = 0.25, 0.03, 30, 9.0
    prof = (c0 + c1 * np.arange(plen)).astype(np.float32)
    b0, b1 = 20.0, 95.0
    det, msk = O.ring_detector(nx, ny, a0, a0, wl, nprotype, b0, b1, profile=prof, refpix=refpix)
    plain, pm = O.ring_detector(nx, ny, a0, a0, wl, nprotype, b0, b1)
    assert np.array_equal(msk, pm)
    k2, _, _ = _k2(nx, ny, a0, a0)
    g = np.sqrt(k2)
    g1 = b0 * 1e-3 / wl if nprotype == 1 else np.sin(b0 * 1e-3) / wl
    want = c0 + c1 * np.minimum(g * (refpix - 1.0) / g1, plen - 1)
    on = plain > 0
    assert on.sum() > 1000 and np.any(g[on] * (refpix - 1.0) / g1 > plen - 1)  # the clamp is reached
    assert np.allclose(det[on], want[on], rtol=2e-6, atol=0)
    assert np.all(det[~on] == 0)


@pytest.mark.parametrize("ndetper,nobj,expect", [
    (0, 10, {10: 0}),
    (3, 10, {0: 0, 3: 1, 6: 2, 9: 3, 10: 4}),
    (5, 10, {0: 0, 5: 1, 10: 2}),
    (1, 3, {0: 0, 1: 1, 2: 2, 3: 3}),
    (4, 0, {0: 0}),
])
def test_detection_planes(ndetper, nobj, expect):
    h, n = O.detection_planes(ndetper, nobj)
    assert n == len(expect)
    for i in range(nobj + 1):
        assert h[i] == expect.get(i, -1)


def test_fdncs2m_sums():
    for n in (1, 2, 3, 31, 32, 33, 1000, 4096, 4097, 10000, 70000):
        a = np.arange(n, dtype=np.float32) % 7  # exact in float
        assert O.fdncs2m(a) == float(a.astype(np.int64).sum())
    rng = np.random.default_rng(1)
    for n in (100, 5000, 300000):
        a = rng.random(n).astype(np.float32)
        s = O.fdncs2m(a)
        assert abs(s - a.astype(np.float64).sum()) / a.sum() < 1e-6


def test_masked_dot():
    rng = np.random.default_rng(2)
    for n in (1, 2, 7, 1001, 50000):
        m = rng.choice(100000, n, replace=False).astype(np.int32)
        a = rng.random(100000).astype(np.float32)
        b = rng.random(100000).astype(np.float32)
        ref = float(np.sum(a[m].astype(np.float64) * b[m]))
        assert abs(O.masked_dot(m, a, b) - ref) / ref < 2e-6


def test_mt19937_known_answer():
    r = O.Rng(5489)
    v = [r.raw() for _ in range(10000)]
    assert v[-1] == 4123659995
    r2 = O.Rng(5489)
    assert r2.rand() == v[0] >> 1


def test_unirand_closed_form():
    r, q = O.Rng(42), O.Rng(42)
    for _ in range(1000):
        x = q.raw() >> 1
        assert r.unirand(2.0, 10.0) == 2.0 + (x / 2147483647.0) * 8.0


def test_normrand_moments():
    r = O.Rng(1234)
    x = np.array([r.normrand(0.0, 1.0) for _ in range(200000)])
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1.0) < 0.01
    # N(0,1) quantile check
    from scipy.stats import kstest
    assert kstest(x, "norm").pvalue > 1e-3


def test_variant_sequence_properties():
    nv = [4, 1, 6]
    obj = np.array([i % 3 for i in range(300)], np.int32)
    seqs = O.variant_sequences(17, nv, obj, 5)
    for s in seqs:
        for t in range(3):
            v = s[obj == t]
            if nv[t] < 2:
                assert np.all(v == 0)
                continue
            win = v[: (len(v) // nv[t]) * nv[t]].reshape(-1, nv[t])
            for w in win:
                assert sorted(w) == list(range(nv[t]))
            assert np.all(win[1:, 0] != win[:-1, -1])  # no repeat across a window boundary
    # different seeds give different sequences, same seed the same
    assert not np.array_equal(seqs, O.variant_sequences(18, nv, obj, 5))
    assert np.array_equal(seqs, O.variant_sequences(17, nv, obj, 5))


def test_aberration_table_order():
    """Interface list (JProbeGen.cpp:187-216): orders 1..8, symmetries of equal
    parity, in that nesting."""
    t = O.aberration_table()
    assert len(t) == 24
    assert t[:10] == [(1, 1), (2, 0), (2, 2), (3, 1), (3, 3), (4, 0), (4, 2), (4, 4), (5, 1), (5, 3)]
    assert t[-1] == (8, 8) and all((m + n) % 2 == 0 and 0 <= n <= m for m, n in t)


def _chi_ref(qx, qy, wl, abrr):
    """Closed form in float64: chi = 2pi/wl sum Re[(ax - i ay)(wx + i wy)^n] |w|^(m-n) / m."""
    wx, wy = wl * qx, wl * qy
    w = np.hypot(wx, wy)
    chi = 0.0
    for (m, n), (ax, ay) in zip(O.aberration_table(), np.asarray(abrr, float).reshape(-1, 2)):
        chi += ((ax - 1j * ay) * (wx + 1j * wy) ** n).real * w ** (m - n) / m
    return 2 * np.pi / wl * chi


def test_aberration_closed_forms():
    wl = 0.0025079
    rng = np.random.default_rng(9)
    for _ in range(50):
        qx, qy = rng.uniform(-8, 8, 2)  # 1/nm, up to ~20 mrad
        # defocus C1: pi wl C1 q^2
        c1 = rng.uniform(-50, 50)
        ab = np.zeros((24, 2)); ab[1] = (c1, 0)
        ref = np.pi * wl * c1 * (qx * qx + qy * qy)
        assert abs(O.aberration_chi(qx, qy, wl, ab) - ref) <= 1e-5 * max(1.0, abs(ref))
        # image shift (a, b): 2 pi (a qx + b qy)
        a, b = rng.uniform(-0.3, 0.3, 2)
        ab = np.zeros((24, 2)); ab[0] = (a, b)
        ref = 2 * np.pi * (a * qx + b * qy)
        assert abs(O.aberration_chi(qx, qy, wl, ab) - ref) <= 1e-5 * max(1.0, abs(ref))
        # spherical C3: pi/2 C3 wl^3 q^4
        c3 = rng.uniform(-2e6, 2e6)
        ab = np.zeros((24, 2)); ab[5] = (c3, 0)
        ref = np.pi / 2 * c3 * wl ** 3 * (qx * qx + qy * qy) ** 2
        assert abs(O.aberration_chi(qx, qy, wl, ab) - ref) <= 1e-5 * max(1.0, abs(ref))
        # random set of all 24 (sizes scaled by order)
        ab = np.array([rng.uniform(-1, 1, 2) * 10.0 ** (m - 1) for m, n in O.aberration_table()])
        ref = _chi_ref(qx, qy, wl, ab)
        assert abs(O.aberration_chi(qx, qy, wl, ab) - ref) <= 2e-5 * max(1.0, abs(ref))


def test_aberrated_probe_phase_and_shift():
    nx = ny = 64
    a0, wl = 3.124, 0.0025079
    ab = np.zeros((24, 2)); ab[1] = (12.0, 0); ab[2] = (3.0, -2.0); ab[5] = (1.0e5, 0)
    p0 = O.probe_wave(nx, ny, a0, a0, wl, 25.0, 0.03)
    p = O.probe_wave(nx, ny, a0, a0, wl, 25.0, 0.03, abrr=ab)
    assert np.allclose(np.abs(p), np.abs(p0), atol=1e-7)      # aberrations change only the phase
    k2, fx, fy = _k2(nx, ny, a0, a0)
    m = np.abs(p0) > 0
    chi = np.array([[_chi_ref(fx[i], fy[j], wl, ab) for i in range(nx)] for j in range(ny)])
    assert np.abs(np.angle(p[m] * np.exp(1j * chi[m]))).max() < 1e-3  # psi = A exp(-i chi)
    # an image shift by whole pixels translates the real-space probe
    s = np.zeros((24, 2)); s[0] = (3 * a0 / nx, -2 * a0 / ny)
    ps = O.probe_wave(nx, ny, a0, a0, wl, 25.0, 0.03, abrr=s)
    r0, r1 = np.fft.ifft2(p0.astype(complex)), np.fft.ifft2(ps.astype(complex))
    assert np.abs(r1 - np.roll(np.roll(r0, -2, 0), 3, 1)).max() < 1e-5 * np.abs(r0).max()
