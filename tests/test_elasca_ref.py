"""The oracle's phase gratings from projected form factors (oracle/elasca_ref.py,
after CJEElaSca::CalculatePhasegratingGPU, JEElaSca.cpp:978-1159), pinned by
what the mathematics fixes: the shift theorem for a point-like form factor,
the continuous Fourier pair of a Gaussian (scale and axes on a non-square
cell), the product of the two physical constants (1 / (4 pi)), absorption
by an imaginary form factor, linearity over atoms and types, and the band
limit as a projection."""
import numpy as np

from oracle import elasca_ref as E

NX, NY, AX, AY, CZ = 48, 32, 1.9, 1.3, 0.195
WL = 0.00250793  # 200 keV [nm]


def gauss_ff(s, amp=1.0, ntype=1):
    """form factor exp(-2 pi^2 s^2 q^2) (FT of the unit 2-D Gaussian of width s)"""
    qx, qy = E.freqs(NX, AX).astype(np.float64), E.freqs(NY, AY).astype(np.float64)
    q2 = qx[None, :] ** 2 + qy[:, None] ** 2
    return np.stack([amp * np.exp(-2 * np.pi ** 2 * s * s * q2)] * ntype).astype(np.complex128)


def gauss_closed(x0, y0, s):
    """periodic sum of the unit 2-D Gaussian of width s around (x0, y0), sampled
    on the grid (pixel (j, i) at (i AX/NX, j AY/NY))"""
    x = np.arange(NX) * AX / NX
    y = np.arange(NY) * AY / NY
    out = np.zeros((NY, NX))
    for mx in range(-2, 3):
        for my in range(-2, 3):
            dx = x[None, :] - x0 - mx * AX
            dy = y[:, None] - y0 - my * AY
            out += np.exp(-(dx * dx + dy * dy) / (2 * s * s)) / (2 * np.pi * s * s)
    return out


def test_constants_product():
    assert abs(E.POT_SCA * E.PGR_SIG - 1 / (4 * np.pi)) < 1e-7


def test_frequencies_fft_order():
    q = E.freqs(7, 2.0)
    assert np.allclose(q, np.array([0, 1, 2, 3, -3, -2, -1]) / 2.0)
    q = E.freqs(8, 4.0)
    assert np.allclose(q, np.array([0, 1, 2, 3, -4, -3, -2, -1]) / 4.0)


def test_point_atom_shift_theorem():
    """ff = 1: V(r) = n POT_SCA / vol at the atom's pixel, zero elsewhere"""
    i0, j0 = 7, 20
    ff = np.ones((1, NY, NX), np.complex128)
    v = E.real_potential(NX, NY, AX, AY, CZ, ff, [[i0 * AX / NX, j0 * AY / NY]], [0])
    want = np.zeros((NY, NX))
    want[j0, i0] = NX * NY * E.POT_SCA / (AX * AY * CZ)
    assert np.abs(v - want).max() < 1e-6 * want.max()  # q is formed in float32


def test_gaussian_form_factor_closed_form():
    s, x0, y0 = 0.08, 0.53, 0.91
    v = E.real_potential(NX, NY, AX, AY, CZ, gauss_ff(s), [[x0, y0]], [0])
    want = E.POT_SCA / CZ * gauss_closed(x0, y0, s)
    assert np.abs(v.imag).max() < 1e-6 * want.max()
    assert np.abs(v.real - want).max() < 1e-6 * want.max()


def test_grating_phase_and_absorption_closed_form():
    s, x0, y0, amp = 0.08, 1.61, 0.12, 200.0
    f = gauss_closed(x0, y0, s)
    t = E.phase_grating(NX, NY, AX, AY, CZ, WL, gauss_ff(s, amp), [[x0, y0]], [0], bwl_pgr=False)
    ph = WL / (4 * np.pi) * amp * f
    assert 0.5 < ph.max() < 3.0  # a phase shift of order one radian
    assert np.abs(np.abs(t) - 1).max() < 1e-9
    assert np.abs(np.angle(t) - ph).max() < 1e-6
    # an imaginary form factor absorbs: |t| = exp(-phase)
    t = E.phase_grating(NX, NY, AX, AY, CZ, WL, 1j * gauss_ff(s, amp), [[x0, y0]], [0], bwl_pgr=False)
    assert np.abs(np.abs(t) - np.exp(-ph)).max() < 1e-6
    assert np.abs(np.angle(t)).max() < 1e-6


def test_linearity_over_atoms_and_types():
    rng = np.random.default_rng(4)
    ff = gauss_ff(0.07, 1.0, 2) * np.array([1.0, 2.5 + 0.3j])[:, None, None]
    xy = rng.uniform(0, 1.2, (9, 2))
    typ = rng.integers(0, 2, 9)
    occ = rng.uniform(0.3, 1.0, 9)
    v = E.real_potential(NX, NY, AX, AY, CZ, ff, xy, typ, occ)
    parts = sum(E.real_potential(NX, NY, AX, AY, CZ, ff, xy[k:k + 1], typ[k:k + 1], occ[k:k + 1]) for k in range(9))
    assert np.abs(v - parts).max() < 1e-9 * np.abs(v).max()
    # each atom is its type's Gaussian, scaled by the type factor and occupancy
    want = sum(occ[k] * [1.0, 2.5 + 0.3j][typ[k]] * gauss_closed(xy[k, 0], xy[k, 1], 0.07) for k in range(9))
    assert np.abs(v - E.POT_SCA / CZ * want).max() < 1e-6 * np.abs(v).max()


def test_band_limit_is_a_projection():
    rng = np.random.default_rng(5)
    ff = gauss_ff(0.03, 30.0)
    xy = rng.uniform(0, 1.2, (6, 2))
    t0 = E.phase_grating(NX, NY, AX, AY, CZ, WL, ff, xy, np.zeros(6, int), bwl_pgr=False)
    t1 = E.phase_grating(NX, NY, AX, AY, CZ, WL, ff, xy, np.zeros(6, int), bwl_pgr=True)
    gm = E.helpers(NX, NY, AX, AY)[2]
    kmax = min(NX // 2 / AX, NY // 2 / AY)
    inside = gm <= 2.0 / 3.0 * kmax * (1 + 1e-6)
    f0, f1 = np.fft.fft2(t0), np.fft.fft2(t1)
    assert np.abs(f1[~inside]).max() < 1e-9 * np.abs(f0).max()
    assert np.abs(f1[inside] - f0[inside]).max() < 1e-9 * np.abs(f0).max()
    assert np.abs(f0[~inside]).max() > 1e-3 * np.abs(f0).max()  # the limit removes something


def test_potential_band_limit_edge():
    """the potential aperture halves the form factor at |q| = 0.9 kmax and
    passes q = 0 unchanged"""
    _, _, gm, kmax, ap, _ = E.helpers(NX, NY, AX, AY)
    assert abs(ap[0, 0] - 1.0) < 1e-12
    assert abs(0.5 - 0.5 * np.tanh((0.9 - 0.9) * 30.0) - 0.5) < 1e-15
    k = np.argmin(np.abs(gm - 0.9 * kmax))
    assert abs(ap.ravel()[k] - 0.5) < 0.5 * 30 * abs(gm.ravel()[k] / kmax - 0.9) + 1e-6
    ff = np.ones((1, NY, NX), np.complex128)
    vq = E.fourier_potential(NX, NY, AX, AY, ff, [[0.0, 0.0]], [0], bwl_pot=True)
    assert np.allclose(vq, ap)
