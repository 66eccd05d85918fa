"""Phase gratings generated on the device (jms_hip_gaussian_gratings) against
the host generator of the same Gaussian projected-potential model
(jms_gaussian_phase_grating, bit-exact to the oracle): both sum every atom
image within 8 sigma per pixel in the same order in double, so they differ
only where the device's exp/cos/sin round differently from the host's; the
bar is 1e-6 absolute on the complex64 values.  Frozen-phonon variants of the
perovskite workload and a random amorphous set with images across a small
cell are covered, and a multislice run on device gratings matches the run
on host gratings."""
import numpy as np
import pytest

from jmultislicelib_amd import Engine, configs
from oracle import oracle as O

from helpers import assert_dets, assert_dets_noise_aware, det_errors

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,ncell", [(64, 2), (256, 4), (512, 8)])
def test_perovskite_variants(n, ncell):
    sp = configs.perovskite_setup(n=n, ncell=ncell, nslices=4, nvar=3)
    eng = Engine().setup(sp)
    cell = ncell * configs.A_PEROVSKITE
    for t in range(2):
        sites = [configs.perovskite_sites(ncell, t, 3, v, 0.08, 1000) for v in range(3)]
        eng.gaussian_gratings(t, sites, configs.SIGMA_POT, configs.SIG_INT)
        dev = eng.gratings(t, 3)
        for v, (xy, z) in enumerate(sites):
            host = configs.gaussian_grating(n, cell, xy, z)
            d = np.abs(dev[v] - host).max()
            print(f"{n}^2 type {t} variant {v}: max |device - host| {d:.2e}")
            assert d < 1e-6
    eng.close()


def test_amorphous_images_small_cell():
    """Atoms near every edge of a cell smaller than the 2 x 8 sigma cut-off
    diameter: each pixel collects several images of the same atom."""
    rng = np.random.default_rng(3)
    n = 48
    sp = configs.perovskite_setup(n=n, ncell=1, nslices=2, nvar=1)  # 3.905 A cell < 4.8 A
    eng = Engine().setup(sp)
    xy = rng.uniform(-1.0, 5.0, (9, 2))
    z = rng.integers(6, 80, 9).astype(np.int32)
    eng.gaussian_gratings(0, [(xy, z)], configs.SIGMA_POT, configs.SIG_INT)
    dev = eng.gratings(0, 1)[0]
    host = configs.gaussian_grating(n, configs.A_PEROVSKITE, xy, z)
    assert np.abs(dev - host).max() < 1e-6
    eng.close()


def test_multislice_on_device_gratings():
    sp = configs.perovskite_setup(n=128, ncell=4, nslices=12, nvar=4)
    var = configs.variant_sequences(sp, 1, seed=2)[0]
    eng = Engine().setup(sp)
    eng.wave_offset(0.1, 0.2)
    eng.multislice(var)
    ref = eng.result()
    # other displacements (seed base 2000) first: the generated gratings are used
    for t in range(2):
        sites = [configs.perovskite_sites(4, t, 4, v, 0.08, 2000) for v in range(4)]
        eng.gaussian_gratings(t, sites, configs.SIGMA_POT, configs.SIG_INT)
    eng.wave_offset(0.1, 0.2)
    eng.multislice(var)
    other = eng.result()
    assert det_errors(other, ref).max() > 1e-4
    for t in range(2):
        sites = [configs.perovskite_sites(4, t, 4, v, 0.08, 1000) for v in range(4)]
        eng.gaussian_gratings(t, sites, configs.SIGMA_POT, configs.SIG_INT)
    eng.wave_offset(0.1, 0.2)
    eng.multislice(var)
    got = eng.result()
    print(f"other displacements: {det_errors(other, ref)}")
    assert_dets(got, ref, 1e-5, "multislice on device vs host gratings")
    eng.close()


def test_thick_amorphous_small_vs_oracle():
    """The large-specimen workload's recipe (distinct amorphous slices, device
    gratings, no frozen phonons) at 256^2 x 24 slices: a scan of 5 probes
    against the oracle run on host-generated gratings."""
    sp = configs.thick_amorphous(n=256, cell_ang=15.625, nslices=24)
    eng = Engine().setup(sp)
    # host gratings for the oracle
    host = []
    for sets in sp.grating_sites:
        host.append(np.stack([configs.gaussian_grating(sp.nx, 15.625, xy, z) for xy, z in sets]))
    sp_h = configs.Setup(**{**sp.__dict__, "gratings": host})
    dx = np.array([0.0, 0.3, 0.7, 1.1, 1.5], np.float32)
    dy = np.array([0.0, 0.2, 0.9, 0.4, 1.3], np.float32)
    var = np.zeros((len(dx), len(sp.objslc)), np.int32)
    w = np.ones(1, np.float32)
    eng.stage_scan(dx, dy, None, w, var)
    eng.run_staged(0, len(dx))
    eng.sync()
    got = eng.scan_results()
    ref = O.stem_scan(sp_h, dx, dy, w, var, prec=32, nthreads=4)
    assert_dets_noise_aware(got, ref, lambda: O.stem_scan(sp_h, dx, dy, w, var, prec=64, nthreads=4), 1e-4,
                            "thick amorphous 256^2 x 24")
    eng.close()
