"""Probe generation on the device (aperture of ApertureFunctionA,
JProbeGen.cpp:471-505, axial aberrations of AberrationFunction :508-571,
normalisation of CalculateProbeWaveFourier :707-777) against the host
generator, which is bit-exact to the oracle.  The device uses its own
tanh/sincos and may fuse multiply-adds, so the bar is float accuracy:
relative L2 <= 1e-5 on the probe, 1e-4 on a scan driven by it."""
import numpy as np
import pytest

from jmultislicelib_amd import Engine, JMultiSlice, configs
from oracle import oracle as O

from helpers import assert_dets, random_setup, rel_l2

pytestmark = pytest.mark.gpu


def _abrr(seed):
    rng = np.random.default_rng(seed)
    ab = np.array([rng.uniform(-1, 1, 2) * 10.0 ** (m - 1) for m, n in O.aberration_table()], np.float32)
    ab[rng.random(24) < 0.5] = 0.0
    return ab


@pytest.mark.parametrize("nx,ny,seed", [(64, 64, 0), (96, 64, 1), (512, 512, 2)])
def test_device_probe_vs_host(nx, ny, seed):
    sp = random_setup(nx, ny, nslices=2, seed=seed)
    eng = Engine().setup(sp)
    J = JMultiSlice()
    J.SetHighTension(sp.ht)
    J.SetGridSize(nx, ny)
    J.SetSupercellSize(sp.a0[0], sp.a0[1], sp.a0[2])
    for ab in (None, _abrr(seed)):
        args = (21.0, 0.03, 0.4, -0.3, 0.04, 20.0, 0.6, -0.2)
        host = J.CalculateProbeWaveFourier(*args, abrr=ab)
        dev = eng.probe(*args, abrr=ab)
        e = rel_l2(dev, host)
        print(f"{nx}x{ny} aberrations={ab is not None}: device vs host probe rel L2 {e:.2e}")
        assert e < 1e-5
        assert abs(np.sum(np.abs(dev.astype(np.complex128)) ** 2) - 1.0) < 1e-5
    eng.close()


def test_scan_with_device_probe():
    """A staged scan whose incoming wave is the device-generated aberrated
    probe matches the oracle run on the host-generated one."""
    sp = configs.perovskite_setup(n=64, ncell=2, nslices=10, nvar=3)
    ab = np.zeros((24, 2), np.float32)
    ab[1] = (8.0, 0.0)      # defocus 8 nm
    ab[5] = (2.0e4, 0.0)    # spherical aberration 20 um
    J = JMultiSlice()
    J.SetHighTension(sp.ht)
    J.SetGridSize(sp.nx, sp.ny)
    J.SetSupercellSize(sp.a0[0], sp.a0[1], sp.a0[2])
    host_probe = J.CalculateProbeWaveFourier(25.0, 0.03, abrr=ab)
    eng = Engine().setup(sp)
    eng.probe(25.0, 0.03, abrr=ab)
    dx, dy = configs.scan_positions(4, 1)
    npass = 2
    w = np.full(npass, 0.5, np.float32)
    var = configs.variant_sequences(sp, len(dx) * npass, seed=3)
    eng.stage_scan(dx, dy, None, w, var)
    eng.run_staged(0, len(dx))
    eng.sync()
    got = eng.scan_results()
    ref = np.zeros_like(got)
    for p in range(len(dx)):
        for r in range(npass):
            win = O.offset_wave(host_probe, sp.a0[0], sp.a0[1], sp.wl, dx[p], dy[p])
            res, _ = O.multislice(sp, win, var[p * npass + r], w[r], prec=32)
            ref[p] += res
    assert_dets(got, ref, 1e-4, "scan with device probe")
    eng.close()
