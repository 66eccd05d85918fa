"""From a structure to a STEM scan on the device: SrTiO3 over 4x4 cells is
sliced into two equidistant slices (CStructure::GetZSliceStructure), its
three atom types get Weickenmeier-Kohl form factors with Debye-Waller
damping and thermal absorptive parts (CAtomType::CalculateRadialFormFactors,
CJEElaSca::PrepareScatteringFunctions), the device generator builds the
phase gratings (CJEElaSca::CalculatePhasegratingGPU), and a scan runs on
them.  The host steps are pinned on the CPU (tests/test_elasca_wk.py); here
the device gratings are checked against the oracle's restatement of the
grating recipe (oracle/elasca_ref.py) on the same 2-D form factors,
relative L2 < 1e-5, and the scan against the oracle's multislice on those
gratings, every detector within 1e-4."""
import numpy as np
import pytest

from jmultislicelib_amd import Engine, JMultiSlice, configs, elasca
from oracle import elasca_ref as E
from oracle import oracle as O

from helpers import assert_dets, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("absorb,bwl_pot", [(3, False), (2, True), (1, False)])
def test_structure_gratings_and_scan(absorb, bwl_pot):
    st = elasca.perovskite_structure(4)
    n, ekv, nslc = 128, 200.0, 2
    a, c = st.cell[0], st.cell[2]
    J = JMultiSlice()
    J.SetHighTension(ekv)
    J.SetGridSize(n, n)
    J.SetSupercellSize(a, a, c)
    wl = J.GetWaveLength()
    cz = c / nslc
    pro = J.CalculatePropagator(cz, 0.0, 0.0, 1)
    dets = [J.CalculateRingDetector(b0, b1) for (b0, b1) in configs.DETECTOR_RANGES]
    probe = J.CalculateProbeWaveFourier(25.0, 0.03)
    nobj = 20
    objslc = np.array([i % nslc for i in range(nobj)], np.int32)
    det_objslc, ndetslc = configs.detection_planes(5, nobj)
    sp = configs.Setup(nx=n, ny=n, a0=(a, a, c), ht=ekv, wl=wl, nvarslc=[1] * nslc, gratings=None,
                       propagators=[pro], proidx=[0] * nslc, objslc=objslc, det_objslc=det_objslc,
                       ndetslc=ndetslc, detectors=dets, probe=probe, flags=1, slice_thickness=cz)
    eng = Engine().setup(sp)
    ff2d, slices, cz2 = elasca.gratings_from_structure(eng, st, nslc, ekv, dwf=True, absorb=absorb, abf=0.1,
                                                       bwl_pot=bwl_pot)
    assert cz2 == pytest.approx(cz, rel=1e-6)
    assert np.all(ff2d.imag[:, 0, 0] > 0)  # absorptive parts present
    dev = [eng.gratings(i, 1)[0] for i in range(nslc)]
    for i, (xy, t, o) in enumerate(slices):
        ref = E.phase_grating(n, n, a, a, cz2, wl, ff2d, xy, t, o, bwl_pot=bwl_pot, bwl_pgr=True)
        e = rel_l2(dev[i], ref)
        print(f"slice {i}: {len(t)} atoms, device vs oracle grating rel L2 {e:.2e}, "
              f"|t| range {np.abs(dev[i]).min():.4f}..{np.abs(dev[i]).max():.4f}")
        assert e < 1e-5
        if absorb != 2:  # absorption damps the transmission on average
            assert np.mean(np.abs(dev[i]) ** 2) < 1.0
    dx = np.array([0.0, 0.0976, 0.19525, 0.3], np.float32)
    dy = np.array([0.0, 0.0976, 0.0, 0.21], np.float32)
    var = np.zeros((len(dx), nobj), np.int32)
    w = np.ones(1, np.float32)
    eng.stage_scan(dx, dy, None, w, var)
    eng.run_staged(0, len(dx))
    eng.sync()
    got = eng.scan_results()
    eng.close()
    sp_h = configs.Setup(**{**sp.__dict__, "gratings": [g[None] for g in dev]})
    ref = O.stem_scan(sp_h, dx, dy, w, var, prec=32, nthreads=4)
    assert_dets(got, ref, 1e-4, f"structure scan absorb={absorb}")
