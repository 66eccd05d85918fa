"""Diffraction de-scan (DescanDifN / DescanDifWavN_d, JMultiSlice.cpp:
2184-2320): every Fourier-space readout (integrated detectors, diffraction
and Fourier-space wave channels) sees the wave shifted cyclically by the
de-scan pixels; real-space channels are unaffected.  GPU against the
oracle's de-scanned readout, relative L2 <= 1e-4, through the engine and the
class API (SetDiffractionDescanN / DiffractionDescan)."""
import numpy as np
import pytest

from jmultislicelib_amd import Engine, JMultiSlice, configs
from oracle import oracle as O

from helpers import assert_dets, random_setup, rel_l2

pytestmark = pytest.mark.gpu

ALL = 1 | 2 | 4 | 8 | 16


@pytest.mark.parametrize("nx,ny,d", [(64, 64, (3, -2)), (96, 64, (-7, 5)), (64, 1024, (2, 11))])
def test_descan_engine(nx, ny, d):
    sp = random_setup(nx, ny, nslices=6, nvar=(2, 3), ndetper=3, seed=nx + 3 * ny, flags=ALL)
    var = configs.variant_sequences(sp, 1, seed=6)[0]
    eng = Engine().setup(sp)
    eng.set_descan(*d)
    eng.wave_offset(0.05, 0.08)
    eng.multislice(var, weight=0.8)
    win = O.offset_wave(sp.probe, sp.a0[0], sp.a0[1], sp.wl, 0.05, 0.08)
    ref = O.multislice_channels(sp, win, var, 0.8, prec=32, flags=ALL, descan=d)
    assert_dets(eng.result(1), ref["int"][:, :len(sp.detectors)], 1e-4, f"{nx}x{ny} descan {d}")
    for k, flag in (("dif", 4), ("wff", 16), ("img", 2), ("wfr", 8)):
        e = rel_l2(eng.result(flag), ref[k])
        print(f"  {k}: {e:.2e}")
        assert e < 1e-4, k
    # off again: the plain readout
    eng.set_descan(0, 0)
    eng.wave_offset(0.05, 0.08)
    eng.multislice(var, weight=0.8)
    plain = O.multislice_channels(sp, win, var, 0.8, prec=32, flags=ALL)
    assert rel_l2(eng.result(4), plain["dif"]) < 1e-4
    eng.close()


def test_descan_class_api():
    sp = configs.perovskite_setup(n=64, ncell=2, nslices=6, nvar=2, flags=1)
    M = JMultiSlice()
    M.SetHighTension(sp.ht)
    M.SetGridSize(sp.nx, sp.ny)
    M.SetSupercellSize(*sp.a0)
    assert M.PhaseGratingSetup(2, sp.nx, sp.ny, 2, 2, sp.nvarslc) == 0
    assert M.ObjectSliceSetup(sp.objslc) == 0
    assert M.PropagatorSetup(2, 1, [0, 0]) == 0
    assert M.DetectorSetup(2, 0, 3, 1) == 0
    for i, g in enumerate(sp.gratings):
        assert M.SetPhaseGratingData(2, i, g) == 0
    assert M.SetPropagatorData(2, 0, sp.propagators[0]) == 0
    for i, (dd, m) in enumerate(sp.detectors):
        assert M.SetDetectorData(2, i, dd, m) == 0
    assert M.InitCore(2, 1) == 0
    assert M.SetIncidentWave(2, sp.probe) == 0
    M.SetDiffractionDescanN(2, 4, -3)
    M.DiffractionDescan(True)
    M.SeedRngEx(5)
    var = M.GetRandomVariantSequence(len(sp.objslc))
    M.SeedRngEx(5)
    assert M.OffsetIncomingWave(2, 0.1, 0.1, 0.0) == 0
    assert M.GPUMultislice(0, 0, 1.0) == 0
    got = M.GetResult(2, 1)
    win = O.offset_wave(sp.probe, sp.a0[0], sp.a0[1], sp.wl, 0.1, 0.1)
    ref = O.multislice_channels(sp, win, var, 1.0, prec=32, flags=1, descan=(4, -3))
    assert_dets(got, ref["int"][:, :3], 1e-4, "class API descan")
    # the flag is a setting of the process-global object (Cleanup keeps it, as
    # in the reference): switch it off for later users
    M.DiffractionDescan(False)
    M.Cleanup()
