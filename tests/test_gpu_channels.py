"""GPU parity of the detection channels of the single-wave path
(GPUMultislice, JMultiSlice.cpp:4513-4677) against the CPU oracle:
  * Fourier-space channels (ReadoutDifDet_d :4218-4367): diffraction
    |psi(k)|^2 and the wave psi(k) at every detection plane;
  * real-space channels (ReadoutImgDet_d :4369-4465): probe image |psi(r)|^2
    and psi(r), the unscaled inverse transform of each plane, read inside the
    row pass and, for the exit plane, after an extra inverse transform;
  * averaging channels (GetAvgResult :2489, Reset*Averaging :2592-2664,
    ReadoutDetAvg :2667): accumulated over runs until reset, with the summed
    run weights.
Tolerance: relative L2 <= 1e-4 against the oracle's float32 run; the FP64
twin's error is printed beside it and asserted: GPU and float32 oracle both
within 1e-4 of it, the GPU's error at most 4x the oracle's or 2e-6.
Integrated detectors are checked one by one (and plane by plane)."""
import numpy as np
import pytest

from jmultislicelib_amd import Engine, JMultiSlice, configs
from oracle import oracle as O

from helpers import assert_dets, random_setup, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-4
ALL = 1 | 2 | 4 | 8 | 16
AVG = 32 | 64 | 128 | 256
KEYS = {"dif": 4, "img": 2, "wfr": 8, "wff": 16}


@pytest.mark.parametrize("nx,ny,ndetper", [(64, 64, 0), (48, 32, 3), (128, 96, 2)])
def test_all_channels_single_wave(nx, ny, ndetper):
    sp = random_setup(nx, ny, nslices=9, nvar=(3, 2), ndetper=ndetper, seed=nx + ny, flags=ALL)
    var = configs.variant_sequences(sp, 2, seed=4)
    eng = Engine().setup(sp)
    weights = [0.65, 0.35]
    ref32 = ref64 = None
    for r, w in enumerate(weights):
        eng.wave_offset(0.07, -0.05)
        eng.multislice(var[r], weight=w, accumulate=r > 0)
        win = O.offset_wave(sp.probe, sp.a0[0], sp.a0[1], sp.wl, 0.07, -0.05)
        ref32 = O.multislice_channels(sp, win, var[r], w, prec=32, flags=ALL, chans=ref32)
        ref64 = O.multislice_channels(sp, win, var[r], w, prec=64, flags=ALL, chans=ref64)
    nd = len(sp.detectors)
    g_int = eng.result(1)
    assert_dets(g_int, ref32["int"][:, :nd], TOL, f"{nx}x{ny} ndetper {ndetper} integrated")
    e_g64 = assert_dets(g_int, ref64["int"][:, :nd], TOL, "integrated gpu vs FP64")
    e_o64 = assert_dets(ref32["int"][:, :nd], ref64["int"][:, :nd], TOL, "integrated oracle f32 vs FP64")
    assert np.all(e_g64 <= np.maximum(4.0 * e_o64, 2e-6))
    for k, flag in KEYS.items():
        g = eng.result(flag)
        e = rel_l2(g, ref32[k])
        e64 = rel_l2(g, ref64[k])
        o64 = rel_l2(ref32[k], ref64[k])
        print(f"  {k}: rel L2 vs oracle f32 {e:.2e}, vs FP64 twin: gpu {e64:.2e}, oracle f32 {o64:.2e}")
        assert e < TOL and e64 < TOL and o64 < TOL, k
        assert e64 <= max(4.0 * o64, 2e-6), k
    eng.close()


def test_averaging_channels():
    sp = random_setup(64, 48, nslices=8, nvar=(4, 3), ndetper=4, seed=5, flags=1 | AVG)
    var = configs.variant_sequences(sp, 3, seed=8)
    eng = Engine().setup(sp)
    weights = [0.5, 0.3, 0.2]
    ref = None
    for r, w in enumerate(weights):
        eng.wave_offset(0.02, 0.11)
        eng.multislice(var[r], weight=w, accumulate=False)  # clears the plain channels only
        win = O.offset_wave(sp.probe, sp.a0[0], sp.a0[1], sp.wl, 0.02, 0.11)
        ref = O.multislice_channels(sp, win, var[r], w, prec=32, flags=ALL, chans=ref)
    wi, ww = eng.averaging_weights()
    assert abs(wi - 1.0) < 1e-6 and abs(ww - 1.0) < 1e-6
    for k, flag in (("img", 32), ("dif", 64), ("wfr", 128), ("wff", 256)):
        g, w = eng.avg_result(flag)
        e = rel_l2(g, ref[k])
        print(f"avg {k}: rel L2 {e:.2e}, weight {w}")
        assert e < TOL and abs(w - 1.0) < 1e-6, k
    # intensities of the averaged (elastic) waves
    wff = ref["wff"].astype(np.complex128)
    det = np.stack([d for d, _ in sp.detectors]).astype(np.float64)
    el_int = np.einsum("pyx,dyx->pd", np.abs(wff) ** 2, det)
    got, w = eng.readout_det_avg(1)
    assert_dets(got, el_int, TOL, "elastic (averaged-wave) detector sums")
    assert abs(w - 1.0) < 1e-6
    got, _ = eng.readout_det_avg(4)
    assert rel_l2(got, np.abs(wff) ** 2) < TOL
    got, _ = eng.readout_det_avg(2)
    assert rel_l2(got, np.abs(ref["wfr"].astype(np.complex128)) ** 2) < TOL
    # the elastic part never exceeds the total (|<psi>|^2 <= <|psi|^2>)
    tot, _ = eng.avg_result(64)
    assert np.all(np.abs(wff) ** 2 <= tot + 1e-6 * tot.max())
    # reset clears the channels and the weights of each group
    eng.reset_averaging(0)
    g, w = eng.avg_result(32)
    assert w == 0.0 and not g.any()
    g, w = eng.avg_result(256)
    assert w > 0.99 and g.any()
    eng.reset_averaging(1)
    g, w = eng.avg_result(256)
    assert w == 0.0 and not g.any()
    eng.close()


def test_class_api_real_space_channels():
    """DetectorSetup with image channels through the class API, as a Dr. Probe
    style caller uses it: GetResult(IMAGE / WAVEREAL) after GPUMultislice."""
    sp = configs.perovskite_setup(n=64, ncell=2, nslices=6, nvar=2, ndetper=2, flags=1 | 2 | 8)
    M = JMultiSlice()
    M.SetHighTension(sp.ht)
    M.SetGridSize(sp.nx, sp.ny)
    M.SetSupercellSize(sp.a0[0], sp.a0[1], sp.a0[2])
    assert M.PhaseGratingSetup(2, sp.nx, sp.ny, len(sp.gratings), max(sp.nvarslc), sp.nvarslc) == 0
    assert M.ObjectSliceSetup(sp.objslc) == 0
    assert M.PropagatorSetup(2, len(sp.propagators), sp.proidx) == 0
    assert M.DetectorSetup(2, 2, len(sp.detectors), 1 | 2 | 8) == 0
    for i, g in enumerate(sp.gratings):
        assert M.SetPhaseGratingData(2, i, g) == 0
    for i, p in enumerate(sp.propagators):
        assert M.SetPropagatorData(2, i, p) == 0
    for i, (d, m) in enumerate(sp.detectors):
        assert M.SetDetectorData(2, i, d, m) == 0
    assert M.InitCore(2, 1) == 0
    assert M.SetIncidentWave(2, sp.probe) == 0
    M.SeedRngEx(99)
    var = M.GetRandomVariantSequence(len(sp.objslc))
    M.SeedRngEx(99)
    assert M.OffsetIncomingWave(2, 0.04, 0.09, 0.0) == 0
    assert M.GPUMultislice(0, 0, 1.0) == 0
    img = M.GetResult(2, 2)
    wfr = M.GetResult(2, 8)
    win = O.offset_wave(sp.probe, sp.a0[0], sp.a0[1], sp.wl, 0.04, 0.09)
    ref = O.multislice_channels(sp, win, var, 1.0, prec=32, flags=1 | 2 | 8)
    assert rel_l2(img, ref["img"]) < TOL
    assert rel_l2(wfr, ref["wfr"]) < TOL
    M.Cleanup()
