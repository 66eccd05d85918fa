"""GPU parity of the HIP slice loop against the CPU oracle.

Tolerances (float32 arithmetic on both sides):
  * 2-D FFT: relative L2 error <= 2e-6 against the oracle's FP64 transform;
  * multislice: every detector at every detection plane within 1e-4
    relative L2 of the oracle's float32 run of the same recipe
    (CPUMultislice restated; a detector whose oracle values are all zero
    must be zero), exit waves within 1e-4 relative L2, variant sequences
    identical (drawn on the host from the same seed);
  * against the oracle's FP64 twin: GPU and float32 oracle both within 1e-4,
    the GPU's error at most 4x the float32 oracle's or 2e-6;
  * a detector whose reading sits at float32's rounding-noise floor (where
    the float32 oracle itself is more than 2.5e-5 from its FP64 twin) and
    misses 1e-4 against the float32 oracle is judged against the FP64 twin:
    the GPU's error there must be below 4x the float32 oracle's.
"""
import numpy as np
import pytest

from jmultislicelib_amd import Engine, JMultiSlice, configs, fft2d
from oracle import oracle as O

from helpers import assert_dets, assert_dets_noise_aware, random_setup, rel_l2

pytestmark = pytest.mark.gpu

GPU = 2
TOL = 1e-4


@pytest.mark.parametrize("ny,nx", [(64, 64), (256, 256), (512, 512), (360, 384), (128, 256),
                                   (720, 640), (96, 480), (2048, 64), (16, 768), (320, 32), (48, 128),
                                   (625, 625), (48, 625)])
def test_fft2d_vs_oracle(ny, nx):
    rng = np.random.default_rng(nx * 7 + ny)
    a = (rng.standard_normal((2, ny, nx)) + 1j * rng.standard_normal((2, ny, nx))).astype(np.complex64)
    for d in (-1, +1):
        g = fft2d(a, d)
        r = O.fft2(a, d, prec=64)
        e = rel_l2(g, r)
        print(f"fft {ny}x{nx} dir {d}: rel L2 {e:.2e}")
        assert e < 2e-6


def _oracle_passes(sp, positions, var, weights):
    """oracle readouts and exit-wave channel for passes over positions"""
    out = np.zeros((len(positions), sp.ndetslc, len(sp.detectors)), np.float32)
    waves = []
    k = 0
    for p, (dx, dy) in enumerate(positions):
        acc = np.zeros((sp.ndetslc, len(sp.detectors)), np.float32)
        wacc = None
        for r, w in enumerate(weights):
            win = O.offset_wave(sp.probe, sp.a0[0], sp.a0[1], sp.wl, dx, dy)
            res, wv = O.multislice(sp, win, var[k], w, prec=32, want_wave=True)
            k += 1
            acc += res
            wacc = wv if wacc is None else wacc + wv
        out[p] = acc
        waves.append(wacc)
    return out, waves


@pytest.mark.parametrize("ndetper", [0, 3])
def test_engine_single_wave_vs_oracle(ndetper):
    sp = configs.perovskite_setup(n=64, ncell=2, nslices=12, nvar=4, ndetper=ndetper, flags=1 | 16)
    var = configs.variant_sequences(sp, 3, seed=11)
    eng = Engine().setup(sp)
    weights = [0.5, 0.3, 0.2]
    for r, w in enumerate(weights):
        eng.wave_offset(0.21, 0.37, 0.0)
        eng.multislice(var[r], weight=w, accumulate=r > 0)
    g_int = eng.result(1)
    g_wff = eng.result(16)
    ref, waves = _oracle_passes(sp, [(0.21, 0.37)], var, weights)
    e_wav = rel_l2(g_wff, waves[0])
    print(f"ndetper {ndetper}: wave channel {e_wav:.2e}")
    assert g_int.shape == ref[0].shape
    assert_dets(g_int, ref[0], TOL, f"ndetper {ndetper}")
    assert e_wav < TOL


def test_class_api_dropin_vs_oracle():
    """Reference setup chain through the flat C-ABI (JMultiSliceLib.h)."""
    sp = configs.perovskite_setup(n=64, ncell=2, nslices=10, nvar=3)
    J = JMultiSlice()
    J.Cleanup()
    J.SetHighTension(sp.ht)
    J.SetGridSize(sp.nx, sp.ny)
    J.SetSupercellSize(*sp.a0)
    assert J.PhaseGratingSetup(GPU, sp.nx, sp.ny, 2, 3, [3, 3]) == 0
    assert J.ObjectSliceSetup(sp.objslc) == 0
    assert J.PropagatorSetup(GPU, 1, [0, 0]) == 0
    assert J.DetectorSetup(GPU, 0, 3, 1, 1) == 0
    for i, g in enumerate(sp.gratings):
        assert J.SetPhaseGratingData(GPU, i, g) == 0
    assert J.SetPropagatorData(GPU, 0, sp.propagators[0]) == 0
    for i, (d, m) in enumerate(sp.detectors):
        assert J.SetDetectorData(GPU, i, d, m) == 0
    assert J.InitCore(GPU, 1) == 0
    assert J.SetIncidentWave(GPU, sp.probe) == 0
    J.SeedRngEx(99)
    pos = [(0.0, 0.0), (0.33, 0.12)]
    npass = 2
    res = []
    for (dx, dy) in pos:
        for r in range(npass):
            assert J.OffsetIncomingWave(GPU, dx, dy, 0.0) == 0
            assert J.GPUMultislice(0, 1 if r else 0, 1.0 / npass) == 0
        res.append(J.GetResult(GPU, 1)[0])
    var = O.variant_sequences(99, sp.nvarslc, sp.objslc, len(pos) * npass)
    ref, _ = _oracle_passes(sp, pos, var, [1.0 / npass] * npass)
    assert_dets(np.array(res), ref[:, 0], TOL, "class API")
    # the batched entry gives the same numbers from the same RNG stream
    J.SeedRngEx(99)
    dx = np.array([p[0] for p in pos], np.float32)
    dy = np.array([p[1] for p in pos], np.float32)
    scan = J.GPUMultisliceScan(dx, dy, None, [1.0 / npass] * npass)
    assert_dets(scan[:, 0], ref[:, 0], TOL, "class API batched scan")
    J.Cleanup()


@pytest.mark.parametrize("islc0", [3, 6])
def test_start_inside_the_stack_vs_oracle(islc0):
    """GPUMultislice(islc0 > 0) (JMultiSlice.cpp:4568-4590): the incident wave
    enters at plane islc0 and only the planes from there on are read out;
    single-wave path and a staged batch against the oracle started at the
    same plane (islc0 = 6 is the exit plane: a readout only)."""
    sp = random_setup(96, 64, nslices=6, nvar=(3, 2), ndetper=2, seed=31)
    var = configs.variant_sequences(sp, 2, seed=6)
    eng = Engine().setup(sp)
    win = O.offset_wave(sp.probe, sp.a0[0], sp.a0[1], sp.wl, 0.21, 0.13)
    eng.wave_offset(0.21, 0.13, 0.0)
    eng.multislice(var[0], 1.0, False, islc0=islc0)
    g = eng.result(1)
    ref, _ = O.multislice(sp, win, var[0], 1.0, islc0=islc0)
    assert_dets(g, ref, TOL, f"single wave from plane {islc0}")
    assert np.count_nonzero(ref) > 0
    dx = np.array([0.21, 0.5], np.float32)
    dy = np.array([0.13, 0.02], np.float32)
    eng.set_batch(2)
    eng.stage_scan(dx, dy, None, np.ones(1, np.float32), var)
    eng.run_staged(0, 2)  # a full run first: its readouts of the early planes must not leak
    eng.sync()
    assert np.count_nonzero(eng.scan_results()[:, 0]) > 0
    eng.run_staged(0, 2, islc0)
    eng.sync()
    gs = eng.scan_results()
    for p in range(2):
        w_p = O.offset_wave(sp.probe, sp.a0[0], sp.a0[1], sp.wl, float(dx[p]), float(dy[p]))
        r_p, _ = O.multislice(sp, w_p, var[p], 1.0, islc0=islc0)
        assert_dets(gs[p], r_p, TOL, f"staged position {p} from plane {islc0}")
    eng.close()


def test_detectors_without_masks_vs_oracle():
    """Detectors uploaded without a pixel mask (msklen 0) weight every pixel
    (the reference's CPU readout, DotProduct_h): a smooth random weight map
    and a ring detector with its mask dropped, beside a masked ring; single
    wave and staged batch against the oracle."""
    sp = random_setup(64, 96, nslices=5, nvar=(2, 3), ndetper=2, seed=41)
    rng = np.random.default_rng(5)
    ky, kx = np.meshgrid(np.fft.fftfreq(sp.ny), np.fft.fftfreq(sp.nx), indexing="ij")
    smooth = (0.5 + 0.5 * np.cos(6.0 * kx) * np.cos(4.0 * ky) + 0.1 * rng.random((sp.ny, sp.nx))).astype(np.float32)
    d0, m0 = sp.detectors[0]
    d2, _ = sp.detectors[2]
    sp.detectors = [(d0, m0), (smooth, None), (d2, None)]
    var = configs.variant_sequences(sp, 2, seed=8)
    eng = Engine().setup(sp)
    win = O.offset_wave(sp.probe, sp.a0[0], sp.a0[1], sp.wl, 0.17, 0.29)
    eng.wave_offset(0.17, 0.29, 0.0)
    eng.multislice(var[0], 1.0, False)
    ref, _ = O.multislice(sp, win, var[0], 1.0)
    assert_dets(eng.result(1), ref, TOL, "single wave, mask-less detectors")
    assert ref[-1, 1] > 0.5  # the weight map sees most of the intensity
    dx = np.array([0.17, 0.6], np.float32)
    dy = np.array([0.29, 0.05], np.float32)
    eng.set_batch(2)
    eng.stage_scan(dx, dy, None, np.ones(1, np.float32), var)
    eng.run_staged(0, 2)
    eng.sync()
    ref_s = O.stem_scan(sp, dx, dy, np.ones(1, np.float32), var, prec=32, nthreads=2)
    assert_dets(eng.scan_results(), ref_s, TOL, "staged, mask-less detectors")
    eng.close()


def test_offset_defocus_tilt_phase_plate_vs_oracle():
    """OffsetIncomingWave with a defocus, a beam-tilt pixel shift and a
    constant phase plate (bcppl on and off) on the single-wave path, and a
    staged batch with a defocus per position and the phase plate (fused into
    the first column pass); every case against the oracle's offset wave run
    through its slice loop."""
    sp = random_setup(96, 64, nslices=5, nvar=(2, 2), ndetper=2, seed=61)
    ax, ay, wl = sp.a0[0], sp.a0[1], sp.wl
    cppl = np.random.default_rng(9).uniform(-0.6, 0.6, (sp.ny, sp.nx)).astype(np.float32)
    var = configs.variant_sequences(sp, 3, seed=2)
    eng = Engine().setup(sp)
    eng._c("jms_hip_set_phaseplate", cppl.ctypes.data)
    for k, (dx, dy, dz, bt, use) in enumerate([(0.21, 0.13, 4.0, (2, -3), True),
                                               (0.05, 0.4, -6.0, (0, 0), False),
                                               (0.33, 0.02, 0.0, (-1, 1), True)]):
        eng.wave_offset(dx, dy, dz, bt[0], bt[1], use)
        eng.multislice(var[k], 1.0, False)
        win = O.offset_wave(sp.probe, ax, ay, wl, dx, dy, dz, cppl=cppl if use else None, ibtx=bt[0], ibty=bt[1])
        ref, _ = O.multislice(sp, win, var[k], 1.0)
        assert_dets(eng.result(1), ref, TOL, f"single wave, case {k}")
    dx = np.array([0.21, 0.05, 0.33], np.float32)
    dy = np.array([0.13, 0.4, 0.02], np.float32)
    dz = np.array([4.0, -6.0, 0.0], np.float32)
    eng.set_batch(3)
    eng.stage_scan(dx, dy, dz, np.ones(1, np.float32), var)
    eng.run_staged(0, 3)
    eng.sync()
    gs = eng.scan_results()
    for p in range(3):
        win = O.offset_wave(sp.probe, ax, ay, wl, float(dx[p]), float(dy[p]), float(dz[p]), cppl=cppl)
        ref, _ = O.multislice(sp, win, var[p], 1.0)
        assert_dets(gs[p], ref, TOL, f"staged position {p}")
    eng.close()


def test_detection_setup_refuses_shared_or_out_of_range_indices():
    sp = random_setup(64, 64, nslices=4, nvar=(2, 1), ndetper=2, seed=3)
    eng = Engine().setup(sp)
    for plane_ids, ndetslc, code in (([0, -1, 0, -1, 1], 3, 6), ([0, -1, 1, -1, 3], 3, 5)):
        ids = np.array(plane_ids, np.int32)
        rc = eng.L.jms_hip_setup_detection(eng.h, ndetslc, ids.ctypes.data, 3, 1)
        assert rc == code, (plane_ids, rc)
    ids = np.array([0, -1, 1, -1, 2], np.int32)
    assert eng.L.jms_hip_setup_detection(eng.h, 3, ids.ctypes.data, 3, 1) == 0
    eng.close()


def test_zero_object_slices_read_the_incident_wave():
    """No object slices (ObjectSliceSetup with an empty list, a valid set-up in
    the reference): the exit plane is the entrance plane, so each detector
    reads sum_mask |probe|^2 * det of the incident probe (a lateral offset is
    a phase ramp and leaves |psi(k)|^2 unchanged); checked per detector
    against that sum in float64."""
    sp = configs.perovskite_setup(n=64, ncell=2, nslices=10, nvar=3)
    J = JMultiSlice()
    J.Cleanup()
    J.SetHighTension(sp.ht)
    J.SetGridSize(sp.nx, sp.ny)
    J.SetSupercellSize(*sp.a0)
    assert J.PhaseGratingSetup(GPU, sp.nx, sp.ny, 2, 3, [3, 3]) == 0
    assert J.ObjectSliceSetup([]) == 0
    assert J.PropagatorSetup(GPU, 1, [0, 0]) == 0
    assert J.DetectorSetup(GPU, 0, 3, 1, 1) == 0
    for i, g in enumerate(sp.gratings):
        assert J.SetPhaseGratingData(GPU, i, g) == 0
    assert J.SetPropagatorData(GPU, 0, sp.propagators[0]) == 0
    for i, (d, m) in enumerate(sp.detectors):
        assert J.SetDetectorData(GPU, i, d, m) == 0
    assert J.InitCore(GPU, 1) == 0
    assert J.SetIncidentWave(GPU, sp.probe) == 0
    p2 = np.abs(sp.probe.astype(np.complex128).ravel()) ** 2
    want = np.array([np.sum(p2[np.asarray(m)] * np.asarray(d, np.float64).ravel()[np.asarray(m)])
                     for d, m in sp.detectors])
    assert want[0] > 0.1
    for dx, dy in ((0.0, 0.0), (0.33, 0.12)):
        assert J.OffsetIncomingWave(GPU, dx, dy, 0.0) == 0
        assert J.GPUMultislice(0, 0, 1.0) == 0
        got = J.GetResult(GPU, 1)[0]
        err = np.abs(got - want) / np.maximum(np.abs(want), 1e-12)
        assert np.all(err < 1e-5), (got, want)
    J.Cleanup()


@pytest.mark.parametrize("graph", [False, True])
def test_staged_scan_vs_oracle(graph):
    sp = configs.perovskite_setup(n=64, ncell=2, nslices=16, nvar=4)
    npos, npass = 7, 3
    rng = np.random.default_rng(5)
    dx = rng.uniform(0, 0.78, npos).astype(np.float32)
    dy = rng.uniform(0, 0.78, npos).astype(np.float32)
    weights = np.full(npass, 1.0 / npass, np.float32)
    var = configs.variant_sequences(sp, npos * npass, seed=42)
    eng = Engine().setup(sp)
    eng.set_batch(6)  # several batches, last one partial
    eng.set_graph(graph)
    eng.stage_scan(dx, dy, None, weights, var)
    eng.run_staged(0, npos)
    eng.sync()
    g = eng.scan_results()
    ref = O.stem_scan(sp, dx, dy, weights, var, prec=32, nthreads=4)
    assert_dets(g, ref, TOL, f"staged scan graph={graph}")


def test_graph_cache_follows_batch_shape_and_buffers():
    """Graph replay stays correct when the batch buffers grow between calls
    and when a scan is staged again with another pass grouping of the same
    wave count: a graph-on engine gives exactly the graph-off engine's
    numbers after every call of the sequence (and the oracle's)."""
    sp = configs.perovskite_setup(n=64, ncell=2, nslices=10, nvar=3)
    rng = np.random.default_rng(17)
    engines = [Engine().setup(sp), Engine().setup(sp)]
    engines[1].set_graph(True)

    def stage(npos, npass, seed):
        dx = rng.uniform(0, 0.78, npos).astype(np.float32)
        dy = rng.uniform(0, 0.78, npos).astype(np.float32)
        w = np.linspace(1.0, 2.0, npass).astype(np.float32)
        w /= w.sum()
        var = configs.variant_sequences(sp, npos * npass, seed=seed)
        for eng in engines:
            eng.stage_scan(dx, dy, None, w, var)
        return dx, dy, w, var

    def run(pos0, npos):
        out = []
        for eng in engines:
            eng.run_staged(pos0, npos)
            eng.sync()
            out.append(eng.scan_results())
        assert np.array_equal(out[0], out[1]), f"graph replay differs after run_staged({pos0}, {npos})"
        return out[0]

    dx, dy, w, var = stage(8, 2, 1)
    run(0, 2)   # batch of 4 waves: captured
    run(0, 8)   # 16 waves: the batch buffers grow
    g = run(0, 2)  # 4 waves again, after the buffers moved
    ref = O.stem_scan(sp, dx, dy, w, var, prec=32, nthreads=4)
    assert_dets(g, ref, TOL, "after buffer growth")
    dx, dy, w, var = stage(4, 4, 2)  # same 16 waves, other grouping and weights
    g = run(0, 4)
    ref = O.stem_scan(sp, dx, dy, w, var, prec=32, nthreads=4)
    assert_dets(g, ref, TOL, "restaged 4 x 4")
    dx, dy, w, var = stage(2, 8, 3)  # more passes: the weight buffer grows
    g = run(0, 2)
    ref = O.stem_scan(sp, dx, dy, w, var, prec=32, nthreads=4)
    assert_dets(g, ref, TOL, "restaged 2 x 8")
    for eng in engines:
        eng.close()


@pytest.mark.parametrize("nx,ny", [(128, 64), (64, 128), (360, 384), (96, 160), (2048, 64), (64, 2048),
                                   (1024, 96), (96, 1024), (768, 640), (360, 1024), (48, 2048),
                                   (384, 360), (96, 480), (480, 720), (720, 64), (320, 48), (48, 320),
                                   (32, 128), (625, 625), (625, 96), (64, 625),
                                   # long columns with a partial last tile (1 or 8 columns)
                                   (625, 1024), (625, 2048), (360, 2048)])
def test_nonsquare_mixed_radix(nx, ny):
    sp = random_setup(nx, ny, nslices=6, nvar=(3, 2), ndetper=2, seed=nx + ny)
    var = configs.variant_sequences(sp, 2, seed=3)
    eng = Engine().setup(sp)
    eng.wave_offset(0.1, 0.05, 0.0)
    eng.multislice(var[0], 0.6, False)
    eng.wave_offset(0.1, 0.05, 0.0)
    eng.multislice(var[1], 0.4, True)
    g = eng.result(1)
    ref, _ = _oracle_passes(sp, [(0.1, 0.05)], var, [0.6, 0.4])
    assert_dets(g, ref[0], TOL, f"{nx}x{ny}")


def test_single_probe_256_vs_oracle_fp32_fp64():
    """256^2, 100 slices, one probe on an A column: detectors and exit wave
    against the oracle's float32 run; both against the FP64 re-run."""
    sp = configs.single_probe_256(flags=1 | 16)
    var = np.zeros((1, len(sp.objslc)), np.int32)
    eng = Engine().setup(sp)
    eng.wave_offset(0.0, 0.0, 0.0)
    eng.multislice(var[0], 1.0, False)
    g_int, g_wav = eng.result(1), eng.result(16)
    win = O.offset_wave(sp.probe, sp.a0[0], sp.a0[1], sp.wl, 0.0, 0.0)
    r32, w32 = O.multislice(sp, win, var[0], 1.0, prec=32, want_wave=True)
    r64, w64 = O.multislice(sp, win, var[0], 1.0, prec=64, want_wave=True)
    print("detectors gpu", g_int.ravel(), "f64", r64.ravel())
    assert_dets(g_int, r32, TOL, "256^2x100 gpu vs oracle f32")
    e_g64 = assert_dets(g_int, r64, TOL, "gpu vs FP64")
    e_o64 = assert_dets(r32, r64, TOL, "oracle f32 vs FP64")
    assert np.all(e_g64 <= np.maximum(4.0 * e_o64, 2e-6))
    e_wav, ew_g64, ew_o64 = rel_l2(g_wav, w32), rel_l2(g_wav, w64), rel_l2(w32, w64)
    print(f"exit wave: gpu vs oracle f32 {e_wav:.2e}, vs FP64: gpu {ew_g64:.2e}, oracle f32 {ew_o64:.2e}")
    assert e_wav < TOL and ew_g64 < TOL and ew_o64 < TOL
    assert ew_g64 <= max(4.0 * ew_o64, 2e-6)


def test_full_2048_grid_batch_vs_oracle():
    """2048^2 (the long-line kernels on a full grid): a staged batch of two
    probe passes with detection planes every 2 slices, against the oracle."""
    sp = random_setup(2048, 2048, nslices=5, nvar=(2, 2), ndetper=2, seed=21)
    npos = 2
    dx = np.array([0.3, 1.1], np.float32)
    dy = np.array([0.7, 0.2], np.float32)
    var = configs.variant_sequences(sp, npos, seed=4)
    eng = Engine().setup(sp)
    eng.set_batch(2)
    eng.stage_scan(dx, dy, None, np.ones(1, np.float32), var)
    eng.run_staged(0, npos)
    eng.sync()
    g = eng.scan_results()
    ref = O.stem_scan(sp, dx, dy, np.ones(1, np.float32), var, prec=32, nthreads=2)
    assert_dets(g, ref, TOL, "2048^2 batch")
    eng.close()


def test_full_1024_grid_batch_vs_oracle():
    """1024^2 (the long-column kernel with 8-column tiles in pairs of waves):
    a staged batch of three probe passes, the last pair of waves short, with
    detection planes every 2 slices, against the oracle."""
    sp = random_setup(1024, 1024, nslices=5, nvar=(2, 2), ndetper=2, seed=23)
    npos = 3
    dx = np.array([0.3, 1.1, 0.55], np.float32)
    dy = np.array([0.7, 0.2, 0.95], np.float32)
    var = configs.variant_sequences(sp, npos, seed=5)
    eng = Engine().setup(sp)
    eng.set_batch(npos)
    eng.stage_scan(dx, dy, None, np.ones(1, np.float32), var)
    eng.run_staged(0, npos)
    eng.sync()
    g = eng.scan_results()
    ref = O.stem_scan(sp, dx, dy, np.ones(1, np.float32), var, prec=32, nthreads=3)
    assert_dets(g, ref, TOL, "1024^2 batch")
    eng.close()


def test_reference_sample_job_shape_625():
    """The shape of the reference's own sample job (test/JMSBench1/test with
    STO_test.slp): 625^2 pixels over 4x4 SrTiO3 cells at 300 kV, 64 slices
    with a detection plane every 4, 25 mrad, BF 0-10 / ABF 12-24 / HAADF
    80-200 mrad, positions on the 40x40 raster over 0.782 nm.  The slice
    files themselves are not part of the reference, so the gratings come
    from the Gaussian potential model, with 3 frozen-phonon variants per
    slice type (the listing has 150); a staged scan of 5 positions x 2
    passes against the oracle."""
    sp = configs.perovskite_setup(n=625, ncell=4, nslices=64, nvar=3, ht=300.0, alpha=25.0, ndetper=4,
                                  det_ranges=((0.0, 10.0), (12.0, 24.0), (80.0, 200.0)))
    assert sp.ndetslc == 17
    step = 0.782 / 40
    idx = np.array([0, 7, 19, 26, 39])
    dx = (idx * step).astype(np.float32)
    dy = (idx[::-1] * step).astype(np.float32)
    npass = 2
    weights = np.full(npass, 1.0 / npass, np.float32)
    var = configs.variant_sequences(sp, len(dx) * npass, seed=7)
    eng = Engine().setup(sp)
    eng.set_batch(4)  # three batches, the last one partial
    eng.stage_scan(dx, dy, None, weights, var)
    eng.run_staged(0, len(dx))
    eng.sync()
    g = eng.scan_results()
    ref = O.stem_scan(sp, dx, dy, weights, var, prec=32, nthreads=8)
    print(f"625^2 sample-job shape: {g.shape} results")
    # the 80-200 mrad ring lies inside the band limit (~262 mrad) but reads only
    # ~1e-10 of the incident intensity (Gaussian potentials scatter little to
    # high angles), at float32's noise floor
    assert np.all(ref[:, -1, 2] > 0)
    assert_dets_noise_aware(g, ref, lambda: O.stem_scan(sp, dx, dy, weights, var, prec=64, nthreads=8), TOL,
                            "625^2 sample-job shape")
    eng.close()
