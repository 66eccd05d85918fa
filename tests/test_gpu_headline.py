"""GPU parity of the headline STEM scan configuration, the one the benchmark
times: a 512x512 grid over 8x8 perovskite cells (31.24 A), 200 slices of two
alternating types, 8 frozen-phonon variants per type (rms 0.08 A, seeds
1000 + 8*type + v), the 0-11 / 11-22 / 70-200 mrad ring detectors at the
exit plane, 8 passes per position with weights 1/8 and variant sequences
from one generator seeded 42.  This runs the 512-point slice-loop kernels
(the row pass and the initial, middle and exit column passes with the tuned
launch shapes) through the staged-scan entry, with and without HIP graph
replay, over several batches (the last one partial).

At 0.061 A sampling the 2/3 band limit is about 137 mrad, so the 70-200 mrad
ring (HAADF) carries signal; every detector is checked on its own.

Tolerances:
  * each detector, GPU vs the oracle's float32 run (CPUMultislice restated):
    relative L2 < 1e-4 over the positions;
  * each detector against the oracle's FP64 twin of the same recipe: the
    GPU's and the float32 oracle's error are both < 1e-4, and the GPU's is
    at most 4x the float32 oracle's or 2e-6, whichever is larger (the GPU
    path is no less accurate than the CPU reference path, up to the
    rounding noise of two different float32 FFTs).
"""
import os

import numpy as np
import pytest

from jmultislicelib_amd import Engine, configs
from oracle import oracle as O

from helpers import assert_dets, det_errors

pytestmark = pytest.mark.gpu

TOL = 1e-4
NPASS = 8
# raster indices of the 64x64 scan over 2x2 cells: the A column at the
# origin, points between columns, a B column, the far corner
IDX = np.array([0, 37, 1056, 2080, 3000, 4095])


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def headline():
    sp = configs.stem_scan_512()
    dx_all, dy_all = configs.scan_positions(64, 2)
    dx, dy = dx_all[IDX], dy_all[IDX]
    weights = np.full(NPASS, 1.0 / NPASS, np.float32)
    var = configs.variant_sequences(sp, len(dx) * NPASS, seed=42)
    r32 = O.stem_scan(sp, dx, dy, weights, var, prec=32, nthreads=_threads())
    r64 = O.stem_scan(sp, dx, dy, weights, var, prec=64, nthreads=_threads())
    return sp, dx, dy, weights, var, r32, r64


@pytest.mark.parametrize("graph", [False, True])
def test_headline_scan_512_vs_oracle(headline, graph):
    sp, dx, dy, weights, var, r32, r64 = headline
    assert sp.nx == sp.ny == 512 and len(sp.objslc) == 200 and sp.nvarslc == [8, 8]
    # HAADF carries signal at every position (inside the band limit)
    assert np.all(r32[:, 0, 2] > 1e-7), r32[:, 0, 2]
    eng = Engine().setup(sp)
    eng.set_batch(4 * NPASS)  # 6 positions: a batch of 4, then a partial batch of 2
    eng.set_graph(graph)
    eng.stage_scan(dx, dy, None, weights, var)
    eng.run_staged(0, len(dx))
    eng.sync()
    g = eng.scan_results()
    if graph:  # a second replay of the captured graphs gives the same numbers
        eng.run_staged(0, len(dx))
        eng.sync()
        assert np.array_equal(eng.scan_results(), g)
    eng.close()
    print("positions x [BF, ABF, HAADF]: gpu", g[:, 0], "oracle f32", r32[:, 0])
    assert_dets(g, r32, TOL, f"graph={graph} gpu vs oracle f32")
    e_g64 = assert_dets(g, r64, TOL, "gpu vs FP64")
    e_o64 = assert_dets(r32, r64, TOL, "oracle f32 vs FP64")
    assert np.all(e_g64 <= np.maximum(4.0 * e_o64, 2e-6)), (e_g64, e_o64)


def test_headline_single_wave_matches_batch(headline):
    """GPUMultislice on one wave (OffsetIncomingWave + the single-wave
    path) at the headline shape gives the batch's numbers for one pass."""
    sp, dx, dy, weights, var, r32, r64 = headline
    eng = Engine().setup(sp)
    p, r = 1, 3
    eng.wave_offset(float(dx[p]), float(dy[p]), 0.0)
    eng.multislice(var[p * NPASS + r], weight=1.0, accumulate=False)
    g1 = eng.result(1)
    eng.set_batch(1)
    eng.stage_scan(dx[p:p + 1], dy[p:p + 1], None, np.ones(1, np.float32), var[p * NPASS + r:p * NPASS + r + 1])
    eng.run_staged(0, 1)
    eng.sync()
    gb = eng.scan_results()[0]
    eng.close()
    win = O.offset_wave(sp.probe, sp.a0[0], sp.a0[1], sp.wl, float(dx[p]), float(dy[p]))
    ref, _ = O.multislice(sp, win, var[p * NPASS + r], 1.0, prec=32)
    assert_dets(g1, ref, TOL, "single wave vs oracle")
    e = det_errors(g1, gb)
    print("single wave vs batch of one", e)
    assert np.all(e < 1e-6)
