"""Uneven batch: several random-crystal jobs of different grid edges and
slice counts run at the same time, one engine (and HIP stream) each; every
job's detector sums match the oracle within 1e-4 relative L2, detector by
detector (float32 on both sides; a detector at float32's noise floor is
judged against the FP64 twin, see helpers.assert_dets_noise_aware).  The edges cover every grid size of the
uneven-batch workload (360, 384, 480, 512, 640, 720)."""
import numpy as np
import pytest

from jmultislicelib_amd import Engine, configs
from oracle import oracle as O

from helpers import assert_dets_noise_aware

pytestmark = pytest.mark.gpu
TOL = 1e-4


def test_concurrent_uneven_jobs_vs_oracle():
    rng = np.random.default_rng(2)
    jobs = [configs.random_job(rng, edge=e, nslices=s, npos=p)
            for e, s, p in ((360, 14, 5), (384, 9, 4), (480, 11, 3), (512, 8, 3), (640, 7, 2),
                                  (720, 6, 2))]
    engines = []
    for j in jobs:
        eng = Engine().setup(j.setup)
        eng.set_batch(len(j.dx))
        var = np.zeros((len(j.dx), len(j.setup.objslc)), np.int32)
        eng.stage_scan(j.dx, j.dy, None, np.ones(1, np.float32), var)
        engines.append(eng)
    for eng, j in zip(engines, jobs):  # all enqueued before any wait
        eng.run_staged(0, len(j.dx))
    for eng in engines:
        eng.sync()
    for eng, j in zip(engines, jobs):
        sp = j.setup
        g = eng.scan_results()
        var = np.zeros((len(j.dx), len(sp.objslc)), np.int32)
        w = np.ones(1, np.float32)
        ref = O.stem_scan(sp, j.dx, j.dy, w, var, prec=32, nthreads=4)
        assert np.isfinite(g).all()
        assert_dets_noise_aware(g, ref, lambda: O.stem_scan(sp, j.dx, j.dy, w, var, prec=64, nthreads=4), TOL,
                                f"job {sp.nx}^2 x {len(sp.objslc)} slices, {len(j.dx)} positions:")
        eng.close()
