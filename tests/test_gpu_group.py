"""Grouped launches over an uneven batch (jms_hip_group): jobs of one grid
edge with different slice counts, positions, passes, variant tables,
detection planes and batch sizes run as one wave map.  Each member's
results equal those of the same engine set-up run on its own
(jms_hip_run_staged: the same arithmetic on the same per-wave data) to
float32 rounding: every intensity within 1e-6 of the largest (the grouped
kernels are instantiations of their own, whose instruction selection may
round a last bit differently).  A second run of the group repeats the first
bit for bit, and every member matches the oracle within 1e-4 relative L2
per detector (noise-aware for rings at float32's noise floor)."""
import numpy as np
import pytest

from jmultislicelib_amd import Engine, Group, JMSError, configs
from oracle import oracle as O

from helpers import assert_dets_noise_aware, random_setup

pytestmark = pytest.mark.gpu


def _jobs(edge):
    rng = np.random.default_rng(7)
    spec = ((12, 5, 1, 0), (7, 3, 2, 3), (7, 4, 1, 2), (3, 2, 1, 0), (10, 6, 2, 4))
    out = []
    for k, (ns, npos, npass, ndetper) in enumerate(spec):
        j = configs.random_job(rng, edge=edge, nslices=ns, npos=npos)
        sp = j.setup
        det, ndetslc = configs.detection_planes(ndetper, ns)
        upd = {"det_objslc": det, "ndetslc": ndetslc}
        if k == 1:  # two variants of slice type 0: the variant table matters
            g0 = sp.gratings[0]
            upd["gratings"] = [np.concatenate([g0, np.roll(g0, 7, axis=2)]), sp.gratings[1]]
            upd["nvarslc"] = [2, 1]
        sp = configs.Setup(**{**sp.__dict__, **upd})
        var = np.zeros((npos * npass, ns), np.int32)
        if k == 1:
            var[:, sp.objslc == 0] = rng.integers(0, 2, (npos * npass, int((sp.objslc == 0).sum())))
        w = np.full(npass, 1.0 / npass, np.float32) if npass == 1 else np.array([0.3, 0.7], np.float32)
        out.append((sp, j.dx, j.dy, w, var))
    return out


def _engine(sp, dx, dy, w, var, batch=None):
    eng = Engine().setup(sp)
    if batch:
        eng.set_batch(batch)
    eng.stage_scan(dx, dy, None, w, var)
    return eng


@pytest.mark.parametrize("edge", [128, 192, 256, 360, 512, 625, 768, 1024])
def test_group_equals_separate_runs_and_oracle(edge):
    jobs = _jobs(edge)
    # member 2 gets a batch of two waves: it runs over several rounds
    grouped = [_engine(*j, batch=2 if k == 2 else None) for k, j in enumerate(jobs)]
    alone = [_engine(*j) for j in jobs]
    grp = Group(grouped)
    grp.run()
    grp.sync()
    for eng, (sp, dx, dy, w, var) in zip(alone, jobs):
        eng.run_staged(0, len(dx))
        eng.sync()
    got = [e.scan_results() for e in grouped]
    for k, (g, e) in enumerate(zip(got, alone)):
        r = e.scan_results()
        assert np.isfinite(g).all() and np.abs(g).max() > 0
        d = np.abs(g - r).max() / np.abs(r).max()
        print(f"member {k}: max |grouped - alone| / max {d:.2e}")
        assert d <= 1e-6, f"member {k}: grouped vs alone {d:.2e}"
    # the group again (tables reused), then after restaging one member with fewer positions (tables rebuilt)
    grp.run()
    grp.sync()
    for k, e in enumerate(grouped):
        assert np.array_equal(e.scan_results(), got[k])
    sp, dx, dy, w, var = jobs[0]
    grouped[0].stage_scan(dx[:2], dy[:2], None, w, var[:2 * len(w)])
    grp.run()
    grp.sync()
    close = lambda a, b: np.abs(a - b).max() <= 1e-6 * np.abs(b).max()
    assert close(grouped[0].scan_results(), got[0][:2])
    for k in range(1, len(jobs)):
        assert close(grouped[k].scan_results(), got[k]), k
    # the oracle
    for k, (sp, dx, dy, w, var) in enumerate(jobs):
        ref = O.stem_scan(sp, dx, dy, w, var, prec=32, nthreads=4)
        assert_dets_noise_aware(got[k], ref, lambda: O.stem_scan(sp, dx, dy, w, var, prec=64, nthreads=4), 1e-4,
                                f"grouped member {k} ({edge}^2, {len(sp.objslc)} slices)")
    grp.close()
    for e in grouped + alone:
        e.close()


def test_group_long_columns_2048():
    """2048^2 (the long-column kernel in its grouped form): three jobs with
    different slice counts, positions and detection planes as one group; each
    equals its own staged run to float32 rounding, and the smallest matches
    the oracle."""
    jobs = []
    for k, (ns, npos, ndetper) in enumerate(((4, 2, 0), (2, 3, 1), (3, 1, 2))):
        sp = random_setup(2048, 2048, nslices=ns, nvar=(2, 1), ndetper=ndetper, seed=50 + k)
        dx = np.linspace(0.1, 0.9, npos).astype(np.float32)
        dy = np.linspace(0.7, 0.2, npos).astype(np.float32)
        jobs.append((sp, dx, dy, np.ones(1, np.float32), configs.variant_sequences(sp, npos, seed=k)))
    grouped = [_engine(*j) for j in jobs]
    alone = [_engine(*j) for j in jobs]
    grp = Group(grouped)
    grp.run()
    grp.sync()
    for eng, j in zip(alone, jobs):
        eng.run_staged(0, len(j[1]))
        eng.sync()
    for k, (g, e) in enumerate(zip(grouped, alone)):
        r = e.scan_results()
        d = np.abs(g.scan_results() - r).max() / np.abs(r).max()
        print(f"member {k}: max |grouped - alone| / max {d:.2e}")
        assert np.abs(r).max() > 0 and d <= 1e-6, f"member {k}: grouped vs alone {d:.2e}"
    sp, dx, dy, w, var = jobs[1]
    ref = O.stem_scan(sp, dx, dy, w, var, prec=32, nthreads=3)
    assert_dets_noise_aware(grouped[1].scan_results(), ref,
                            lambda: O.stem_scan(sp, dx, dy, w, var, prec=64, nthreads=3), 1e-4,
                            "grouped member 1 (2048^2)")
    grp.close()
    for e in grouped + alone:
        e.close()


def test_group_refuses_mixed_grids_and_stale_members():
    a = Engine()
    b = Engine()
    c = Engine()
    a._c("jms_hip_set_grid", 360, 360, 2.0, 2.0, 0.0025)
    b._c("jms_hip_set_grid", 384, 384, 2.0, 2.0, 0.0025)
    c._c("jms_hip_set_grid", 1024, 1024, 2.0, 2.0, 0.0025)
    with pytest.raises(JMSError) as ex:
        Group([a, b])
    assert ex.value.code == 4
    with pytest.raises(JMSError) as ex:
        Group([a, c])
    assert ex.value.code == 4
    Group([c]).close()  # every supported grid groups (1024: the long-column kernel)
    with pytest.raises(JMSError) as ex:
        Group([a, a])
    assert ex.value.code == 3
    g = Group([a])  # set up grid only: running it reports the missing set-up
    with pytest.raises(JMSError):
        g.run()
    g.close()
    d = Engine()  # a member destroyed before the group: the group refuses to run, then closes cleanly
    e2 = Engine()
    d._c("jms_hip_set_grid", 360, 360, 2.0, 2.0, 0.0025)
    e2._c("jms_hip_set_grid", 360, 360, 2.0, 2.0, 0.0025)
    g = Group([d, e2])
    d.close()
    with pytest.raises(JMSError) as ex:
        g.run()
    assert ex.value.code == 19
    g.close()
    e2.close()
    g = Group([b])  # a member's grid changed after the group was made
    b._c("jms_hip_set_grid", 360, 360, 2.0, 2.0, 0.0025)
    with pytest.raises(JMSError) as ex:
        g.run()
    assert ex.value.code == 18
    g.close()
    for e in (a, b, c):
        e.close()
