"""A specimen loaded from EMS/SLI slice files runs through the HIP slice
loop exactly as the same specimen built in memory, and matches the oracle
within 1e-4 relative L2 (float32 on both sides)."""
import numpy as np
import pytest

from jmultislicelib_amd import Engine, configs, io as jio
from oracle import oracle as O

from helpers import assert_dets

pytestmark = pytest.mark.gpu


def test_slice_file_specimen_vs_oracle(tmp_path):
    sp = configs.perovskite_setup(n=64, ncell=2, nslices=12, nvar=3, ndetper=4, flags=1)
    paths = []
    for i, g in enumerate(sp.gratings):
        p = str(tmp_path / configs.slice_file_name("pv", i))
        jio.write_ems(p, g, sp.a0[0], sp.a0[1], sp.slice_thickness, sp.ht)
        paths.append(p)
    s2 = configs.setup_from_slice_files(paths, 12, ndetper=4)
    var = configs.variant_sequences(s2, 2, seed=5)
    res = []
    for s in (sp, s2):
        eng = Engine().setup(s)
        for r in range(2):
            eng.wave_offset(0.2, 0.1, 0.0)
            eng.multislice(var[r], 0.5, r > 0)
        res.append(eng.result(1))
    assert np.array_equal(res[0], res[1])
    ref = np.zeros_like(res[1])
    for r in range(2):
        win = O.offset_wave(s2.probe, s2.a0[0], s2.a0[1], s2.wl, 0.2, 0.1)
        ref += O.multislice(s2, win, var[r], 0.5, prec=32)[0]
    assert_dets(res[1], ref, 1e-4, "slice-file specimen")
