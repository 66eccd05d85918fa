"""The allocation-free batch entry (jms_hip_run_batch_noalloc) through the
C-ABI against the existing staged-scan entry (jms_hip_stage_scan +
jms_hip_run_staged), which runs the same kernels with library-allocated
buffers: identical detector sums on two calls in a row, the second with a
smaller batch, every buffer owned by the caller (device memory allocated
here with hipMalloc of the HIP runtime the library uses) and the readout
partial sums in the library's fixed scratch.  Both are checked against the
oracle too (every detector within 1e-4)."""
import ctypes as C

import numpy as np
import pytest

from jmultislicelib_amd import Engine, configs
from oracle import oracle as O

from helpers import assert_dets

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime the library is linked
    with (the already-loaded libamdhip64.so.7)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(1, a.nbytes)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def get(self, p, shape, dtype):
        out = np.zeros(shape, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _noalloc(eng, hip, npos, npass, dx, dy, w, var):
    sp = eng.sp
    nwave = npos * npass
    ppar = np.zeros((npos, npass, 4), np.float32)
    ppar[:, :, 0] = np.asarray(dx, np.float32)[:, None]
    ppar[:, :, 1] = np.asarray(dy, np.float32)[:, None]
    d_ppar = hip.put(ppar)
    d_var = hip.put(np.asarray(var, np.int32))
    d_w = hip.put(np.asarray(w, np.float32))
    d_waves = hip.alloc(8 * nwave * sp.nx * sp.ny)
    nres = npos * sp.ndetslc * len(sp.detectors)
    d_res = hip.put(np.full(nres, -1.0, np.float32))
    rc = eng.L.jms_hip_run_batch_noalloc(eng.h, 0, nwave, d_waves, d_ppar, d_var, npos, npass, d_w, d_res)
    assert rc == 0, eng.L.jms_hip_last_error(eng.h)
    eng.sync()
    out = hip.get(d_res, (npos, sp.ndetslc, len(sp.detectors)), np.float32)
    hip.free()
    return out


def test_noalloc_batch_matches_staged_scan():
    hip = _Hip()
    sp = configs.perovskite_setup(n=128, ncell=4, nslices=12, nvar=3, ndetper=4)
    rng = np.random.default_rng(31)
    npass = 3
    eng = Engine().setup(sp)
    for npos in (6, 2):  # the second call with a smaller batch
        dx = rng.uniform(0, 0.78, npos).astype(np.float32)
        dy = rng.uniform(0, 0.78, npos).astype(np.float32)
        w = np.array([0.5, 0.3, 0.2], np.float32)
        var = configs.variant_sequences(sp, npos * npass, seed=npos)
        got = _noalloc(eng, hip, npos, npass, dx, dy, w, var)
        eng.set_batch(npos * npass)
        eng.stage_scan(dx, dy, None, w, var)
        eng.run_staged(0, npos)
        eng.sync()
        want = eng.scan_results()
        print(f"npos {npos}: noalloc vs staged max |diff| {np.abs(got - want).max():.3e}")
        assert np.array_equal(got, want)
        ref = O.stem_scan(sp, dx, dy, w, var, prec=32, nthreads=4)
        assert_dets(got, ref, 1e-4, f"noalloc npos {npos}")
    eng.close()


def test_noalloc_refuses_oversize_and_bad_shapes():
    hip = _Hip()
    sp = configs.perovskite_setup(n=64, ncell=2, nslices=4, nvar=1)
    eng = Engine().setup(sp)
    L = eng.L
    p = hip.put(np.zeros(16, np.float32))
    # nwave != npos * npass, and over the grid limit: refused before any launch
    assert L.jms_hip_run_batch_noalloc(eng.h, 0, 5, p, p, p, 2, 2, p, p) == 2
    assert L.jms_hip_run_batch_noalloc(eng.h, 0, 70000, p, p, p, 70000, 1, p, p) == 2
    # partial sums beyond the fixed scratch (64 columns: 4 tiles x 3 detectors
    # per wave; 65535 waves fit, so a larger plane count is needed): refused
    sp2 = configs.perovskite_setup(n=64, ncell=2, nslices=40, nvar=1, ndetper=1)
    eng2 = Engine().setup(sp2)
    assert eng2.L.jms_hip_run_batch_noalloc(eng2.h, 0, 60000, p, p, p, 60000, 1, p, p) == 5
    eng2.close()
    eng.close()
    hip.free()
