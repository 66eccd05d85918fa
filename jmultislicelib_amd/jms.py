"""Python host-side mirror of the reference multislice interface.

``JMultiSlice`` forwards to the flat C-ABI of libjmultislice_hip.so (the
reference's exported function names, src/JMultiSliceLib.h:111-380) and keeps
their argument meaning and 0/>0 error codes.  ``Engine`` wraps the thin
engine ABI (include/jms_hip.h) used for batched scans and timing.
All arrays are numpy; complex arrays are complex64 of shape (ny, nx) in FFT
order, the reference's memory layout.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import CJProbeParams, lib, ptr

_JMS_CODE_CPU = 1
_JMS_CODE_GPU = 2
_JMS_ACCMODE_NONE = 0
_JMS_ACCMODE_INTEGRATE = 1
_JMS_DETECT_INTEGRATED = 1
_JMS_DETECT_IMAGE = 2
_JMS_DETECT_DIFFRACTION = 4
_JMS_DETECT_WAVEREAL = 8
_JMS_DETECT_WAVEFOURIER = 16
_JMS_DETECT_IMAGE_AVG = 32
_JMS_DETECT_DIFFR_AVG = 64
_JMS_DETECT_WAVER_AVG = 128
_JMS_DETECT_WAVEF_AVG = 256


class JMSError(RuntimeError):
    def __init__(self, fn: str, code: int):
        super().__init__(f"{fn} returned error code {code}")
        self.code = code


def _chk(fn: str, rc: int) -> int:
    if rc != 0:
        raise JMSError(fn, rc)
    return rc


def _c64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex64)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


class JMultiSlice:
    """Facade over the library's process-global CJMultiSlice (the flat C-ABI)."""

    def __init__(self):
        self.L = lib()

    # -- physics / numerics ------------------------------------------------
    def SetHighTension(self, ht: float) -> float:
        return self.L.SetHighTension(ht)

    def GetWaveLength(self) -> float:
        return self.L.GetWaveLength()

    def SetSupercellSize(self, a: float, b: float, c: float) -> None:
        self.L.SetSupercellSizeABC(a, b, c)

    def SetGridSize(self, nx: int, ny: int) -> None:
        self.L.SetGridSize(nx, ny)

    def GetGridSize(self):
        a, b = C.c_int(), C.c_int()
        self.L.GetGridSize(C.byref(a), C.byref(b))
        return a.value, b.value

    def SeedRngEx(self, seed: int) -> None:
        self.L.SeedRngEx(seed)

    def DiffractionDescan(self, on: bool) -> None:
        self.L.DiffractionDescan(int(bool(on)))

    def SetDiffractionDescanN(self, whichcode, ndescanx, ndescany) -> None:
        self.L.SetDiffractionDescanN(whichcode, int(ndescanx), int(ndescany), 0)

    def SetDiffractionDescan(self, whichcode, descanx, descany) -> None:
        self.L.SetDiffractionDescan(whichcode, descanx, descany, 0)

    def SetDiffractionDescanMRad(self, whichcode, descanx, descany) -> None:
        self.L.SetDiffractionDescanMRad(whichcode, descanx, descany, 0)

    def GetDetetionSliceNum(self) -> int:
        return self.L.GetDetetionSliceNum()

    def GetDetNum(self) -> int:
        return self.L.GetDetNum()

    # -- calculators ---------------------------------------------------------
    def CalculatePropagator(self, fthick: float, otx: float, oty: float, ntype: int = 0) -> np.ndarray:
        nx, ny = self.GetGridSize()
        pro = np.zeros((ny, nx), np.complex64)
        _chk("CalculatePropagator", self.L.CalculatePropagator(fthick, otx, oty, ptr(pro), ntype))
        return pro

    def CalculateRingDetector(self, beta0, beta1, phi0=0.0, phi1=0.0, theta0x=0.0, theta0y=0.0,
                              sdsprofile: str = "", kmom1: int = 0, mask: bool = True):
        nx, ny = self.GetGridSize()
        det = np.zeros((ny, nx), np.float32)
        msk = np.zeros(nx * ny, np.int32) if mask else None
        ml = C.c_int(0)
        _chk("CalculateRingDetector",
             self.L.CalculateRingDetector(beta0, beta1, phi0, phi1, theta0x, theta0y,
                                          sdsprofile.encode(), ptr(det), C.byref(ml), ptr(msk), kmom1))
        return det, (msk[: ml.value].copy() if mask else None)

    def CalculateProbeWaveFourier(self, alpha_mrad: float, alpha_rs: float = 0.03, alpha_x0=0.0,
                                  alpha_y0=0.0, alpha_asym=0.0, alpha_adir=0.0, btx=0.0,
                                  bty=0.0, abrr=None) -> np.ndarray:
        """Probe in Fourier space; abrr: [k][x, y] axial aberration coefficients
        [nm] in the library's interface order (image shift, defocus, 2-fold
        astigmatism, coma, 3-fold astigmatism, spherical, ...)."""
        nx, ny = self.GetGridSize()
        prm = CJProbeParams()
        ab = None
        if abrr is not None:
            ab = np.ascontiguousarray(abrr, np.float32).reshape(-1)
            prm.m_abrr_coeff = ab.ctypes.data_as(C.POINTER(C.c_float))
            prm.m_abrr_num = len(ab) // 2
        prm.m_alpha = alpha_mrad
        prm.m_alpha_rs = alpha_rs
        prm.m_alpha_x0 = alpha_x0
        prm.m_alpha_y0 = alpha_y0
        prm.m_alpha_asym = alpha_asym
        prm.m_alpha_adir = alpha_adir
        prm.m_btx = btx
        prm.m_bty = bty
        if ab is None:
            prm.m_abrr_num = 0
        wav = np.zeros((ny, nx), np.complex64)
        _chk("CalculateProbeWaveFourier", self.L.CalculateProbeWaveFourier(C.byref(prm), ptr(wav)))
        return wav

    def GetUnscrambleHash(self) -> np.ndarray:
        nx, ny = self.GetGridSize()
        h = np.zeros(nx * ny, np.uint32)
        _chk("GetUnscrambleHash", self.L.GetUnscrambleHash(ptr(h)))
        return h

    # -- setup chain -----------------------------------------------------------
    def PhaseGratingSetup(self, whichcode, nx, ny, nslc, nvarmax, nslcvar) -> int:
        v = _i32(nslcvar)
        return self.L.PhaseGratingSetup(whichcode, nx, ny, nslc, nvarmax, ptr(v))

    def ObjectSliceSetup(self, objslc) -> int:
        o = _i32(objslc)
        return self.L.ObjectSliceSetup(len(o), ptr(o))

    def PropagatorSetup(self, whichcode, npro, proidx) -> int:
        p = _i32(proidx)
        return self.L.PropagatorSetup(whichcode, npro, ptr(p))

    def DetectorSetup(self, whichcode, ndetper, ndetint, imagedet, nthreads_CPU=1) -> int:
        return self.L.DetectorSetup(whichcode, ndetper, ndetint, imagedet, nthreads_CPU)

    def SetPhaseGratingData(self, whichcode, islc, pgr) -> int:
        g = _c64(pgr)
        nvar = g.shape[0] if g.ndim == 3 else 1
        return self.L.SetSlicePhaseGratings(whichcode, islc, nvar, ptr(g))

    def SetPropagatorData(self, whichcode, ipro, pro) -> int:
        p = _c64(pro)
        return self.L.SetPropagatorData(whichcode, ipro, ptr(p))

    def SetDetectorData(self, whichcode, idet, det, msk=None) -> int:
        d = _f32(det)
        m = _i32(msk) if msk is not None else None
        return self.L.SetDetectorData(whichcode, idet, ptr(d), 0 if m is None else len(m), ptr(m))

    def InitCore(self, whichcode, ncpu=1) -> int:
        return self.L.InitCore(whichcode, ncpu)

    def SetIncidentWave(self, whichcode, wav, transpose=False) -> int:
        w = _c64(wav)
        return self.L.SetIncidentWave(whichcode, ptr(w), int(transpose))

    def SetIncidentWavePhaseplate(self, whichcode, ppl, transpose=False) -> int:
        p = _f32(ppl)
        return self.L.SetIncidentWavePhaseplate(whichcode, ptr(p), int(transpose))

    # -- multislice --------------------------------------------------------------
    def OffsetIncomingWave(self, whichcode, dx, dy, dz, iThread=0, ibtx=0, ibty=0, bcppl=True) -> int:
        return self.L.OffsetIncomingWave(whichcode, dx, dy, dz, iThread, ibtx, ibty, int(bcppl))

    def SetIncomingWaveGPU(self, wav, transpose=False) -> int:
        w = _c64(wav)
        return self.L.SetIncomingWaveGPU(ptr(w), int(transpose))

    def GPUMultislice(self, islc0: int, accmode: int, weight: float = 1.0) -> int:
        return self.L.GPUMultislice(islc0, accmode, weight)

    def _chan_buffer(self, which):
        nx, ny = self.GetGridSize()
        nds = self.GetDetetionSliceNum()
        if which == _JMS_DETECT_INTEGRATED:
            return np.zeros((nds, self.GetDetNum()), np.float32)
        if which in (_JMS_DETECT_DIFFRACTION, _JMS_DETECT_IMAGE, _JMS_DETECT_IMAGE_AVG, _JMS_DETECT_DIFFR_AVG):
            return np.zeros((nds, ny, nx), np.float32)
        if which in (_JMS_DETECT_WAVEFOURIER, _JMS_DETECT_WAVEREAL, _JMS_DETECT_WAVER_AVG, _JMS_DETECT_WAVEF_AVG):
            return np.zeros((nds, ny, nx), np.complex64)
        raise ValueError("unsupported result channel")

    def GetResult(self, whichcode, whichresult) -> np.ndarray:
        out = self._chan_buffer(whichresult)
        _chk("GetResult", self.L.GetResult(whichcode, whichresult, ptr(out), 0))
        return out

    def GetAvgResult(self, whichcode, whichresult):
        """(averaged channel, weight) -- GetAvgResult (JMultiSliceLib.h:354)."""
        out = self._chan_buffer(whichresult)
        w = C.c_float(0.0)
        _chk("GetAvgResult", self.L.GetAvgResult(whichcode, whichresult, ptr(out), C.byref(w), 0))
        return out, w.value

    def GetImageAveragingWeight(self, whichcode) -> float:
        return self.L.GetImageAveragingWeight(whichcode, 0)

    def GetWaveAveragingWeight(self, whichcode) -> float:
        return self.L.GetWaveAveragingWeight(whichcode, 0)

    def ResetImageAveraging(self, whichcode) -> int:
        return self.L.ResetImageAveraging(whichcode, 0)

    def ResetWaveAveraging(self, whichcode) -> int:
        return self.L.ResetWaveAveraging(whichcode, 0)

    def GetRandomVariantSequence(self, nobjslc: int) -> np.ndarray:
        v = np.zeros(nobjslc, np.int32)
        _chk("GetRandomVariantSequence", self.L.GetRandomVariantSequence(ptr(v)))
        return v

    def GPUMultisliceScan(self, dx, dy, dz, weights, var=None) -> np.ndarray:
        dx, dy = _f32(dx), _f32(dy)
        dz = _f32(dz) if dz is not None else None
        w = _f32(weights)
        npos, npass = len(dx), len(w)
        out = np.zeros((npos, self.GetDetetionSliceNum(), self.GetDetNum()), np.float32)
        if var is None:
            rc = self.L.GPUMultisliceScan(npos, ptr(dx), ptr(dy), ptr(dz), npass, ptr(w), ptr(out))
        else:
            v = _i32(var)
            rc = self.L.GPUMultisliceScanVar(npos, ptr(dx), ptr(dy), ptr(dz), npass, ptr(w), ptr(v),
                                             ptr(out))
        _chk("GPUMultisliceScan", rc)
        return out

    def Cleanup(self) -> int:
        return self.L.Cleanup()


class Engine:
    """Batched multislice engine (include/jms_hip.h) on the current GPU."""

    def __init__(self):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.jms_hip_create(C.byref(h))
        if rc != 0:
            raise JMSError("jms_hip_create", rc)
        self.h = h

    def close(self):
        if self.h:
            self.L.jms_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, name, *args):
        rc = getattr(self.L, name)(self.h, *args)
        if rc != 0:
            msg = self.L.jms_hip_last_error(self.h)
            raise JMSError(f"{name} ({msg.decode() if msg else ''})", rc)

    def setup(self, sp) -> "Engine":
        """Upload a specimen/optics setup (jmultislicelib_amd.configs.Setup)."""
        self._c("jms_hip_set_grid", sp.nx, sp.ny, sp.a0[0], sp.a0[1], sp.wl)
        nv = _i32(sp.nvarslc)
        self._c("jms_hip_setup_gratings", sp.nx, sp.ny, len(nv), ptr(nv))
        self.sp = sp
        if sp.gratings is None:  # generated on the device from the atom sets, or later by the caller
            from .configs import SIG_INT, SIGMA_POT
            for i, sets in enumerate(sp.grating_sites or []):
                self.gaussian_gratings(i, sets, SIGMA_POT, SIG_INT)
        else:
            for i, g in enumerate(sp.gratings):
                g = _c64(g)
                self._c("jms_hip_upload_gratings", i, g.shape[0], ptr(g))
        pi = _i32(sp.proidx)
        self._c("jms_hip_setup_propagators", len(sp.propagators), len(pi), ptr(pi))
        for i, p in enumerate(sp.propagators):
            p = _c64(p)
            self._c("jms_hip_upload_propagator", i, ptr(p))
        o = _i32(sp.objslc)
        self._c("jms_hip_setup_objects", len(o), ptr(o))
        d = _i32(sp.det_objslc)
        self._c("jms_hip_setup_detection", sp.ndetslc, ptr(d), len(sp.detectors), sp.flags)
        for i, (det, msk) in enumerate(sp.detectors):
            det = _f32(det)
            m = _i32(msk) if msk is not None else None
            self._c("jms_hip_upload_detector", i, ptr(det), 0 if m is None else len(m), ptr(m))
        w0 = _c64(sp.probe)
        self._c("jms_hip_set_probe", ptr(w0))
        self.sp = sp
        return self

    def gaussian_gratings(self, islc, sites, sigma, sig_int):
        """Generate the variants of structure slice islc on the device from
        per-variant site lists [(xy [A] (k, 2), z (k,)), ...]."""
        nat = _i32([len(z) for _, z in sites])
        xy = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64).reshape(-1, 2) for x, _ in sites]),
                                  np.float64)
        z = _i32(np.concatenate([np.asarray(z) for _, z in sites]))
        self._c("jms_hip_gaussian_gratings", islc, len(sites), ptr(nat), ptr(xy), ptr(z), float(sigma),
                float(sig_int))

    def formfactor_gratings(self, islc, ff2d, sites, cz, bwl_pot=False, bwl_pgr=True):
        """Generate the variants of structure slice islc on the device from
        projected atomic form factors ff2d (ntype, ny, nx) [nm] (FFT order)
        and per-variant atom lists [(xy [nm] (k, 2), type (k,), occ (k,) or
        None), ...]; cz = slice thickness [nm] (the device path of
        CJEElaSca::CalculatePhasegratingGPU, JEElaSca.cpp:978-1159)."""
        ff = np.ascontiguousarray(ff2d, np.complex64)
        if ff.ndim != 3 or ff.shape[1:] != (self.sp.ny, self.sp.nx):
            raise ValueError(f"ff2d must be (ntype, {self.sp.ny}, {self.sp.nx})")
        if not sites:
            raise ValueError("no atom sets")
        for x, t, o in sites:
            k = len(np.asarray(t).reshape(-1))
            if np.asarray(x).size != 2 * k or (o is not None and np.asarray(o).size != k):
                raise ValueError("each atom set needs 2 coordinates and one occupancy per type entry")
        nat = _i32([len(t) for _, t, _ in sites])
        xy = _f32(np.concatenate([np.asarray(x, np.float32).reshape(-1, 2) for x, _, _ in sites]).reshape(-1))
        typ = _i32(np.concatenate([np.asarray(t).reshape(-1) for _, t, _ in sites]))
        occ = _f32(np.concatenate([np.ones(len(t), np.float32) if o is None else np.asarray(o, np.float32)
                                   for _, t, o in sites]))
        flags = (1 if bwl_pot else 0) | (2 if bwl_pgr else 0)
        self._c("jms_hip_formfactor_gratings", islc, len(sites), ff.shape[0], ptr(ff), ptr(nat), ptr(xy), ptr(typ),
                ptr(occ), float(cz), flags)

    def gratings(self, islc, nvar) -> np.ndarray:
        out = np.zeros((nvar, self.sp.ny, self.sp.nx), np.complex64)
        self._c("jms_hip_get_gratings", islc, nvar, ptr(out))
        return out

    def probe(self, alpha_mrad, alpha_rs=0.03, alpha_x0=0.0, alpha_y0=0.0, alpha_asym=0.0, alpha_adir=0.0,
              btx=0.0, bty=0.0, abrr=None) -> np.ndarray:
        """Generate the incident probe on the device (it becomes the scan's
        incoming wave) and return a host copy."""
        prm = _f32([alpha_mrad, alpha_rs, alpha_x0, alpha_y0, alpha_asym, alpha_adir, btx, bty])
        ab = None if abrr is None else _f32(np.asarray(abrr, np.float32).reshape(-1))
        out = np.zeros((self.sp.ny, self.sp.nx), np.complex64)
        self._c("jms_hip_probe", ptr(prm), 0 if ab is None else len(ab) // 2, ptr(ab), ptr(out))
        return out

    def stage_scan(self, dx, dy, dz, weights, var):
        dx, dy = _f32(dx), _f32(dy)
        dz = _f32(dz) if dz is not None else None
        w = _f32(weights)
        v = _i32(var)
        self.npos = len(dx)
        self._c("jms_hip_stage_scan", len(dx), ptr(dx), ptr(dy), ptr(dz), len(w), ptr(w), ptr(v))

    def run_staged(self, pos0: int, npos: int, islc0: int = 0):
        self._c("jms_hip_run_staged", pos0, npos, islc0)

    def sync(self):
        self._c("jms_hip_sync")

    def scan_results(self) -> np.ndarray:
        out = np.zeros((self.npos, self.sp.ndetslc, len(self.sp.detectors)), np.float32)
        self._c("jms_hip_get_scan_results", ptr(out))
        return out

    def set_batch(self, nwave: int):
        self._c("jms_hip_set_batch", nwave)

    def set_graph(self, on: bool):
        self._c("jms_hip_set_graph", int(on))

    def set_timing(self, on: bool):
        self._c("jms_hip_set_timing", int(on))

    def timing(self, kind: int):
        ms, cnt = C.c_double(), C.c_longlong()
        self._c("jms_hip_get_timing", kind, C.byref(ms), C.byref(cnt))
        return ms.value, cnt.value

    def wave_offset(self, dx, dy, dz=0.0, ibtx=0, ibty=0, use_cppl=True):
        self._c("jms_hip_wave_offset", dx, dy, dz, ibtx, ibty, int(use_cppl))

    def wave_set(self, wav):
        w = _c64(wav)
        self._c("jms_hip_wave_set", ptr(w))

    def wave_get(self) -> np.ndarray:
        out = np.zeros((self.sp.ny, self.sp.nx), np.complex64)
        self._c("jms_hip_wave_get", ptr(out))
        return out

    def multislice(self, var, weight=1.0, accumulate=False, islc0=0):
        v = _i32(var) if var is not None and len(var) else np.zeros(1, np.int32)
        self._c("jms_hip_multislice", islc0, ptr(v), weight, int(accumulate))

    def _buffer(self, which):
        sp = self.sp
        if which == _JMS_DETECT_INTEGRATED:
            return np.zeros((sp.ndetslc, len(sp.detectors)), np.float32)
        if which in (_JMS_DETECT_DIFFRACTION, _JMS_DETECT_IMAGE, _JMS_DETECT_IMAGE_AVG, _JMS_DETECT_DIFFR_AVG):
            return np.zeros((sp.ndetslc, sp.ny, sp.nx), np.float32)
        return np.zeros((sp.ndetslc, sp.ny, sp.nx), np.complex64)

    def result(self, which=_JMS_DETECT_INTEGRATED) -> np.ndarray:
        out = self._buffer(which)
        self._c("jms_hip_get_result", which, ptr(out))
        return out

    def avg_result(self, which):
        """(averaged channel, weight) of the single-wave path."""
        out = self._buffer(which)
        w = C.c_float(0.0)
        self._c("jms_hip_get_avg_result", which, ptr(out), C.byref(w))
        return out, w.value

    def set_descan(self, dnx: int, dny: int):
        self._c("jms_hip_set_descan", int(dnx), int(dny))

    def averaging_weights(self):
        wi, ww = C.c_float(0.0), C.c_float(0.0)
        self._c("jms_hip_get_averaging_weights", C.byref(wi), C.byref(ww))
        return wi.value, ww.value

    def reset_averaging(self, group: int):
        self._c("jms_hip_reset_averaging", group)

    def readout_det_avg(self, which):
        """(intensities of the averaged waves, wave-group weight)."""
        out = self._buffer(which)
        w = C.c_float(0.0)
        self._c("jms_hip_readout_det_avg", which, ptr(out), C.byref(w))
        return out, w.value


class Group:
    """Grouped launches over engines with one grid (jms_hip_group): run()
    runs every member's staged scan in full, the waves of all members
    advancing through one row and one column launch per slice step.  The
    results are read from each member (Engine.scan_results) as after
    Engine.run_staged; Engine.sync on any member waits for the group."""

    _CREATE_ERRORS = {2: "invalid engine list", 3: "an engine given twice", 4: "grids or devices differ",
                      5: "grid not supported by the grouped kernels"}

    def __init__(self, engines):
        self.L = lib()
        self.members = list(engines)
        arr = (C.c_void_p * len(self.members))(*[e.h.value for e in self.members])
        h = C.c_void_p()
        rc = self.L.jms_hip_group_create(C.byref(h), arr, len(self.members))
        if rc != 0:
            raise JMSError(f"jms_hip_group_create ({self._CREATE_ERRORS.get(rc, '')})", rc)
        self.h = h

    def run(self):
        rc = self.L.jms_hip_group_run(self.h)
        if rc != 0:
            msg = self.L.jms_hip_group_last_error(self.h)
            raise JMSError(f"jms_hip_group_run ({msg.decode() if msg else ''})", rc)

    def sync(self):
        for e in self.members:
            e.sync()

    def close(self):
        if self.h:
            self.L.jms_hip_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fft2d(a: np.ndarray, direction: int) -> np.ndarray:
    """2-D FFT on the GPU of (..., ny, nx) complex64 data; direction < 0
    forward exp(-i), > 0 backward exp(+i), unscaled."""
    d = _c64(a).copy()
    ny, nx = d.shape[-2:]
    nb = int(np.prod(d.shape[:-2])) if d.ndim > 2 else 1
    _chk("jms_hip_fft2d", lib().jms_hip_fft2d(nx, ny, nb, ptr(d), direction))
    return d


def grid_supported(nx: int, ny: int) -> bool:
    return bool(lib().jms_hip_grid_supported(nx, ny))
