"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker.  The product library never uses
it.  See jms_oracle.cpp for the restated reference functions.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(_HERE, "liboracle.so")
_L = None

i_, f_, d_, p_, ll_ = C.c_int, C.c_float, C.c_double, C.c_void_p, C.c_longlong
_SIGS = {
    "orc_wavelength": (f_, [f_]),
    "orc_fft1d_f32": (i_, [p_, i_, i_, i_]),
    "orc_fft1d_f64": (i_, [p_, i_, i_, i_]),
    "orc_fft2d_f32": (i_, [p_, i_, i_, i_]),
    "orc_fft2d_f64": (i_, [p_, i_, i_, i_]),
    "orc_propagator": (i_, [i_, i_, f_, f_, f_, f_, f_, f_, i_, p_]),
    "orc_ring_detector": (i_, [i_, i_, f_, f_, f_, i_, f_, f_, f_, f_, f_, f_, p_, i_, f_, p_, p_, p_, i_]),
    "orc_probe_wave": (i_, [i_, i_, f_, f_, f_, f_, f_, f_, f_, f_, f_, f_, f_, p_]),
    "orc_probe_wave_ab": (i_, [i_, i_, f_, f_, f_, f_, f_, f_, f_, f_, f_, f_, f_, i_, p_, p_]),
    "orc_aberration_chi": (f_, [f_, f_, f_, i_, p_]),
    "orc_aberration_table": (i_, [p_, p_]),
    "orc_detection_planes": (i_, [i_, i_, p_]),
    "orc_rng_create": (p_, [i_]),
    "orc_rng_destroy": (None, [p_]),
    "orc_rng_seed": (None, [p_, i_]),
    "orc_rng_rand": (i_, [p_]),
    "orc_rng_unirand": (d_, [p_, d_, d_]),
    "orc_rng_normrand": (d_, [p_, d_, d_]),
    "orc_rng_raw": (C.c_uint, [p_]),
    "orc_variant_sequence": (i_, [p_, i_, p_, i_, p_, p_]),
    "orc_fdncs2m": (f_, [p_, ll_]),
    "orc_masked_dot": (f_, [p_, p_, p_, ll_]),
    "orc_plain_dot": (f_, [p_, p_, ll_]),
    "orc_offset_wave": (i_, [i_, i_, f_, f_, f_, p_, p_, f_, f_, f_, i_, i_, p_]),
    "orc_multislice": (i_, [i_, i_, i_, i_, p_, p_, p_, i_, p_, p_, i_, p_, p_, i_, i_, p_, p_, p_, i_, p_,
                            p_, i_, f_, p_, p_]),
    "orc_multislice_ch": (i_, [i_, i_, i_, i_, p_, p_, p_, i_, p_, p_, i_, p_, p_, i_, i_, p_, p_, p_, i_, p_,
                               p_, i_, f_, p_, p_, p_, p_, p_, i_, i_]),
    "orc_stem_scan": (i_, [i_, i_, i_, i_, f_, f_, f_, i_, p_, p_, p_, i_, p_, p_, i_, p_, p_, i_, i_, p_, p_,
                           p_, i_, p_, p_, i_, p_, p_, p_, i_, p_, p_, p_, p_]),
    "orc_perovskite_sites": (i_, [i_, i_, d_, i_, i_, i_, i_, d_, i_, i_, p_, p_]),
    "orc_gaussian_grating": (i_, [i_, i_, d_, d_, i_, p_, p_, d_, d_, p_]),
}


def L() -> C.CDLL:
    global _L
    if _L is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"{ORACLE_PATH} missing: run make -C oracle")
        lib = C.CDLL(ORACLE_PATH)
        for k, (r, a) in _SIGS.items():
            f = getattr(lib, k)
            f.restype = r
            f.argtypes = a
        _L = lib
    return _L


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def fft2(a: np.ndarray, direction: int, prec: int = 32) -> np.ndarray:
    d = np.array(a, dtype=np.complex64 if prec == 32 else np.complex128, order="C", copy=True)
    ny, nx = d.shape[-2:]
    f = L().orc_fft2d_f32 if prec == 32 else L().orc_fft2d_f64
    for blk in d.reshape(-1, ny, nx):
        assert f(_p(blk), nx, ny, direction) == 0
    return d


def fft1(a: np.ndarray, direction: int, prec: int = 32) -> np.ndarray:
    """transform along the last axis"""
    d = np.array(a, dtype=np.complex64 if prec == 32 else np.complex128, order="C", copy=True)
    n = d.shape[-1]
    rows = d.reshape(-1, n)
    # howmany interleaved sequences: transpose to (n, howmany)
    t = np.ascontiguousarray(rows.T)
    f = L().orc_fft1d_f32 if prec == 32 else L().orc_fft1d_f64
    assert f(_p(t), n, t.shape[1], direction) == 0
    return np.ascontiguousarray(t.T).reshape(d.shape)


def wavelength(ht: float) -> float:
    return L().orc_wavelength(ht)


def propagator(nx, ny, a0x, a0y, wl, fthick, otx=0.0, oty=0.0, ntype=1) -> np.ndarray:
    out = np.zeros((ny, nx), np.complex64)
    assert L().orc_propagator(nx, ny, a0x, a0y, wl, fthick, otx, oty, ntype, _p(out)) == 0
    return out


def ring_detector(nx, ny, a0x, a0y, wl, nprotype, beta0, beta1, phi0=0.0, phi1=0.0, theta0x=0.0,
                  theta0y=0.0, profile=None, refpix=0.0, kmom1=0):
    det = np.zeros((ny, nx), np.float32)
    msk = np.zeros(nx * ny, np.int32)
    ml = C.c_int(0)
    prof = None if profile is None else np.ascontiguousarray(profile, np.float32)
    assert L().orc_ring_detector(nx, ny, a0x, a0y, wl, nprotype, beta0, beta1, phi0, phi1, theta0x, theta0y,
                                 _p(prof), 0 if prof is None else len(prof), refpix, _p(det), C.byref(ml),
                                 _p(msk), kmom1) == 0
    return det, msk[: ml.value].copy()


def probe_wave(nx, ny, ax, ay, wl, alpha, alpha_rs=0.03, alpha_asym=0.0, alpha_adir=0.0, alpha_x0=0.0,
               alpha_y0=0.0, btx=0.0, bty=0.0, abrr=None):
    """Probe in Fourier space (JProbeGen.cpp:707-777); abrr: [k][x, y] aberration
    coefficients [nm] in the order of aberration_table()."""
    out = np.zeros((ny, nx), np.complex64)
    ab = None if abrr is None else np.ascontiguousarray(abrr, np.float32).reshape(-1)
    nab = 0 if ab is None else len(ab) // 2
    assert L().orc_probe_wave_ab(nx, ny, ax, ay, wl, alpha, alpha_rs, alpha_asym, alpha_adir, alpha_x0,
                                 alpha_y0, btx, bty, nab, _p(ab), _p(out)) == 0
    return out


def aberration_chi(qx, qy, wl, abrr):
    ab = np.ascontiguousarray(abrr, np.float32).reshape(-1)
    return L().orc_aberration_chi(qx, qy, wl, len(ab) // 2, _p(ab))


def aberration_table():
    """[(order m, rotational symmetry n)] of the aberration interface list."""
    num = C.c_int(0)
    L().orc_aberration_table(C.byref(num), None)
    w = np.zeros(2 * num.value, np.int32)
    L().orc_aberration_table(None, _p(w))
    return [tuple(x) for x in w.reshape(-1, 2)]


def detection_planes(ndetper, nobjslc):
    h = np.zeros(max(0, nobjslc) + 1, np.int32)
    n = L().orc_detection_planes(ndetper, nobjslc, _p(h))
    return h, n


class Rng:
    def __init__(self, seed: int):
        self.h = L().orc_rng_create(seed)

    def __del__(self):
        try:
            L().orc_rng_destroy(self.h)
        except Exception:
            pass

    def rand(self):
        return L().orc_rng_rand(self.h)

    def raw(self):
        return L().orc_rng_raw(self.h)

    def unirand(self, a=0.0, b=1.0):
        return L().orc_rng_unirand(self.h, a, b)

    def normrand(self, mu=0.0, s=1.0):
        return L().orc_rng_normrand(self.h, mu, s)


def variant_sequences(seed, nvarslc, objslc, nseq):
    rng = Rng(seed)
    nv = np.ascontiguousarray(nvarslc, np.int32)
    ob = np.ascontiguousarray(objslc, np.int32)
    out = np.zeros((nseq, len(ob)), np.int32)
    for k in range(nseq):
        row = out[k]
        assert L().orc_variant_sequence(rng.h, len(nv), _p(nv), len(ob), _p(ob), _p(row)) == 0
    return out


def fdncs2m(a):
    b = np.ascontiguousarray(a, np.float32).copy()
    return L().orc_fdncs2m(_p(b), len(b))


def masked_dot(mask, in1, in2):
    m = np.ascontiguousarray(mask, np.int32)
    return L().orc_masked_dot(_p(m), _p(np.ascontiguousarray(in1, np.float32)),
                              _p(np.ascontiguousarray(in2, np.float32)), len(m))


def offset_wave(wav0, a0x, a0y, wl, dx, dy, dz=0.0, cppl=None, ibtx=0, ibty=0):
    ny, nx = wav0.shape
    w0 = np.ascontiguousarray(wav0, np.complex64)
    out = np.zeros_like(w0)
    cp = None if cppl is None else np.ascontiguousarray(cppl, np.float32)
    assert L().orc_offset_wave(nx, ny, a0x, a0y, wl, _p(w0), _p(cp), dx, dy, dz, ibtx, ibty, _p(out)) == 0
    return out


class _Packed:
    """A Setup flattened into the oracle's contiguous layout."""

    def __init__(self, sp):
        self.nv = np.ascontiguousarray(sp.nvarslc, np.int32)
        n = sp.nx * sp.ny
        offs, off = [], 0
        for g in sp.gratings:
            offs.append(off)
            off += g.shape[0] * n
        self.slcoff = np.ascontiguousarray(offs, np.int64)
        self.pgr = np.ascontiguousarray(np.concatenate([g.reshape(-1) for g in sp.gratings]), np.complex64)
        self.pro = np.ascontiguousarray(np.stack(sp.propagators), np.complex64)
        self.proidx = np.ascontiguousarray(sp.proidx, np.int32)
        self.objslc = np.ascontiguousarray(sp.objslc, np.int32)
        self.det_objslc = np.ascontiguousarray(sp.det_objslc, np.int32)
        self.det = np.ascontiguousarray(np.stack([d for d, _ in sp.detectors]), np.float32) if sp.detectors \
            else np.zeros((1, n), np.float32)
        self.msk = np.zeros((max(1, len(sp.detectors)), n), np.int32)
        self.msklen = np.zeros(max(1, len(sp.detectors)), np.int32)
        for i, (_, m) in enumerate(sp.detectors):
            if m is not None:
                self.msk[i, : len(m)] = m
                self.msklen[i] = len(m)


def multislice(sp, wave_in, var, weight=1.0, prec=32, islc0=0, want_wave=False, out_int=None):
    """One CPUMultislice pass (JMultiSlice.cpp:3589-3720) of wave_in (Fourier
    space, already offset) with variant sequence var.  Returns (integrated
    readouts [ndetslc, ndet], Fourier-space wave channel or None)."""
    P = _Packed(sp)
    n = sp.nx * sp.ny
    nd = len(sp.detectors)
    oi = np.zeros((sp.ndetslc, max(1, nd)), np.float32) if out_int is None else out_int
    flags = 1 | (16 if want_wave else 0)
    ow = None
    if want_wave:
        ow = np.zeros((sp.ndetslc, sp.ny, sp.nx), np.complex64 if prec == 32 else np.complex128)
    w = np.ascontiguousarray(wave_in, np.complex64)
    v = np.ascontiguousarray(var if len(var) else [0], np.int32)
    rc = L().orc_multislice(prec, sp.nx, sp.ny, len(sp.gratings), _p(P.nv), _p(P.slcoff), _p(P.pgr),
                            len(sp.propagators), _p(P.proidx), _p(P.pro), len(sp.objslc), _p(P.objslc),
                            _p(P.det_objslc), sp.ndetslc, nd, _p(P.det), _p(P.msk), _p(P.msklen), flags,
                            _p(w), _p(v), islc0, weight, _p(oi), _p(ow))
    assert rc == 0
    return oi[:, :nd], ow


def multislice_channels(sp, wave_in, var, weight=1.0, prec=32, islc0=0, flags=1 | 4 | 2 | 8 | 16,
                        chans=None, descan=(0, 0)):
    """One CPUMultislice pass with the channels selected by `flags`
    (1 integrated, 2 image, 4 diffraction, 8 real-space wave, 16 Fourier-space
    wave; ReadoutDifDet_h / ReadoutImgDet_h, JMultiSlice.cpp:3369-3538).
    Accumulates into `chans` (a dict from a previous call) when given;
    descan: (dnx, dny) pixel de-scan of the Fourier-space readouts.
    Returns dict: int [ndetslc, ndet], dif/img [ndetslc, ny, nx],
    wfr/wff [ndetslc, ny, nx] complex."""
    P = _Packed(sp)
    nd = len(sp.detectors)
    rt = np.float32 if prec == 32 else np.float64
    ct = np.complex64 if prec == 32 else np.complex128
    shp = (sp.ndetslc, sp.ny, sp.nx)
    if chans is None:
        chans = {"int": np.zeros((sp.ndetslc, max(1, nd)), np.float32),
                 "dif": np.zeros(shp, rt) if flags & 4 else None,
                 "img": np.zeros(shp, rt) if flags & 2 else None,
                 "wfr": np.zeros(shp, ct) if flags & 8 else None,
                 "wff": np.zeros(shp, ct) if flags & 16 else None}
    w = np.ascontiguousarray(wave_in, np.complex64)
    v = np.ascontiguousarray(var if len(var) else [0], np.int32)
    rc = L().orc_multislice_ch(prec, sp.nx, sp.ny, len(sp.gratings), _p(P.nv), _p(P.slcoff), _p(P.pgr),
                               len(sp.propagators), _p(P.proidx), _p(P.pro), len(sp.objslc), _p(P.objslc),
                               _p(P.det_objslc), sp.ndetslc, nd, _p(P.det), _p(P.msk), _p(P.msklen), flags,
                               _p(w), _p(v), islc0, weight, _p(chans["int"]), _p(chans["dif"]),
                               _p(chans["img"]), _p(chans["wfr"]), _p(chans["wff"]), int(descan[0]),
                               int(descan[1]))
    assert rc == 0
    return chans


def stem_scan(sp, dx, dy, weights, var_all, prec=32, nthreads=1, dz=None, want_wave=False):
    """Threaded scan as the reference bench drives CPUMultislice
    (test/JMSBench1/multislice.cpp:469-688).  Returns [npos, ndetslc, ndet]."""
    P = _Packed(sp)
    dx = np.ascontiguousarray(dx, np.float32)
    dy = np.ascontiguousarray(dy, np.float32)
    dzz = None if dz is None else np.ascontiguousarray(dz, np.float32)
    w = np.ascontiguousarray(weights, np.float32)
    v = np.ascontiguousarray(var_all, np.int32)
    npos, nd = len(dx), len(sp.detectors)
    out = np.zeros((npos, sp.ndetslc, nd), np.float32)
    ow = None
    if want_wave:
        ow = np.zeros((npos, sp.ndetslc, sp.ny, sp.nx), np.complex64 if prec == 32 else np.complex128)
    probe = np.ascontiguousarray(sp.probe, np.complex64)
    flags = 1 | (16 if want_wave else 0)
    rc = L().orc_stem_scan(prec, nthreads, sp.nx, sp.ny, sp.a0[0], sp.a0[1], sp.wl, len(sp.gratings), _p(P.nv),
                           _p(P.slcoff), _p(P.pgr), len(sp.propagators), _p(P.proidx), _p(P.pro),
                           len(sp.objslc), _p(P.objslc), _p(P.det_objslc), sp.ndetslc, nd, _p(P.det),
                           _p(P.msk), _p(P.msklen), flags, _p(probe), None, npos, _p(dx), _p(dy), _p(dzz),
                           len(w), _p(w), _p(v), _p(out), _p(ow))
    assert rc == 0
    return (out, ow) if want_wave else out


def perovskite_sites(ncell, slice_type, rms, seed, displace, a=3.905):
    ncs = ncell * ncell * (2 if slice_type == 0 else 3)
    xy = np.zeros((ncs, 2), np.float64)
    z = np.zeros(ncs, np.int32)
    k = L().orc_perovskite_sites(ncell, ncell, a, slice_type, 38, 22, 8, rms, seed, displace, _p(xy), _p(z))
    assert k == ncs
    return xy, z


def gaussian_grating(nx, ny, ax, ay, xy, z, sigma=0.3, sig_int=0.00729):
    out = np.zeros((ny, nx), np.complex64)
    xy = np.ascontiguousarray(xy, np.float64)
    z = np.ascontiguousarray(z, np.int32)
    assert L().orc_gaussian_grating(nx, ny, ax, ay, len(z), _p(xy), _p(z), sigma, sig_int, _p(out)) == 0
    return out
