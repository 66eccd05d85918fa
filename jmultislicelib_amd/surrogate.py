"""Detector-intensity surrogate with multislice fallback.

A small convolutional network (include/jms_nn.h: four 3x3 convolutions with
32 channels, a 512 -> 64 -> 3 dense head; bf16 activations and compute
weights, FP32 accumulation, trained with Adam by the library's own HIP
kernels) maps a window of projected potential around a probe position to
the three annular-detector intensities.  It has no counterpart in the
reference.  Training data are intensities the multislice kernels compute on
random perovskite-like lattices; at scan time an ensemble of networks
predicts every position, and the plain multislice re-simulates a random
audit subset plus every position where the ensemble disagrees.

Input window: 8 depth bins x 32 x 32 pixels at 0.1 A around the probe; bin b
holds the projected potential of the slices in the b-th eighth of the stack
(each slice type weighted by how many of its slices fall in the bin),
scaled by 1e-3 / (V A).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import configs
from ._lib import lib, ptr
from .jms import Engine, JMultiSlice

NBIN, WIN, NOUT = 8, 32, 3
STEP_ANG = 0.1
WIN_SCALE = 1e-3
DET_NAMES = ("BF 0-11 mrad", "ABF 11-22 mrad", "HAADF 70-200 mrad")


class NNError(RuntimeError):
    pass


def _chk(name, rc):
    if rc != 0:
        raise NNError(f"{name} returned error code {rc}")


def bf16_bits(a: np.ndarray) -> np.ndarray:
    """float32 -> bf16 bit patterns (round to nearest even)"""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


class Network:
    """One surrogate network on the current GPU (jms_nn)."""

    def __init__(self, max_batch: int = 256, seed: int = 1):
        h = C.c_void_p()
        _chk("jms_nn_create", lib().jms_nn_create(C.byref(h), max_batch, seed))
        self.h, self.max_batch = h, max_batch

    def close(self):
        if self.h:
            lib().jms_nn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_kernels(self, mfma: bool):
        """convolutions on the matrix cores (default) or the FMA kernels"""
        _chk("jms_nn_set_kernels", lib().jms_nn_set_kernels(self.h, int(mfma)))

    @staticmethod
    def num_params() -> int:
        return lib().jms_nn_num_params()

    def params(self) -> np.ndarray:
        p = np.zeros(self.num_params(), np.float32)
        _chk("jms_nn_get_params", lib().jms_nn_get_params(self.h, ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _chk("jms_nn_set_params", lib().jms_nn_set_params(self.h, ptr(p)))

    def forward(self, x_bits: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x_bits, np.uint16)
        out = np.zeros((x.shape[0], NOUT), np.float32)
        _chk("jms_nn_forward", lib().jms_nn_forward(self.h, ptr(x), x.shape[0], ptr(out)))
        return out

    def gradients(self, x_bits, t, scale):
        x = np.ascontiguousarray(x_bits, np.uint16)
        t = np.ascontiguousarray(t, np.float32)
        sc = np.ascontiguousarray(scale, np.float32)
        g = np.zeros(self.num_params(), np.float32)
        loss = C.c_float()
        _chk("jms_nn_gradients", lib().jms_nn_gradients(self.h, ptr(x), ptr(t), ptr(sc), x.shape[0], ptr(g),
                                                        C.byref(loss)))
        return loss.value, g

    def train_epoch(self, data: "WindowSet", batch: int, lr: float, scale, seed: int) -> float:
        sc = np.ascontiguousarray(scale, np.float32)
        loss = C.c_float()
        _chk("jms_nn_train_epoch", lib().jms_nn_train_epoch(self.h, data.h, batch, lr, ptr(sc), seed,
                                                            C.byref(loss)))
        return loss.value

    def predict(self, data: "WindowSet") -> np.ndarray:
        out = np.zeros((data.n, NOUT), np.float32)
        _chk("jms_nn_predict", lib().jms_nn_predict(self.h, data.h, ptr(out)))
        return out


class WindowSet:
    """Device-resident input windows and targets (jms_nn_data)."""

    def __init__(self, h, n):
        self.h, self.n = h, n

    @classmethod
    def from_arrays(cls, x_bits, t):
        x = np.ascontiguousarray(x_bits, np.uint16)
        t = np.ascontiguousarray(t, np.float32)
        h = C.c_void_p()
        _chk("jms_nn_data_create", lib().jms_nn_data_create(C.byref(h), ptr(x), ptr(t), x.shape[0]))
        return cls(h, x.shape[0])

    @classmethod
    def from_crystals(cls, maps, cells, cid, px_ang, py_ang, binw, t=None, step=STEP_ANG):
        """maps [ncrys][2][m][m] (potential_maps), cells [ncrys][2] [A], per
        sample crystal index cid and probe position [A]; window pixel
        spacing step [A]"""
        maps = np.ascontiguousarray(maps, np.float32)
        cells = np.ascontiguousarray(cells, np.float32)
        cid = np.ascontiguousarray(cid, np.int32)
        px, py = np.ascontiguousarray(px_ang, np.float32), np.ascontiguousarray(py_ang, np.float32)
        bw = np.ascontiguousarray(binw, np.float32)
        tt = None if t is None else np.ascontiguousarray(t, np.float32)
        h = C.c_void_p()
        _chk("jms_nn_data_windows", lib().jms_nn_data_windows(
            C.byref(h), maps.shape[0], maps.shape[1], maps.shape[2], ptr(maps), ptr(cells), len(px), ptr(cid),
            ptr(px), ptr(py), ptr(bw), float(step), WIN_SCALE, ptr(tt)))
        return cls(h, len(px))

    def windows(self) -> np.ndarray:
        x = np.zeros((self.n, NBIN, WIN, WIN), np.uint16)
        _chk("jms_nn_data_get", lib().jms_nn_data_get(self.h, ptr(x)))
        return x

    def close(self):
        if self.h:
            lib().jms_nn_data_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# crystals
# ---------------------------------------------------------------------------
@dataclass
class Lattice:
    """Perovskite-like cubic cell: A (corners), B (body centre), O (faces)."""
    a: float        # [A]
    za: int
    zb: int
    zo: int = 8


def random_lattice(rng: np.random.Generator) -> Lattice:
    """a uniform in [3.6, 4.2] A, A-site Z uniform in [20, 82], B-site Z
    uniform in [22, 74], O."""
    return Lattice(a=float(rng.uniform(3.6, 4.2)), za=int(rng.integers(20, 83)), zb=int(rng.integers(22, 75)))


def cell_sites(lat: Lattice, ncell: int, slice_type: int):
    """Undisplaced sites (xy [A], Z) of one slice type over ncell x ncell cells."""
    ncs = ncell * ncell * (2 if slice_type == 0 else 3)
    xy = np.zeros((ncs, 2), np.float64)
    z = np.zeros(ncs, np.int32)
    k = lib().jms_perovskite_sites(ncell, ncell, lat.a, slice_type, lat.za, lat.zb, lat.zo, 0.0, 0, 0, ptr(xy),
                                   ptr(z))
    assert k == ncs
    return xy, z


def lattice_setup(lat: Lattice, n: int = 512, ncell: int = 8, nslices: int = 200, ht: float = 200.0,
                  alpha: float = 25.0, alpha_rs: float = 0.03, det_ranges=configs.DETECTOR_RANGES):
    """Set-up of a static lattice (the configs' recipe: Gaussian potential, two
    slice types of a/2 alternating, 25 mrad probe, three annular detectors);
    the gratings are generated on the device."""
    cell = ncell * lat.a
    a_nm, dz_nm = cell * 0.1, 0.5 * lat.a * 0.1
    J = JMultiSlice()
    J.SetHighTension(ht)
    J.SetGridSize(n, n)
    J.SetSupercellSize(a_nm, a_nm, lat.a * 0.1)
    wl = J.GetWaveLength()
    gsites = [[cell_sites(lat, ncell, t)] for t in (0, 1)]
    pro = J.CalculatePropagator(dz_nm, 0.0, 0.0, 1)
    dets = [J.CalculateRingDetector(b0, b1) for (b0, b1) in det_ranges]
    probe = J.CalculateProbeWaveFourier(alpha, alpha_rs)
    objslc = np.array([i % 2 for i in range(nslices)], np.int32)
    det_objslc, ndetslc = configs.detection_planes(0, nslices)
    return configs.Setup(nx=n, ny=n, a0=(a_nm, a_nm, lat.a * 0.1), ht=ht, wl=wl, nvarslc=[1, 1], gratings=None,
                         propagators=[pro], proidx=[0, 0], objslc=objslc, det_objslc=det_objslc,
                         ndetslc=ndetslc, detectors=dets, probe=probe, flags=1, slice_thickness=dz_nm,
                         grating_sites=gsites)


MAP_M = 40  # potential-map pixels per cell edge (0.09-0.105 A for a in [3.6, 4.2] A)


def potential_maps(lat: Lattice, m: int = MAP_M) -> np.ndarray:
    """Projected potential [V A] of slice types 0 and 1 over one cell on an
    m x m grid (pixel (i, j) at (j a/m, i a/m)), Gaussian model, periodic."""
    g = np.arange(m) * lat.a / m
    X, Y = np.meshgrid(g, g)
    maps = np.zeros((2, m, m))
    s2 = 2.0 * configs.SIGMA_POT ** 2
    for t in (0, 1):
        xy, z = cell_sites(lat, 1, t)
        for (x0, y0), zz in zip(xy, z):
            amp = 2.0 * zz ** 0.7 / (np.pi * s2)
            for ix in (-1, 0, 1):
                for iy in (-1, 0, 1):
                    dx, dy = X - (x0 + ix * lat.a), Y - (y0 + iy * lat.a)
                    maps[t] += amp * np.exp(-(dx * dx + dy * dy) / s2)
    return maps.astype(np.float32)


def bin_weights(nslices: int = 200) -> np.ndarray:
    """[8][2]: number of type-0 and type-1 slices in each eighth of the stack"""
    w = np.zeros((NBIN, 2), np.float32)
    for j in range(nslices):
        w[min(NBIN - 1, j * NBIN // nslices), j % 2] += 1
    return w


def plain_scan(eng: Engine, px_ang, py_ang) -> np.ndarray:
    """Exit-plane intensities [npos][3] of the plain batched multislice."""
    sp = eng.sp
    var = np.zeros((len(px_ang), len(sp.objslc)), np.int32)
    eng.stage_scan(np.asarray(px_ang, np.float32) * 0.1, np.asarray(py_ang, np.float32) * 0.1, None,
                   np.ones(1, np.float32), var)
    eng.run_staged(0, len(px_ang))
    eng.sync()
    return eng.scan_results()[:, -1, :]


# the point group 4mm of the perovskite-like cells about the A column at the
# origin: (x, y) -> (m00 x + m01 y, m10 x + m11 y).  Each slice type, and so
# the whole stack, maps onto itself, and so does the round aberration-free
# probe: the detector intensities at the 8 images of a position are equal.
SYM_4MM = ((1, 0, 0, 1), (0, -1, 1, 0), (-1, 0, 0, -1), (0, 1, -1, 0),
           (1, 0, 0, -1), (-1, 0, 0, 1), (0, 1, 1, 0), (0, -1, -1, 0))


def symmetric_images(px, py, a):
    """The 8 images [8][n] of positions (px, py) [A] under 4mm about the
    origin, wrapped into the cell [0, a)."""
    px, py = np.asarray(px, np.float64), np.asarray(py, np.float64)
    xs = np.stack([np.mod(m00 * px + m01 * py, a) for m00, m01, _, _ in SYM_4MM])
    ys = np.stack([np.mod(m10 * px + m11 * py, a) for _, _, m10, m11 in SYM_4MM])
    return xs.astype(np.float32), ys.astype(np.float32)


@dataclass
class Samples:
    """Positions on several crystals with their multislice intensities."""
    lattices: list
    cid: np.ndarray     # crystal of each sample
    px: np.ndarray      # [A]
    py: np.ndarray
    t: np.ndarray       # [n][3]

    def augmented(self) -> "Samples":
        """The samples at all 8 symmetry images of their positions, with the
        same intensities (8x the training data without new simulations)."""
        xs, ys = [], []
        for k, lat in enumerate(self.lattices):
            sel = self.cid == k
            x8, y8 = symmetric_images(self.px[sel], self.py[sel], lat.a)
            xs.append((x8, y8, sel))
        px = np.zeros((8, len(self.px)), np.float32)
        py = np.zeros((8, len(self.py)), np.float32)
        for x8, y8, sel in xs:
            px[:, sel], py[:, sel] = x8, y8
        return Samples(self.lattices, np.tile(self.cid, 8), px.ravel(), py.ravel(), np.tile(self.t, (8, 1)))

    def windows(self, nslices: int = 200, scale=None, step: float = STEP_ANG) -> WindowSet:
        """device windows (pixel spacing step [A]) with targets t, or
        log(t / scale) per detector when scale is given"""
        maps = np.stack([potential_maps(lat) for lat in self.lattices])
        cells = np.array([[lat.a, lat.a] for lat in self.lattices], np.float32)
        t = self.t if scale is None else np.log(np.maximum(self.t, 1e-30) / np.asarray(scale, np.float32))
        return WindowSet.from_crystals(maps, cells, self.cid, self.px, self.py, bin_weights(nslices), t, step)


def simulate(lattices, npos: int, seed: int, eng: Engine = None, **kw) -> Samples:
    """npos random positions per lattice (uniform over one cell) and their
    exit-plane intensities from the plain multislice."""
    rng = np.random.default_rng(seed)
    own = eng is None
    eng = eng or Engine()
    px, py, t, cid = [], [], [], []
    for k, lat in enumerate(lattices):
        eng.setup(lattice_setup(lat, **kw))
        eng.set_batch(npos)
        x = rng.uniform(0.0, lat.a, npos).astype(np.float32)
        y = rng.uniform(0.0, lat.a, npos).astype(np.float32)
        t.append(plain_scan(eng, x, y))
        px.append(x)
        py.append(y)
        cid.append(np.full(npos, k, np.int32))
    if own:
        eng.close()
    return Samples(list(lattices), np.concatenate(cid), np.concatenate(px), np.concatenate(py), np.concatenate(t))


class Ensemble:
    """Networks trained on the same data from different initial weights, on
    windows of pixel spacing step [A] (training and scans use the same)."""

    def __init__(self, nmembers: int = 4, max_batch: int = 256, seed: int = 1, step: float = STEP_ANG):
        self.nets = [Network(max_batch, seed + i) for i in range(nmembers)]
        self.scale = np.ones(NOUT, np.float32)
        self.step = float(step)

    def train(self, samples: Samples, epochs: int, batch: int = 256, lr: float = 1e-3, seed: int = 0,
              nslices: int = 200, log=None, augment: bool = False):
        """Adam on the mean squared error of log(intensity / detector mean):
        the three detectors differ by orders of magnitude and the loss then
        measures relative errors; the learning rate decays linearly to lr/10.
        augment: train on the 8 symmetry images of every position.
        Returns the per-epoch losses [epochs][members]."""
        self.scale = np.asarray(samples.t, np.float64).mean(axis=0).astype(np.float32)
        if augment:
            samples = samples.augmented()
        data = samples.windows(nslices, self.scale, self.step)
        one = np.ones(NOUT, np.float32)
        hist = []
        for ep in range(epochs):
            lre = lr * (1.0 - 0.9 * ep / max(1, epochs - 1))
            ls = [net.train_epoch(data, batch, lre, one, seed + 1000 * ep + i) for i, net in enumerate(self.nets)]
            hist.append(ls)
            if log:
                log(f"epoch {ep}: loss " + " ".join(f"{x:.3e}" for x in ls))
        data.close()
        return np.array(hist)

    def predict(self, data: WindowSet, nsym: int = 1):
        """(mean [n][3], relative spread [n][3] = std / |mean|) in intensity
        units.  nsym = 8: data holds the 8 symmetry images of n positions
        (image-major); each member's log-predictions are averaged over the
        images before the ensemble statistics."""
        lp = np.stack([net.predict(data) for net in self.nets]).astype(np.float64)
        if nsym > 1:
            lp = lp.reshape(len(self.nets), nsym, -1, NOUT).mean(axis=1)
        p = np.exp(lp) * self.scale
        mean = p.mean(axis=0)
        spread = p.std(axis=0) / np.maximum(np.abs(mean), 1e-30)
        return mean, spread

    def close(self):
        for n in self.nets:
            n.close()


def _cell_basis(u, v, kmax: int) -> np.ndarray:
    """Periodic functions of the fractional position (u, v) in the cell that
    are invariant under 4mm about the origin: cos(2 pi h u) cos(2 pi k v) +
    cos(2 pi k u) cos(2 pi h v), 0 <= h <= k <= kmax.  [n][nbasis]"""
    u, v = np.asarray(u, np.float64), np.asarray(v, np.float64)
    cols = []
    for k in range(kmax + 1):
        for h in range(k + 1):
            cols.append(np.cos(2 * np.pi * h * u) * np.cos(2 * np.pi * k * v) +
                        np.cos(2 * np.pi * k * u) * np.cos(2 * np.pi * h * v))
    return np.stack(cols, axis=1)


def calibrate(pred, truth, u_fit, v_fit, u_all, v_all, kmax: int = 6, ridge: float = 1e-6):
    """Correction factors [n_all][3] for predictions at (u_all, v_all) from
    the re-simulated positions (u_fit, v_fit) where both the prediction and
    the multislice value are known: log(truth / pred) is fitted per detector
    by least squares (small ridge) on the 4mm-symmetric periodic basis of the
    cell, which is where a surrogate's error on one crystal varies (the
    network's error is a smooth function of the probe position in the cell)."""
    A = _cell_basis(u_fit, v_fit, kmax)
    r = np.log(np.maximum(truth, 1e-30) / np.maximum(pred, 1e-30))
    G = A.T @ A + ridge * len(A) * np.eye(A.shape[1])
    coef = np.linalg.solve(G, A.T @ r)
    return np.exp(_cell_basis(u_all, v_all, kmax) @ coef)


def surrogate_scan(ens: Ensemble, eng: Engine, lat: Lattice, px_ang, py_ang, audit: float = 0.05,
                   spread_max: float = 0.02, seed: int = 0, nslices: int = 200, dets=(1, 2), sym: bool = False,
                   calib: int = 0):
    """Intensities [npos][3] of a scan on lat (engine set up for it): the
    ensemble predicts every position (sym: averaged over the 8 symmetry
    images of the position); a random fraction `audit` and every position
    whose ensemble spread exceeds spread_max on the detectors `dets` are
    re-simulated by the plain multislice, which replaces the prediction
    there.  calib > 0: the re-simulated positions also calibrate the other
    predictions (`calibrate` with kmax = calib).  Returns (intensities,
    indices re-simulated, timings)."""
    t0 = time.perf_counter()
    n = len(px_ang)
    px, py = np.asarray(px_ang, np.float32), np.asarray(py_ang, np.float32)
    nsym = 8 if sym else 1
    if sym:
        x8, y8 = symmetric_images(px, py, lat.a)
        px, py = x8.ravel(), y8.ravel()
    ws = Samples([lat], np.zeros(n * nsym, np.int32), px, py,
                 np.zeros((n * nsym, NOUT), np.float32)).windows(nslices, step=ens.step)
    t1 = time.perf_counter()
    mean, spread = ens.predict(ws, nsym)
    ws.close()
    t2 = time.perf_counter()
    rng = np.random.default_rng(seed)
    pick = rng.random(n) < audit
    pick |= (spread[:, list(dets)] > spread_max).any(axis=1)
    idx = np.nonzero(pick)[0]
    out = mean.copy()
    if len(idx):
        eng.set_batch(len(idx))
        out[idx] = plain_scan(eng, np.asarray(px_ang)[idx], np.asarray(py_ang)[idx])
    t3 = time.perf_counter()
    raw = out.copy()
    if calib > 0 and len(idx) > 0:
        u = np.mod(np.asarray(px_ang, np.float64) / lat.a, 1.0)
        v = np.mod(np.asarray(py_ang, np.float64) / lat.a, 1.0)
        rest = np.setdiff1d(np.arange(n), idx)
        fac = calibrate(mean[idx], out[idx], u[idx], v[idx], u[rest], v[rest], kmax=calib)
        out[rest] = mean[rest] * fac
    t4 = time.perf_counter()
    return out, idx, {"windows_s": t1 - t0, "inference_s": t2 - t1, "fallback_s": t3 - t2,
                      "calibration_s": t4 - t3, "total_s": t4 - t0, "uncalibrated": raw}
