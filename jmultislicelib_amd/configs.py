"""Benchmark workloads: synthetic perovskite specimens and STEM scans.

``perovskite_setup`` builds everything the multislice needs for the cubic
ABO3 recipe of the benchmark configurations (A = Sr Z38 at the cell corners,
B = Ti Z22 at the body centre, O Z8 at the face centres; two slices of a/2
per cell; Gaussian projected potentials; Fresnel propagator; ring
detectors; aberration-free probe) with the library's own host calculators
(CalculatePropagator, CalculateRingDetector, CalculateProbeWaveFourier via
the C-ABI) and specimen builder.  Lengths are in nm at the library interface
(the reference's unit), in Angstrom in the recipe.

Named workloads:
  * single_probe_256: 256x256 grid, 4x4 cells (15.62 A), 100 slices, one
    variant per slice type, one probe on an A column;
  * stem_scan_512: 512x512 grid, 8x8 cells (31.24 A), 8 frozen-phonon
    variants per slice type (rms 0.08 A, seeds 1000 + 8*type + v), 200 slices,
    64x64 probe positions over 2x2 cells, 8 passes per position (weights 1/8),
    variant sequences from one generator seeded 42.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import lib, ptr
from .jms import JMultiSlice

A_PEROVSKITE = 3.905          # lattice constant [A]
SIGMA_POT = 0.3               # Gaussian width of the projected potential [A]
SIG_INT = 0.00729             # interaction constant [1/(V A)] used by the recipe
Z_A, Z_B, Z_O = 38, 22, 8
DETECTOR_RANGES = ((0.0, 11.0), (11.0, 22.0), (70.0, 200.0))  # mrad


@dataclass
class Setup:
    nx: int
    ny: int
    a0: tuple                     # super-cell size [nm] (x, y, z)
    ht: float                     # [keV]
    wl: float                     # [nm]
    nvarslc: list
    gratings: list                # per structure slice: (nvar, ny, nx) complex64
    propagators: list             # (ny, nx) complex64
    proidx: list
    objslc: np.ndarray
    det_objslc: np.ndarray
    ndetslc: int
    detectors: list               # (det (ny, nx) float32, mask int32 or None)
    probe: np.ndarray             # (ny, nx) complex64, Fourier space
    flags: int = 1                # detection channels
    slice_thickness: float = 0.0  # [nm]
    sites: list = field(default_factory=list)  # per (type, variant): (xy [A], z)
    # per structure slice, per variant: (xy [A], z) atom sets whose gratings
    # are generated on the device when `gratings` is None
    grating_sites: list = None


def detection_planes(ndetper: int, nobjslc: int) -> tuple[np.ndarray, int]:
    """Plane -> detection index hash (CJMultiSlice::SetDetectionPlanes,
    JMultiSlice.cpp:579-627)."""
    nos = max(0, nobjslc)
    h = np.full(nos + 1, -1, np.int32)
    if ndetper <= 0:
        h[nos] = 0
        return h, 1
    if nos == 0:
        h[0] = 0
        return h, 1
    k = 0
    for i in range(0, nos + 1, ndetper):
        h[i] = k
        k += 1
    if nos % ndetper:
        h[nos] = k
        k += 1
    return h, k


def perovskite_sites(ncell: int, slice_type: int, nvar: int, v: int, rms: float, seed_base: int):
    """Sites (xy [A], z) of one slice type and frozen-phonon variant."""
    ncs = ncell * ncell * (2 if slice_type == 0 else 3)
    xy = np.zeros((ncs, 2), np.float64)
    z = np.zeros(ncs, np.int32)
    seed = seed_base + 8 * slice_type + v
    k = lib().jms_perovskite_sites(ncell, ncell, A_PEROVSKITE, slice_type, Z_A, Z_B, Z_O, rms, seed,
                                   1 if nvar > 1 else 0, ptr(xy), ptr(z))
    assert k == ncs
    return xy, z


def gaussian_grating(n: int, cell_ang: float, xy: np.ndarray, z: np.ndarray) -> np.ndarray:
    out = np.zeros((n, n), np.complex64)
    rc = lib().jms_gaussian_phase_grating(n, n, cell_ang, cell_ang, len(z), ptr(xy), ptr(z), SIGMA_POT,
                                          SIG_INT, ptr(out))
    assert rc == 0
    return out


def perovskite_setup(n: int = 512, ncell: int = 8, nslices: int = 200, nvar: int = 8, rms: float = 0.08,
                     seed_base: int = 1000, ht: float = 200.0, alpha: float = 25.0,
                     alpha_rs: float = 0.03, ndetper: int = 0, ptype: int = 1,
                     det_ranges=DETECTOR_RANGES, flags: int = 1) -> Setup:
    cell_ang = ncell * A_PEROVSKITE
    a_nm = cell_ang * 0.1
    dz_nm = 0.5 * A_PEROVSKITE * 0.1
    J = JMultiSlice()
    J.SetHighTension(ht)
    J.SetGridSize(n, n)
    J.SetSupercellSize(a_nm, a_nm, A_PEROVSKITE * 0.1)
    wl = J.GetWaveLength()
    gratings, sites = [], []
    for t in (0, 1):
        g = np.zeros((nvar, n, n), np.complex64)
        for v in range(nvar):
            xy, z = perovskite_sites(ncell, t, nvar, v, rms, seed_base)
            sites.append((xy, z))
            g[v] = gaussian_grating(n, cell_ang, xy, z)
        gratings.append(g)
    pro = J.CalculatePropagator(dz_nm, 0.0, 0.0, ptype)  # also selects the detector angle mapping
    dets = [J.CalculateRingDetector(b0, b1) for (b0, b1) in det_ranges]
    probe = J.CalculateProbeWaveFourier(alpha, alpha_rs)
    objslc = np.array([i % 2 for i in range(nslices)], np.int32)
    det_objslc, ndetslc = detection_planes(ndetper, nslices)
    return Setup(nx=n, ny=n, a0=(a_nm, a_nm, A_PEROVSKITE * 0.1), ht=ht, wl=wl, nvarslc=[nvar, nvar],
                 gratings=gratings, propagators=[pro], proidx=[0, 0], objslc=objslc, det_objslc=det_objslc,
                 ndetslc=ndetslc, detectors=dets, probe=probe, flags=flags, slice_thickness=dz_nm,
                 sites=sites)


def single_probe_256(**kw) -> Setup:
    """256^2, 4x4 cells, 100 slices, no frozen phonons."""
    return perovskite_setup(n=256, ncell=4, nslices=100, nvar=1, **kw)


def stem_scan_512(**kw) -> Setup:
    """512^2, 8x8 cells, 200 slices, 8 frozen-phonon variants per slice type."""
    return perovskite_setup(n=512, ncell=8, nslices=200, nvar=8, **kw)


def thick_amorphous(n: int = 2048, cell_ang: float = 125.0, nslices: int = 1000, dz_ang: float = 2.0,
                    density: float = 0.06, zs=(6, 8, 14), pz=(0.5, 0.3, 0.2), seed: int = 3,
                    ht: float = 200.0, alpha: float = 25.0, alpha_rs: float = 0.03,
                    det_ranges=DETECTOR_RANGES, flags: int = 1) -> Setup:
    """Large thick specimen: amorphous random atoms (density [1/A^3], Z drawn
    from zs with probabilities pz, uniform positions) in `nslices` distinct
    slices of dz_ang over a cell_ang^2 cell, the Gaussian potential model;
    the gratings are generated on the device (grating_sites)."""
    rng = np.random.default_rng(seed)
    a_nm = cell_ang * 0.1
    dz_nm = dz_ang * 0.1
    J = JMultiSlice()
    J.SetHighTension(ht)
    J.SetGridSize(n, n)
    J.SetSupercellSize(a_nm, a_nm, dz_nm)
    wl = J.GetWaveLength()
    nat = int(round(density * cell_ang * cell_ang * dz_ang))
    gsites = []
    for _ in range(nslices):
        xy = rng.uniform(0.0, cell_ang, (nat, 2))
        z = rng.choice(np.asarray(zs), nat, p=np.asarray(pz)).astype(np.int32)
        gsites.append([(xy, z)])
    pro = J.CalculatePropagator(dz_nm, 0.0, 0.0, 1)
    dets = [J.CalculateRingDetector(b0, b1) for (b0, b1) in det_ranges]
    probe = J.CalculateProbeWaveFourier(alpha, alpha_rs)
    objslc = np.arange(nslices, dtype=np.int32)
    det_objslc, ndetslc = detection_planes(0, nslices)
    return Setup(nx=n, ny=n, a0=(a_nm, a_nm, dz_nm), ht=ht, wl=wl, nvarslc=[1] * nslices, gratings=None,
                 propagators=[pro], proidx=[0] * nslices, objslc=objslc, det_objslc=det_objslc,
                 ndetslc=ndetslc, detectors=dets, probe=probe, flags=flags, slice_thickness=dz_nm,
                 grating_sites=gsites)


def scan_positions(nscan: int = 64, ncells_scan: int = 2):
    """Probe positions [nm] of an nscan x nscan raster over ncells_scan^2 cells
    (row by row), starting on the A column at the origin."""
    step = ncells_scan * A_PEROVSKITE * 0.1 / nscan
    iy, ix = np.meshgrid(np.arange(nscan), np.arange(nscan), indexing="ij")
    return (ix.ravel() * step).astype(np.float32), (iy.ravel() * step).astype(np.float32)


def variant_sequences(sp: Setup, nseq: int, seed: int = 42) -> np.ndarray:
    """nseq variant sequences from one generator (serial scan order)."""
    nv = np.ascontiguousarray(sp.nvarslc, np.int32)
    out = np.zeros((nseq, len(sp.objslc)), np.int32)
    rc = lib().jms_variant_sequences(seed, len(nv), ptr(nv), len(sp.objslc), ptr(sp.objslc), nseq,
                                     ptr(out))
    assert rc == 0
    return out


def slice_file_name(prefix: str, islc: int, suffix: str = ".sli") -> str:
    """Name of structure slice islc (0-based) in a numbered series:
    prefix_001.sli, prefix_002.sli, ... (params.cpp:110-115,
    prm_sample.cpp:599-615)."""
    return f"{prefix}_{islc + 1:03d}{suffix}"


def setup_from_slice_files(paths, nobj: int, ndetper: int = 0, alpha: float = 25.0, alpha_rs: float = 0.03,
                           tilt_mrad=(0.0, 0.0), ptype: int = 1, det_ranges=DETECTOR_RANGES,
                           flags: int = 1) -> Setup:
    """Set-up from EMS/SLI phase-grating files, one per structure slice, as
    the reference bench prepares it: grid, super-cell and high tension from
    the first file's header, one propagator per structure slice from its own
    header thickness (prm_slice.cpp:495-521, prm_sample.cpp:1231-1282) with
    the tilt given in mrad, periodic stacking of nobj object slices
    (prm_sample.cpp:283-300) and detection planes every ndetper slices."""
    from .io import read_ems
    if not paths:
        raise ValueError("no slice files")
    heads, gratings = [], []
    for p in paths:
        h, d = read_ems(p)
        heads.append(h)
        gratings.append(d)
    h0 = heads[0]
    nx, ny = int(h0.grid_x), int(h0.grid_y)
    for p, h in zip(paths, heads):
        if (h.grid_x, h.grid_y) != (nx, ny):
            raise ValueError(f"{p}: grid {h.grid_x}x{h.grid_y} differs from {nx}x{ny}")
    J = JMultiSlice()
    J.SetHighTension(float(h0.ekv))
    J.SetGridSize(nx, ny)
    J.SetSupercellSize(float(h0.sx), float(h0.sy), float(h0.sz))
    wl = J.GetWaveLength()
    ttf = np.float32(180.0 / np.pi) * np.float32(0.001)  # mrad -> degree
    otx, oty = float(ttf * np.float32(tilt_mrad[0])), float(ttf * np.float32(tilt_mrad[1]))
    pros = [J.CalculatePropagator(float(h.sz), otx, oty, ptype) for h in heads]
    dets = [J.CalculateRingDetector(b0, b1) for (b0, b1) in det_ranges]
    probe = J.CalculateProbeWaveFourier(alpha, alpha_rs)
    nslc = len(paths)
    objslc = np.array([i % nslc for i in range(nobj)], np.int32)
    det_objslc, ndetslc = detection_planes(ndetper, nobj)
    return Setup(nx=nx, ny=ny, a0=(float(h0.sx), float(h0.sy), float(h0.sz)), ht=float(h0.ekv), wl=wl,
                 nvarslc=[int(h.num_var) for h in heads], gratings=gratings, propagators=pros,
                 proidx=list(range(nslc)), objslc=objslc, det_objslc=det_objslc, ndetslc=ndetslc,
                 detectors=dets, probe=probe, flags=flags, slice_thickness=float(h0.sz))


@dataclass
class Job:
    """One job of an uneven batch: its set-up and probe positions [nm]."""
    setup: Setup
    dx: np.ndarray
    dy: np.ndarray
    cell_ang: tuple  # lattice cell used (a, b, c) [A]
    species: np.ndarray


JOB_EDGES = (360, 384, 480, 512, 640, 720)


def random_job(rng: np.random.Generator, edge: int = None, nslices: int = None, npos: int = None,
               sampling_ang: float = 0.06, ht: float = 200.0, det_ranges=DETECTOR_RANGES,
               flags: int = 1) -> Job:
    """A random orthorhombic crystal job: grid edge from JOB_EDGES at
    sampling_ang, a, b, c uniform in [3.5, 6.0] A, 1-4 species with Z uniform
    in [6, 79], 2-8 atoms per cell at uniform fractional positions, the
    Gaussian potential model, two slice types of c/2 (fractional z below /
    above 1/2) alternating for nslices in [50, 400], 16-256 probe positions
    on a raster over one cell, a 20-30 mrad aperture.  The square super-cell
    (edge * sampling_ang) holds a whole number of cells along x and y; the
    cell is stretched to fit it."""
    e = int(edge if edge is not None else rng.choice(JOB_EDGES))
    a, b, c = rng.uniform(3.5, 6.0, 3)
    nsp = int(rng.integers(1, 5))
    species = rng.integers(6, 80, nsp).astype(np.int32)
    nat = int(rng.integers(2, 9))
    frac = rng.uniform(0.0, 1.0, (nat, 3))
    kind = rng.integers(0, nsp, nat)
    ns = int(nslices if nslices is not None else rng.integers(50, 401))
    np_ = int(npos if npos is not None else rng.integers(16, 257))
    alpha = float(rng.uniform(20.0, 30.0))
    L = e * sampling_ang
    ncx, ncy = max(1, int(round(L / a))), max(1, int(round(L / b)))
    ax, by = L / ncx, L / ncy
    dz_nm = 0.5 * c * 0.1
    J = JMultiSlice()
    J.SetHighTension(ht)
    J.SetGridSize(e, e)
    J.SetSupercellSize(L * 0.1, L * 0.1, c * 0.1)
    wl = J.GetWaveLength()
    gratings = []
    for t in (0, 1):
        sel = (frac[:, 2] < 0.5) if t == 0 else (frac[:, 2] >= 0.5)
        fx, fy, zs = frac[sel, 0], frac[sel, 1], species[kind[sel]]
        iy, ix = np.meshgrid(np.arange(ncy), np.arange(ncx), indexing="ij")
        ox, oy = ix.ravel() * ax, iy.ravel() * by
        xy = np.stack([(ox[:, None] + fx[None, :] * ax).ravel(), (oy[:, None] + fy[None, :] * by).ravel()], 1)
        z = np.tile(zs, len(ox)).astype(np.int32)
        gratings.append(gaussian_grating(e, L, np.ascontiguousarray(xy, np.float64), z)[None])
    pro = J.CalculatePropagator(dz_nm, 0.0, 0.0, 1)
    dets = [J.CalculateRingDetector(b0, b1) for (b0, b1) in det_ranges]
    probe = J.CalculateProbeWaveFourier(alpha, 0.03)
    objslc = np.array([i % 2 for i in range(ns)], np.int32)
    det_objslc, ndetslc = detection_planes(0, ns)
    sp = Setup(nx=e, ny=e, a0=(L * 0.1, L * 0.1, c * 0.1), ht=ht, wl=wl, nvarslc=[1, 1], gratings=gratings,
               propagators=[pro], proidx=[0, 0], objslc=objslc, det_objslc=det_objslc, ndetslc=ndetslc,
               detectors=dets, probe=probe, flags=flags, slice_thickness=dz_nm)
    side = int(np.ceil(np.sqrt(np_)))
    i = np.arange(np_)
    dx = ((i % side) * (ax * 0.1 / side)).astype(np.float32)
    dy = ((i // side) * (by * 0.1 / side)).astype(np.float32)
    return Job(setup=sp, dx=dx, dy=dy, cell_ang=(ax, by, c), species=species)


def uneven_jobs(njobs: int = 64, seed: int = 2, **kw) -> list:
    """The uneven batch: njobs random crystal jobs from one seeded generator."""
    rng = np.random.default_rng(seed)
    return [random_job(rng, **kw) for _ in range(njobs)]
