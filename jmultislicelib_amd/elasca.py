"""From a structure to phase gratings: the host-side scattering factors and
slicing of the library (include/jms_elasca.h) and the device grating
generator, in the order the reference's bench drives them
(test/JMSBench1/prm_sample.cpp:728-790, 1290-1420):

  structure --(equidistant z-slices, CStructure::GetZSliceStructure)-->
  slices --(atom types: Weickenmeier-Kohl radial form factors with
  Debye-Waller and absorptive parts, CAtomType::CalculateRadialFormFactors;
  sampling of CJEElaSca::CalculateFormFactors)--> radial tables --(2-D
  projection on the grid, CJEElaSca::PrepareScatteringFunctions)--> 2-D
  form factors --(jms_hip_formfactor_gratings: structure factor sum, inverse
  transform, exp(i sigma V), band limit)--> phase gratings in HBM.

Units: cell lengths and positions [nm], Biso [nm^2], spatial frequencies
[1/nm], high tension [kV], form factors [nm].
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import lib, ptr


@dataclass
class Structure:
    """Orthorhombic structure: cell (a, b, c) [nm], per atom the atomic
    number z, fractional position frac (n, 3), Biso [nm^2] and occupancy."""
    cell: tuple
    z: np.ndarray
    frac: np.ndarray
    biso: np.ndarray
    occ: np.ndarray

    def types(self):
        """Atom types in order of first appearance: (z, biso) pairs, and the
        type index of every atom (CAtomType equality: same Z and Biso)."""
        keys, idx = [], np.zeros(len(self.z), np.int32)
        for i, (zz, bb) in enumerate(zip(self.z, self.biso)):
            k = (int(zz), float(np.float32(bb)))
            if k not in keys:
                keys.append(k)
            idx[i] = keys.index(k)
        return keys, idx


def weko_params(z: int):
    a, b = np.zeros(2), np.zeros(6)
    if lib().jms_weko_params(int(z), ptr(a), ptr(b)) != 0:
        raise ValueError(f"no Weickenmeier-Kohl parameters for Z = {z}")
    return a, b


def weko_form_factor(g: float, z: int, ekv: float, dw: float = 0.0, dwflg: bool = False, absflg: int = 0) -> complex:
    """Atomic form factor [nm] at g [1/nm] (CWEKOScat::wekosca); absflg 1:
    absorptive part as the reference computes it, 2: with the squared rms
    displacement in the thermal Debye-Waller term."""
    out = np.zeros(2)
    rc = lib().jms_weko_form_factor(float(g), float(dw), int(z), float(ekv), int(dwflg), int(absflg), ptr(out))
    if rc != 0:
        raise ValueError(f"jms_weko_form_factor: error {rc}")
    return complex(out[0], out[1])


def form_factor_sampling(nx: int, ny: int, ax: float, ay: float, nz: int = 1, az: float = 1.0):
    """(nk, sk, bwlk) of the radial tables for an nx x ny grid over ax x ay
    [nm]; nz = 1 and az = the structure's c in the projection mode."""
    nk, sk, bw = C.c_longlong(0), C.c_float(0), C.c_float(0)
    if lib().jms_form_factor_sampling(nx, ny, nz, ax, ay, az, C.byref(nk), C.byref(sk), C.byref(bw)) != 0:
        raise ValueError("invalid grid")
    return nk.value, sk.value, bw.value


def radial_form_factors(z: int, biso: float, ekv: float, nk: int, sk: float, bwlk: float, dwf: bool = True,
                        absorb: int = 0, abf: float = 0.0) -> np.ndarray:
    """Radial table (nk,) complex64 at q = ik sk (CAtomType::
    CalculateRadialFormFactors); absorb 0 none, 1 abf * elastic, 2 computed
    as the reference does, 3 computed with the squared displacement in the
    thermal Debye-Waller term."""
    out = np.zeros(nk, np.complex64)
    rc = lib().jms_radial_form_factors(int(z), float(biso), float(ekv), int(dwf), int(absorb), float(abf), int(nk),
                                       float(sk), float(bwlk), ptr(out))
    if rc != 0:
        raise ValueError(f"jms_radial_form_factors: error {rc}")
    return out


def form_factors_2d(nx: int, ny: int, ax: float, ay: float, tables, sk: float, bwl_pot: bool = False) -> np.ndarray:
    """Projected 2-D form factors (ntype, ny, nx) complex64 (FFT order) of
    radial tables (CJEElaSca::PrepareScatteringFunctions)."""
    tab = np.ascontiguousarray(np.stack([np.asarray(t, np.complex64) for t in tables]))
    out = np.zeros((len(tab), ny, nx), np.complex64)
    rc = lib().jms_form_factors_2d(nx, ny, ax, ay, int(bwl_pot), len(tab), tab.shape[1], float(sk), ptr(tab),
                                   ptr(out))
    if rc != 0:
        raise ValueError(f"jms_form_factors_2d: error {rc}")
    return out


def zslice(fz, fz0: float, fz1: float):
    """(indices, in-slice fractional z) of the atoms with wrapped fractional
    z in [fz0, fz1) (CStructure::GetZSliceStructure)."""
    f = np.ascontiguousarray(fz, np.float32)
    idx = np.zeros(max(1, len(f)), np.int32)
    fzn = np.zeros(max(1, len(f)), np.float32)
    k = lib().jms_zslice(len(f), ptr(f), float(fz0), float(fz1), ptr(idx), ptr(fzn))
    if k < 0:
        raise ValueError("invalid slice range")
    return idx[:k].copy(), fzn[:k].copy()


def suggest_zslices(st: Structure, dmin: float = 0.05) -> int:
    """Suggested number of equidistant slices (CStructure::
    SuggestEquidistantZSlices; the reference bench asks with dmin = 0.05 nm)."""
    z = np.ascontiguousarray(st.z, np.int32)
    fz = np.ascontiguousarray(np.asarray(st.frac)[:, 2], np.float32)
    return int(lib().jms_suggest_zslices(len(z), ptr(z), ptr(fz), float(st.cell[2]), float(dmin)))


def slice_structure(st: Structure, nslc: int):
    """Equidistant slices: per slice the atoms' (xy [nm] (k, 2), type index
    (k,), occupancy (k,)), and the slice thickness [nm] (the bench's
    setup_slices, prm_sample.cpp:752-790)."""
    _, tidx = st.types()
    frac = np.asarray(st.frac, np.float64)
    fz = np.float32(1.0) / np.float32(nslc)
    out = []
    for i in range(nslc):
        idx, _ = zslice(frac[:, 2], float(fz * np.float32(i)), float(fz * np.float32(i + 1)))
        xy = np.stack([frac[idx, 0] * st.cell[0], frac[idx, 1] * st.cell[1]], 1).astype(np.float32)
        out.append((xy, tidx[idx], np.asarray(st.occ, np.float32)[idx]))
    return out, float(np.float32(st.cell[2]) * fz)


def structure_form_factors(st: Structure, nx: int, ny: int, ekv: float, dwf: bool = True, absorb: int = 0,
                           abf: float = 0.0, bwl_pot: bool = False) -> np.ndarray:
    """2-D form factors of every atom type of st on the nx x ny grid over the
    cell's a x b (CJEElaSca::CalculateFormFactors + PrepareScatteringFunctions
    in projection mode)."""
    keys, _ = st.types()
    nk, sk, bwlk = form_factor_sampling(nx, ny, float(st.cell[0]), float(st.cell[1]), 1, float(st.cell[2]))
    tabs = [radial_form_factors(z, b, ekv, nk, sk, bwlk, dwf, absorb, abf) for z, b in keys]
    return form_factors_2d(nx, ny, float(st.cell[0]), float(st.cell[1]), tabs, sk, bwl_pot)


def gratings_from_structure(eng, st: Structure, nslc: int, ekv: float, dwf: bool = True, absorb: int = 0,
                            abf: float = 0.0, bwl_pot: bool = False, bwl_pgr: bool = True, variants=None):
    """Slice st into nslc equidistant slices and generate each slice's phase
    grating on the engine's device into structure slice i.  variants: an
    optional callable (islc, xy, type, occ) -> list of displaced xy arrays
    (frozen-phonon draws are the caller's); default one undisplaced variant.
    The engine's grid must span the cell's a x b."""
    ff2d = structure_form_factors(st, eng.sp.nx, eng.sp.ny, ekv, dwf, absorb, abf, bwl_pot)
    slices, cz = slice_structure(st, nslc)
    for i, (xy, t, o) in enumerate(slices):
        sets = [(x, t, o) for x in variants(i, xy, t, o)] if variants else [(xy, t, o)]
        eng.formfactor_gratings(i, ff2d, sets, cz, bwl_pot=bwl_pot, bwl_pgr=bwl_pgr)
    return ff2d, slices, cz


def perovskite_structure(ncell: int, a: float = 0.3905, biso=(0.0056, 0.0070, 0.0045)) -> Structure:
    """Cubic ABO3 (SrTiO3: Sr at the cell corners, Ti at the body centre, O at
    the face centres) over ncell x ncell cells and one cell along z; a [nm];
    Biso [nm^2] of Sr, O and Ti."""
    base = [(38, (0, 0, 0)), (8, (0.5, 0.5, 0)), (22, (0.5, 0.5, 0.5)), (8, (0.5, 0, 0.5)), (8, (0, 0.5, 0.5))]
    bz = {38: biso[0], 8: biso[1], 22: biso[2]}
    z, frac = [], []
    for iy in range(ncell):
        for ix in range(ncell):
            for zz, (fx, fy, fz) in base:
                z.append(zz)
                frac.append(((ix + fx) / ncell, (iy + fy) / ncell, fz))
    z = np.array(z, np.int32)
    return Structure(cell=(ncell * a, ncell * a, a), z=z, frac=np.array(frac),
                     biso=np.array([bz[v] for v in z]), occ=np.ones(len(z)))
