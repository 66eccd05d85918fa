"""EMS/SLI slice files and .bin STEM result channels (include/jms_io.h):
the on-disk formats of the reference's JMSBench1 driver, read and written
by the library's host code."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import lib, ptr


class EmsHeader(C.Structure):
    """jms_ems_header (prm_slice_file_info.h:37-55, 117-132)."""
    _fields_ = [
        ("grid_x", C.c_uint32), ("grid_y", C.c_uint32), ("version", C.c_uint32), ("num_var", C.c_uint32),
        ("data_type", C.c_uint32), ("data_offset", C.c_uint64), ("structure_offset", C.c_uint64),
        ("ekv", C.c_float), ("sx", C.c_float), ("sy", C.c_float), ("sz", C.c_float),
        ("title", C.c_char * 40), ("byte_swapped", C.c_int32),
    ]


def _chk(name, rc):
    if rc != 0:
        raise RuntimeError(f"{name} returned error code {rc}")


def read_ems(path: str):
    """(header, data (num_var, grid_y, grid_x) complex64) of an EMS/SLI file."""
    h = EmsHeader()
    _chk("jms_ems_read_header", lib().jms_ems_read_header(path.encode(), C.byref(h)))
    data = np.zeros((h.num_var, h.grid_y, h.grid_x), np.complex64)
    _chk("jms_ems_read_data", lib().jms_ems_read_data(path.encode(), C.byref(h), ptr(data)))
    return h, data


def write_ems(path: str, data: np.ndarray, sx_nm: float, sy_nm: float, sz_nm: float, ekv: float,
              title: str = "", data_type: int = 0):
    """Write a phase-grating file of data (num_var, ny, nx) complex64."""
    d = np.ascontiguousarray(data, np.complex64)
    if d.ndim == 2:
        d = d[None]
    h = EmsHeader()
    h.grid_x, h.grid_y, h.num_var = d.shape[2], d.shape[1], d.shape[0]
    h.data_type = data_type
    h.sx, h.sy, h.sz, h.ekv = sx_nm, sy_nm, sz_nm, ekv
    h.title = title.encode()[:40]
    _chk("jms_ems_write", lib().jms_ems_write(path.encode(), C.byref(h), ptr(d)))


def stem_images(res: np.ndarray, weight: float = 1.0) -> np.ndarray:
    """STEM images [idet][idetslc][pos] from scan results [pos][idetslc][idet]
    (the layout of the reference bench's result buffers)."""
    r = np.ascontiguousarray(res, np.float32)
    npos, nds, nd = r.shape
    out = np.zeros((nd, nds, npos), np.float32)
    _chk("jms_stem_images", lib().jms_stem_images(npos, nds, nd, ptr(r), weight, ptr(out)))
    return out


def write_bin_channel(path: str, images: np.ndarray, idet: int):
    """Raw float32 image series of detector channel idet (prm_result::save)."""
    im = np.ascontiguousarray(images, np.float32)
    nd, nds, npos = im.shape
    _chk("jms_bin_write_channel", lib().jms_bin_write_channel(path.encode(), ptr(im), npos, nds, nd, idet))
