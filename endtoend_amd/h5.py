"""ctypes mirror of include/cccs_h5.h: the HDF5 file layer (endtoend_amd/lib/libcccs_h5.so).

Dose.h5 volumes, Dose_Coefficients.h5 coefficient files (written straight from the beamlet
pipeline through the library's native sink), fluence-map files and ROI lists, in the reference's
on-disk format (DoseCalcIO/volume.h:118-182, sparse_io.cpp:178-326, beam.cpp:134-298,
fmaps_io.cpp:10-184, roi.cpp:16-185, dosecalc-beamlet/sparsify_manager.cpp:11-221).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Sequence

import numpy as np

from . import abi

LIB_PATH = os.path.join(abi.HERE, "lib", "libcccs_h5.so")
WRITE_CENTRAL, WRITE_PER_BEAMLET = 0, 1


class CalcSpecs(C.Structure):
    _fields_ = [
        ("version", C.c_char_p),
        ("nbeams", C.c_int),
        ("sparsity_thresh", C.c_float),
        ("beam_spectrum", C.c_char_p),
        ("target_structure", C.c_char_p),
        ("nroi", C.c_int),
        ("roi_names", C.POINTER(C.c_char_p)),
        ("roi_capacities", C.POINTER(C.c_uint64)),
    ]


class BeamLabels(C.Structure):
    _fields_ = [("isocenter_type", C.c_char_p), ("isocenter_location", C.c_char_p), ("orientation_type", C.c_char_p)]


_P = C.c_void_p
_u32p = C.POINTER(C.c_uint32)
_f32p = C.POINTER(C.c_float)
EXPORTS = [
    ("cccs_h5_last_error", C.c_char_p, []),
    ("cccs_h5_write_volume", C.c_int, [C.c_char_p, C.c_char_p, _f32p, _u32p, _f32p, _f32p]),
    ("cccs_h5_read_volume", C.c_int, [C.c_char_p, C.c_char_p, _f32p, C.c_size_t, _u32p, _f32p, _f32p]),
    ("cccs_h5_coeff_create", C.c_int, [C.c_char_p, C.POINTER(abi.Problem), C.POINTER(CalcSpecs), C.c_int, C.c_uint64,
                                       C.POINTER(_P)]),
    ("cccs_h5_coeff_set_write_strategy", C.c_int, [_P, C.c_int]),
    ("cccs_h5_coeff_add_beams", C.c_int, [_P, C.POINTER(abi.Beam), C.c_int, C.POINTER(_f32p), C.c_int,
                                          C.POINTER(BeamLabels)]),
    ("cccs_h5_coeff_add_column", C.c_int, [_P, C.POINTER(abi.SparseColumn)]),
    ("cccs_h5_coeff_sink", C.c_int, [_P, C.POINTER(abi.SparseColumn)]),
    ("cccs_h5_coeff_counts", C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("cccs_h5_coeff_close", C.c_int, [_P]),
    ("cccs_h5_write_fmaps", C.c_int, [C.c_char_p, C.POINTER(abi.Beam), C.c_int, C.POINTER(_f32p),
                                      C.POINTER(BeamLabels)]),
    ("cccs_h5_read_fmaps", C.c_int, [C.c_char_p, C.POINTER(abi.Beam), C.POINTER(C.c_int), C.c_int,
                                     C.POINTER(C.c_int), _f32p, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("cccs_h5_write_roi_list", C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.POINTER(C.c_uint8)),
                                         _u32p, _u32p, _u32p]),
    ("cccs_h5_read_roi_list", C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_char * 64), _u32p,
                                        _u32p, _u32p, C.POINTER(C.c_uint8), C.c_size_t, C.POINTER(C.c_size_t)]),
]

_lib = None


def load() -> C.CDLL:
    """Load libcccs_h5.so (raises FileNotFoundError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `python -m endtoend_amd.build` "
                                "(needs the HDF5 C library)")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in EXPORTS:
        fn = getattr(lib, name, None)
        if fn is None:
            raise ImportError(f"{LIB_PATH} does not export {name}")
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int, where: str) -> None:
    if code != abi.CCCS_OK:
        raise abi.CCCSError(code, where, load().cccs_h5_last_error().decode(errors="replace"))


def _u32(v):
    return (C.c_uint32 * 3)(*[int(x) for x in v])


def _f32(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def _labels(labels, n):
    if labels is None:
        return None
    if len(labels) != n:
        raise ValueError("one label triple per beam")
    arr = (BeamLabels * n)()
    for i, (t, loc, o) in enumerate(labels):
        arr[i] = BeamLabels(t.encode() if t else None, loc.encode() if loc else None, o.encode() if o else None)
    return arr


def _fluence(fmaps):
    fm = [np.ascontiguousarray(f, np.float32) for f in fmaps]
    return fm, (_f32p * max(1, len(fm)))(*[abi.fptr(f) for f in fm])


# ---- volumes ----------------------------------------------------------------------------------------
def write_volume(path: str, vol: np.ndarray, start, voxel, dset: str = "volume") -> None:
    """Dose.h5 (Volume<float>::writeToFile): `vol` is [Nz][Ny][Nx] (x fastest)."""
    v = np.ascontiguousarray(vol, np.float32)
    if v.ndim != 3:
        raise ValueError("volume must be 3-D [Nz][Ny][Nx]")
    size = _u32((v.shape[2], v.shape[1], v.shape[0]))
    check(load().cccs_h5_write_volume(path.encode(), dset.encode(), abi.fptr(v), size, _f32(start), _f32(voxel)),
          "cccs_h5_write_volume")


def read_volume(path: str, dset: str = "data"):
    """(volume [Nz][Ny][Nx] float32, start (3,), voxel (3,)) of a Volume file."""
    lib = load()
    size, start, voxel = (C.c_uint32 * 3)(), (C.c_float * 3)(), (C.c_float * 3)()
    check(lib.cccs_h5_read_volume(path.encode(), dset.encode(), None, 0, size, start, voxel), "cccs_h5_read_volume")
    out = np.empty((size[2], size[1], size[0]), np.float32)
    check(lib.cccs_h5_read_volume(path.encode(), dset.encode(), abi.fptr(out), out.size, size, start, voxel),
          "cccs_h5_read_volume")
    return out, np.array(start[:], np.float32), np.array(voxel[:], np.float32)


# ---- coefficient files ------------------------------------------------------------------------------
class CoeffFile:
    """A Dose_Coefficients.h5 being written.  Declare each request's beams with add_beams, then
    run the beamlet pipeline with `sink`/`handle` as its sink (Engine.beamlet_dose_h5 does both),
    or add columns by hand with add_column."""

    def __init__(self, path: str, problem: abi.Problem, nbeams: int, sparsity_thresh: float = 1e-4,
                 beam_spectrum: str = "", target_structure: str = "", version: str = "endtoend_amd",
                 rois: tuple | None = None, compress_lvl: int = -1, chunksize: int = 50,
                 write_strategy: int = 0):
        self.lib = load()
        self._keep = []
        specs = CalcSpecs(version.encode(), nbeams, sparsity_thresh, beam_spectrum.encode(), target_structure.encode(),
                          0, None, None)
        if rois:
            names, caps = rois
            nm = (C.c_char_p * len(names))(*[n.encode() for n in names])
            cp = (C.c_uint64 * len(caps))(*[int(c) for c in caps])
            self._keep += [nm, cp]
            specs.nroi, specs.roi_names, specs.roi_capacities = len(names), nm, cp
        self._problem = problem
        h = C.c_void_p()
        check(self.lib.cccs_h5_coeff_create(path.encode(), C.byref(problem), C.byref(specs), compress_lvl, chunksize,
                                            C.byref(h)), "cccs_h5_coeff_create")
        self.handle = h.value
        if write_strategy:
            st = self.lib.cccs_h5_coeff_set_write_strategy(self.handle, write_strategy)
            if st != 0:
                msg = self.lib.cccs_h5_last_error().decode(errors="replace")
                self.lib.cccs_h5_coeff_close(self.handle)
                self.handle = None
                raise abi.CCCSError(st, "cccs_h5_coeff_set_write_strategy", msg)
        self.sink = abi.SPARSE_SINK(C.cast(self.lib.cccs_h5_coeff_sink, C.c_void_p).value)

    def add_beams(self, beams: Sequence[abi.Beam], fmaps: Sequence[np.ndarray], uid_base: int = 0,
                  labels=None) -> None:
        from .engine import beams_array
        fm, ptrs = _fluence(fmaps)
        lab = _labels(labels, len(beams))
        check(self.lib.cccs_h5_coeff_add_beams(self.handle, beams_array(beams), len(beams), ptrs, uid_base, lab),
              "cccs_h5_coeff_add_beams")

    def add_column(self, beam_index: int, beamlet_index: int, keys: np.ndarray, values: np.ndarray,
                   perc_nonzero: float = 0.0, perc_dropped: float = 0.0, blocks=None,
                   perc_reducedrop: float = 0.0) -> None:
        k = np.ascontiguousarray(keys, np.uint64)
        v = np.ascontiguousarray(values, np.float32)
        if k.shape != v.shape:
            raise ValueError("keys and values differ in length")
        col = abi.SparseColumn()
        col.beam_index, col.beamlet_index, col.nnz = beam_index, beamlet_index, k.size
        col.keys = k.ctypes.data_as(C.POINTER(C.c_uint64))
        col.values = v.ctypes.data_as(C.POINTER(C.c_float))
        col.perc_nonzero, col.perc_dropped, col.perc_reducedrop = perc_nonzero, perc_dropped, perc_reducedrop
        if blocks is not None:
            b = np.ascontiguousarray(blocks, np.uint64)
            col.nblocks, col.row_block_sizes = b.size, b.ctypes.data_as(C.POINTER(C.c_uint64))
        check(self.lib.cccs_h5_coeff_add_column(self.handle, C.byref(col)), "cccs_h5_coeff_add_column")

    def counts(self) -> tuple:
        cols, coeffs = C.c_uint64(), C.c_uint64()
        check(self.lib.cccs_h5_coeff_counts(self.handle, C.byref(cols), C.byref(coeffs)), "cccs_h5_coeff_counts")
        return int(cols.value), int(coeffs.value)

    def close(self) -> None:
        if self.handle:
            h, self.handle = self.handle, None
            check(self.lib.cccs_h5_coeff_close(h), "cccs_h5_coeff_close")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- fluence-map files ------------------------------------------------------------------------------
def write_fmaps(path: str, beams: Sequence[abi.Beam], fmaps: Sequence[np.ndarray], labels=None) -> None:
    from .engine import beams_array
    fm, ptrs = _fluence(fmaps)
    check(load().cccs_h5_write_fmaps(path.encode(), beams_array(beams) if beams else None, len(beams), ptrs,
                                     _labels(labels, len(beams))), "cccs_h5_write_fmaps")


def read_fmaps(path: str):
    """(beams, uids, fluence maps) of an fmaps or coefficient file, in beam_uid order.  Each map
    is returned flat reshaped to (fmap_y, fmap_x), x fastest."""
    lib = load()
    n, nw = C.c_int(), C.c_size_t()
    check(lib.cccs_h5_read_fmaps(path.encode(), None, None, 0, C.byref(n), None, 0, C.byref(nw)), "cccs_h5_read_fmaps")
    beams = (abi.Beam * max(1, n.value))()
    uids = (C.c_int * max(1, n.value))()
    w = np.zeros(max(1, nw.value), np.float32)
    check(lib.cccs_h5_read_fmaps(path.encode(), beams, uids, n.value, C.byref(n), abi.fptr(w), w.size, C.byref(nw)),
          "cccs_h5_read_fmaps")
    out, off = [], 0
    for i in range(n.value):
        fx, fy = beams[i].fmap_size[0], beams[i].fmap_size[1]
        out.append(w[off:off + fx * fy].reshape(fy, fx).copy())
        off += fx * fy
    return [beams[i] for i in range(n.value)], [uids[i] for i in range(n.value)], out


# ---- ROI lists --------------------------------------------------------------------------------------
def write_roi_list(path: str, names: Sequence[str], masks: Sequence[np.ndarray], size, crop_starts) -> None:
    """roi_list.h5: masks[r] is ROI r's crop box [cz][cy][cx] (1 = inside) starting at
    crop_starts[r] (x, y, z voxels) of a volume of `size` (x, y, z)."""
    ms = [np.ascontiguousarray(m, np.uint8) for m in masks]
    n = len(ms)
    nm = (C.c_char_p * max(1, n))(*[s.encode() for s in names])
    mp = (C.POINTER(C.c_uint8) * max(1, n))(*[m.ctypes.data_as(C.POINTER(C.c_uint8)) for m in ms])
    c0 = (C.c_uint32 * max(3, 3 * n))(*[int(v) for s in crop_starts for v in s])
    cs = (C.c_uint32 * max(3, 3 * n))(*[int(v) for m in ms for v in (m.shape[2], m.shape[1], m.shape[0])])
    check(load().cccs_h5_write_roi_list(path.encode(), n, nm, mp, _u32(size), c0, cs), "cccs_h5_write_roi_list")


def read_roi_list(path: str):
    """(size, [(name, crop_start, mask [cz][cy][cx] uint8), ...]) in the list's index order."""
    lib = load()
    n, nb = C.c_int(), C.c_size_t()
    size = (C.c_uint32 * 3)()
    check(lib.cccs_h5_read_roi_list(path.encode(), 0, C.byref(n), None, size, None, None, None, 0, C.byref(nb)),
          "cccs_h5_read_roi_list")
    k = max(1, n.value)
    names = ((C.c_char * 64) * k)()
    c0, cs = (C.c_uint32 * (3 * k))(), (C.c_uint32 * (3 * k))()
    buf = np.zeros(max(1, nb.value), np.uint8)
    check(lib.cccs_h5_read_roi_list(path.encode(), k, C.byref(n), names, size, c0, cs,
                                    buf.ctypes.data_as(C.POINTER(C.c_uint8)), buf.size, C.byref(nb)),
          "cccs_h5_read_roi_list")
    rois, off = [], 0
    for r in range(n.value):
        sx, sy, sz = cs[3 * r], cs[3 * r + 1], cs[3 * r + 2]
        m = buf[off:off + sx * sy * sz].reshape(sz, sy, sx).copy()
        off += sx * sy * sz
        rois.append((names[r].value.decode(), (c0[3 * r], c0[3 * r + 1], c0[3 * r + 2]), m))
    return tuple(size[:]), rois
