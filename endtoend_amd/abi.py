"""ctypes mirror of include/cccs.h and the loader for endtoend_amd/lib/libcccs.so.

The product path is the HIP library: there is no Python or CPU fallback.  `load()` raises if
the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcccs.so")

CCCS_OK = 0
CCCS_EINVAL = -1
CCCS_EGEOM = -2
CCCS_ENOMEM = -3
CCCS_EIO = -4
CCCS_EDEVICE = -5
CCCS_EBUSY = -6

CCCS_BEAMLET_FAITHFUL = 0
CCCS_BEAMLET_PLAIN = 2

_fp = C.POINTER(C.c_float)


class Problem(C.Structure):
    _fields_ = [
        ("size", C.c_uint32 * 3),
        ("voxel", C.c_float * 3),
        ("start", C.c_float * 3),
        ("bbox_start", C.c_uint32 * 3),
        ("bbox_size", C.c_uint32 * 3),
        ("max_rev_size", C.c_uint32 * 3),
        ("rev_latspacing", C.c_float),
        ("rev_longspacing", C.c_float),
        ("kernel_extent", C.c_float),
        ("penumbra", C.c_float),
        ("ss_factor", C.c_uint32),
        ("nphi", C.c_uint32),
        ("ntheta", C.c_uint32),
        ("nradii", C.c_uint32),
        ("beamhard_correct", C.c_float),
        ("density", _fp),
        ("kernel", _fp),
        ("radii", _fp),
        ("conv_theta_deg", _fp),
        ("conv_phi_deg", _fp),
        ("nspec", C.c_uint32),
        ("spec_fluence", _fp),
        ("spec_energy", _fp),
        ("spec_mu_en", _fp),
        ("spec_mu", _fp),
    ]


class Beam(C.Structure):
    _fields_ = [
        ("azimuth", C.c_float),
        ("zenith", C.c_float),
        ("coll", C.c_float),
        ("source", C.c_float * 3),
        ("direction", C.c_float * 3),
        ("isocenter", C.c_float * 3),
        ("sad", C.c_float),
        ("fmap_size", C.c_uint32 * 2),
        ("beamlet_size", C.c_float * 2),
    ]

    def copy(self) -> "Beam":
        b = Beam()
        C.memmove(C.addressof(b), C.addressof(self), C.sizeof(Beam))
        return b


class BeamGeometry(C.Structure):
    _fields_ = [
        ("lim_min", C.c_float * 3),
        ("bev_size", C.c_uint32 * 3),
        ("bev_min", C.c_float * 3),
        ("bev_max", C.c_float * 3),
        ("nrays", C.c_int),
        ("rev_min", (C.c_float * 3) * 48),
        ("rev_size", (C.c_uint32 * 3) * 48),
    ]


class SparseColumn(C.Structure):
    _fields_ = [
        ("beam_index", C.c_int),
        ("beamlet_index", C.c_int),
        ("nnz", C.c_uint64),
        ("keys", C.POINTER(C.c_uint64)),
        ("values", _fp),
        ("perc_nonzero", C.c_float),
        ("perc_dropped", C.c_float),
        ("nblocks", C.c_int),
        ("row_block_sizes", C.POINTER(C.c_uint64)),
        ("perc_reducedrop", C.c_float),
    ]


class CTInfo(C.Structure):
    _fields_ = [("size", C.c_uint32 * 3), ("voxel_mm", C.c_float * 3), ("start_mm", C.c_float * 3),
                ("orientation", C.c_float * 6), ("rescale_slope", C.c_float), ("rescale_intercept", C.c_float),
                ("sop_type", C.c_int)]


class TuneReport(C.Structure):
    _fields_ = [
        ("nkeys", C.c_int),
        ("key", (C.c_char * 32) * 16),
        ("value", C.c_int * 16),
        ("default_ms", C.c_double),
        ("best_ms", C.c_double),
        ("trials", C.c_int),
    ]


CCCS_TUNE_BEAM = 0
CCCS_TUNE_BEAMLET = 1


class KernelStats(C.Structure):
    _fields_ = [
        ("name", C.c_char * 32),
        ("launches", C.c_uint64),
        ("ms_total", C.c_double),
        ("alg_bytes", C.c_double),
        ("alg_flops", C.c_double),
        ("work_items", C.c_uint64),
    ]


SPARSE_SINK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(SparseColumn))


class ServerInfo(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double),
                ("batches", C.c_uint64), ("mean_batch", C.c_double), ("queued", C.c_uint64)]


CCCS_REQ_BEAM_DOSE = 0
CCCS_REQ_BEAMLET = 1


class SinkTotals(C.Structure):
    _fields_ = [("columns", C.c_uint64), ("nnz", C.c_uint64), ("value_sum", C.c_double)]

# (name, restype, argtypes) of every entry point declared in include/cccs.h
_P = C.c_void_p
EXPORTS = [
    ("cccs_ctx_create", C.c_int, [C.POINTER(Problem), C.c_int, C.POINTER(C.c_void_p)]),
    ("cccs_ctx_destroy", C.c_int, [_P]),
    ("cccs_ctx_memory", C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("cccs_ctx_set_schedule", C.c_int, [_P, C.c_char_p, C.c_int]),
    ("cccs_ctx_get_schedule", C.c_int, [_P, C.c_char_p, C.POINTER(C.c_int)]),
    ("cccs_autotune", C.c_int, [_P, C.c_int, C.POINTER(Beam), C.c_int, C.POINTER(_fp), C.POINTER(TuneReport)]),
    ("cccs_fluence_projection", C.c_int, [_P, C.c_void_p, C.POINTER(Beam), C.c_int, _fp]),
    ("cccs_fluence_convolve", C.c_int, [_P, _fp, C.c_int, C.c_int, C.c_int, _fp, C.c_int, C.c_int, _fp]),
    ("cccs_set_roi_masks", C.c_int, [_P, C.c_int, C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("cccs_fmap_aperture_fill", C.c_int, [_fp, C.c_int, C.c_int, C.c_float]),
    ("cccs_fmap_is_aperture_ready", C.c_int, [_fp, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    ("cccs_write_omni_beam_list", C.c_int, [C.c_char_p, C.POINTER(Beam), C.c_int, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_void_p]),
    ("cccs_beam_dose", C.c_int, [_P, C.POINTER(Beam), C.c_int, C.POINTER(_fp), _fp]),
    ("cccs_beam_dose_device", C.c_int, [_P, C.POINTER(Beam), C.c_int, _P, _P]),
    ("cccs_beam_dose_device_noalloc", C.c_int, [_P, C.POINTER(Beam), C.c_int, _P, _P]),
    ("cccs_angle_cache_create", C.c_int, [_P, C.POINTER(Beam), C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("cccs_angle_cache_info", C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
    ("cccs_ctx_use_angle_cache", C.c_int, [_P, _P]),
    ("cccs_angle_cache_destroy", C.c_int, [_P]),
    ("cccs_beam_terma", C.c_int, [_P, C.POINTER(Beam), _fp, C.c_int, _fp]),
    ("cccs_beam_geometry_of", C.c_int, [_P, C.POINTER(Beam), _fp, C.POINTER(BeamGeometry)]),
    ("cccs_siddon_paths", C.c_int, [_P, _fp, _fp, C.c_int, C.c_int, C.POINTER(C.c_int32), _fp, _fp,
                                    C.POINTER(C.c_int32), _fp]),
    ("cccs_beamlet_dose", C.c_int, [_P, C.POINTER(Beam), C.c_int, C.POINTER(_fp), C.c_float, C.c_int,
                                    SPARSE_SINK, _P]),
    ("cccs_beamlet_dense", C.c_int, [_P, C.POINTER(Beam), _fp, C.c_int, _fp, C.c_int, C.POINTER(C.c_int)]),
    ("cccs_sink_count", C.c_int, [_P, C.POINTER(SparseColumn)]),
    ("cccs_beamlet_footprint", C.c_int, [_P, C.POINTER(Beam), _fp, C.POINTER(C.c_uint64)]),
    ("cccs_server_create", C.c_int, [C.POINTER(Problem), C.c_int, C.c_int, C.c_uint64, C.c_int, C.POINTER(C.c_void_p)]),
    ("cccs_server_submit", C.c_int, [_P, C.c_int, C.POINTER(Beam), _fp, C.c_float, C.c_int, C.POINTER(C.c_uint64)]),
    ("cccs_server_wait", C.c_int, [_P, C.c_uint64, _fp, SPARSE_SINK, _P, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    ("cccs_server_stats", C.c_int, [_P, C.POINTER(ServerInfo)]),
    ("cccs_server_preload_angles", C.c_int, [_P, C.POINTER(Beam), C.c_int, C.c_uint64, C.POINTER(C.c_int)]),
    ("cccs_server_destroy", C.c_int, [_P]),
    ("cccs_dicom_ct_info", C.c_int, [C.c_char_p, C.POINTER(CTInfo)]),
    ("cccs_dicom_ct_read", C.c_int, [C.c_char_p, _fp, C.c_uint64]),
    ("cccs_dicom_rtstruct_open", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("cccs_rtstruct_count", C.c_int, [_P, C.POINTER(C.c_int)]),
    ("cccs_rtstruct_roi_info", C.c_int, [_P, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(C.c_uint64)]),
    ("cccs_rtstruct_roi_points", C.c_int, [_P, C.c_int, C.POINTER(C.c_uint32), _fp]),
    ("cccs_rtstruct_find", C.c_int, [_P, C.c_char_p, C.c_int, C.POINTER(C.c_int)]),
    ("cccs_rtstruct_close", C.c_int, [_P]),
    ("cccs_roi_extents", C.c_int, [_fp, C.c_uint64, C.POINTER(C.c_uint32), _fp, _fp, C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint32)]),
    ("cccs_roi_centroid", C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), _fp, _fp, _fp]),
    ("cccs_contour_rasterize", C.c_int, [_fp, C.POINTER(C.c_uint32), C.c_int, C.POINTER(C.c_uint32), _fp, _fp,
                                         C.c_void_p]),
    ("cccs_ct_lut_load", C.c_int, [C.c_char_p, _fp, _fp, C.c_int, C.POINTER(C.c_int)]),
    ("cccs_ct_to_density", C.c_int, [C.c_int, _fp, C.POINTER(C.c_uint32), _fp, C.c_float, _fp, _fp, C.c_int, _fp,
                                     C.c_uint64, C.POINTER(C.c_uint32)]),
    ("cccs_mask_resample", C.c_int, [C.c_int, C.c_void_p, C.POINTER(C.c_uint32), _fp, _fp, C.POINTER(C.c_uint32),
                                     _fp, _fp, C.c_void_p]),
    ("cccs_load_beam_list", C.c_int, [C.c_char_p, C.c_int, C.POINTER(Beam), C.c_int, C.POINTER(C.c_int)]),
    ("cccs_load_omni_beam_list", C.c_int, [C.c_char_p, C.c_int, C.POINTER(Beam), C.c_char_p, C.c_int,
                                           C.POINTER(C.c_int)]),
    ("cccs_load_omni_header", C.c_int, [C.c_char_p, C.POINTER(Problem), C.POINTER(C.c_int), C.c_char_p, C.c_char_p,
                                        C.POINTER(C.c_int)]),
    ("cccs_write_omni_header", C.c_int, [C.c_char_p, C.POINTER(Problem), C.c_int, C.c_char_p, C.c_char_p, C.c_int]),
    ("cccs_read_f32", C.c_int, [C.c_char_p, _fp, C.c_uint64]),
    ("cccs_write_f32", C.c_int, [C.c_char_p, _fp, C.c_uint64]),
    ("cccs_beam_reconfigure", C.c_int, [C.POINTER(Beam), C.c_int]),
    ("cccs_build_kernel_from_files", C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int, _fp, _fp,
                                               _fp, _fp, _fp, C.POINTER(C.c_int)]),
    ("cccs_poly_kernel_from_files", C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, _fp, _fp, _fp,
                                              C.POINTER(C.c_int)]),
    ("cccs_cumulative_kernel", C.c_int, [_fp, C.c_int, C.c_int, _fp, _fp, _fp]),
    ("cccs_dev_alloc", C.c_int, [_P, C.c_size_t, C.POINTER(C.c_void_p)]),
    ("cccs_dev_free", C.c_int, [_P, _P]),
    ("cccs_dev_upload", C.c_int, [_P, _P, _P, C.c_size_t]),
    ("cccs_dev_download", C.c_int, [_P, _P, _P, C.c_size_t]),
    ("cccs_dev_zero", C.c_int, [_P, _P, C.c_size_t]),
    ("cccs_dev_copy", C.c_int, [_P, _P, _P, C.c_size_t]),
    ("cccs_ctx_sync", C.c_int, [_P]),
    ("cccs_ctx_profile", C.c_int, [_P, C.c_int]),
    ("cccs_ctx_profile_reset", C.c_int, [_P]),
    ("cccs_ctx_profile_read", C.c_int, [_P, C.POINTER(KernelStats), C.c_int, C.POINTER(C.c_int)]),
    ("cccs_last_error", C.c_char_p, []),
]

_lib = None


class CCCSError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


def load(path: str | None = None) -> C.CDLL:
    """Load the HIP library (raises FileNotFoundError when it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is missing: build it with `python -m endtoend_amd.build`")
    lib = C.CDLL(p)
    for name, res, args in EXPORTS:
        fn = getattr(lib, name, None)
        if fn is None:
            raise ImportError(f"{p} does not export {name}")
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(code: int, where: str) -> None:
    if code != CCCS_OK:
        msg = load().cccs_last_error().decode(errors="replace")
        raise CCCSError(code, where, msg)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], "need C-contiguous float32"
    return a.ctypes.data_as(_fp)
