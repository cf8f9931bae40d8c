"""DICOM ingestion: the dosecalc-preprocess steps from a DICOM directory to the inputs of a dose
calculation (CCCS/dosecalc-preprocess/main.cu:139-320), over the library's C entry points.

  read_ct            CT series (RTImage, RTClasses/rtimages.cpp:20-400)
  RTStruct           structure set (RTStruct, RTClasses/rtstruct.cpp:100-380; getROIIndex)
  ct_to_density      HU -> density on the isotropic grid, on the GPU (read_dicom / cudaCreateTexIso)
  contour_mask       ROI mask on the calculation grid (generateContourMask: host rasteriser on the
                     CT grid, GPU resample + threshold)
  roi_extents, roi_centroid   (DoseCalcIO/roi.cpp:245-345)
  preprocess         the whole chain: density, calc_bbox (with the reference's edge margin),
                     PTV mask and its centroid (the isocentre of beams without one)
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Optional, Sequence

import numpy as np

from . import abi

# HU -> mass density (g/cm^3) of the reference's default table ("Siemens (default)",
# main.cu:186-201; from sbcrowe.net/ct-density-tables), sorted by HU
SIEMENS_LUT = (
    np.array([-969.8, -712.9, -536.5, -95.6, -45.6, -5.6, -1.9, 25.7, 65.6, 207.5, 220.7, 429.9, 775.3, 1173.7],
             np.float32),
    np.array([0.0, 0.290, 0.450, 0.943, 0.985, 1.000, 1.016, 1.052, 1.089, 1.145, 1.159, 1.335, 1.560, 1.823],
             np.float32),
)


@dataclasses.dataclass
class Frame:
    """A grid: size (x, y, z), start and spacing in cm."""
    size: tuple
    start: tuple
    spacing: tuple


@dataclasses.dataclass
class CTVolume:
    hu: np.ndarray            # [z, y, x] FP32
    frame: Frame              # cm (the reference scales the DICOM mm by 0.1)
    info: abi.CTInfo


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def _u3(v):
    return (C.c_uint32 * 3)(*[int(x) for x in v])


def read_ct(dicom_dir: str) -> CTVolume:
    lib = abi.load()
    info = abi.CTInfo()
    abi.check(lib.cccs_dicom_ct_info(dicom_dir.encode(), C.byref(info)), "cccs_dicom_ct_info")
    nx, ny, nz = info.size
    hu = np.zeros((nz, ny, nx), np.float32)
    abi.check(lib.cccs_dicom_ct_read(dicom_dir.encode(), abi.fptr(hu), hu.size), "cccs_dicom_ct_read")
    f32 = np.float32
    frame = Frame((nx, ny, nz), tuple(float(f32(0.1) * f32(v)) for v in info.start_mm),
                  tuple(float(f32(0.1) * f32(v)) for v in info.voxel_mm))
    return CTVolume(hu, frame, info)


class RTStruct:
    """The first RT Structure Set of a directory."""

    def __init__(self, dicom_dir: str):
        self.lib = abi.load()
        h = C.c_void_p()
        abi.check(self.lib.cccs_dicom_rtstruct_open(dicom_dir.encode(), C.byref(h)), "cccs_dicom_rtstruct_open")
        self.h = h
        n = C.c_int(0)
        abi.check(self.lib.cccs_rtstruct_count(self.h, C.byref(n)), "cccs_rtstruct_count")
        self.names, self.numbers = [], []
        for r in range(n.value):
            buf = C.create_string_buffer(256)
            num, nsub, npts = C.c_int(0), C.c_int(0), C.c_uint64(0)
            abi.check(self.lib.cccs_rtstruct_roi_info(self.h, r, buf, 256, C.byref(num), C.byref(nsub), C.byref(npts)),
                      "cccs_rtstruct_roi_info")
            self.names.append(buf.value.decode(errors="replace"))
            self.numbers.append(num.value)

    def find(self, search: str, exact: bool = False) -> int:
        r = C.c_int(-1)
        abi.check(self.lib.cccs_rtstruct_find(self.h, search.encode(), int(exact), C.byref(r)), "cccs_rtstruct_find")
        return r.value

    def contours(self, roi: int):
        """(points per contour [uint32], points [n, 3] FP32 in mm)."""
        num, nsub, npts = C.c_int(0), C.c_int(0), C.c_uint64(0)
        abi.check(self.lib.cccs_rtstruct_roi_info(self.h, roi, None, 0, C.byref(num), C.byref(nsub), C.byref(npts)),
                  "cccs_rtstruct_roi_info")
        counts = np.zeros(nsub.value, np.uint32)
        pts = np.zeros((npts.value, 3), np.float32)
        abi.check(self.lib.cccs_rtstruct_roi_points(self.h, roi, counts.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                    abi.fptr(pts)), "cccs_rtstruct_roi_points")
        return counts, pts

    def close(self):
        if self.h:
            self.lib.cccs_rtstruct_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_lut(path: str):
    lib = abi.load()
    n = C.c_int(0)
    abi.check(lib.cccs_ct_lut_load(path.encode(), None, None, 0, C.byref(n)), "cccs_ct_lut_load")
    hu, de = np.zeros(n.value, np.float32), np.zeros(n.value, np.float32)
    abi.check(lib.cccs_ct_lut_load(path.encode(), abi.fptr(hu), abi.fptr(de), n.value, C.byref(n)), "cccs_ct_lut_load")
    return hu, de


def ct_to_density(ct: CTVolume, iso_voxel_cm: float, lut=SIEMENS_LUT, device: int = 0):
    """(density [z, y, x], Frame) on the isotropic grid; lut None converts by (HU + 1024) / 1024."""
    lib = abi.load()
    h = np.ascontiguousarray(ct.hu, np.float32)
    size, vox = _u3(ct.frame.size), _f3(ct.frame.spacing)
    osz = (C.c_uint32 * 3)()
    lh = ld = None
    n = 0
    if lut is not None:
        lh, ld = np.ascontiguousarray(lut[0], np.float32), np.ascontiguousarray(lut[1], np.float32)
        n = len(lh)
    args = (abi.fptr(lh) if n else None, abi.fptr(ld) if n else None, n)
    abi.check(lib.cccs_ct_to_density(device, abi.fptr(h), size, vox, C.c_float(iso_voxel_cm), *args, None, 0, osz),
              "cccs_ct_to_density")
    out = np.zeros((osz[2], osz[1], osz[0]), np.float32)
    abi.check(lib.cccs_ct_to_density(device, abi.fptr(h), size, vox, C.c_float(iso_voxel_cm), *args, abi.fptr(out),
                                     out.size, osz), "cccs_ct_to_density")
    f = np.float32(iso_voxel_cm)
    return out, Frame(tuple(osz), ct.frame.start, (float(f),) * 3)


def rasterize(counts: np.ndarray, points_mm: np.ndarray, ct_frame: Frame) -> np.ndarray:
    """Filled contours on the CT grid ([z, y, x] bytes), host side of generateContourMask."""
    lib = abi.load()
    c = np.ascontiguousarray(counts, np.uint32)
    p = np.ascontiguousarray(points_mm, np.float32)
    nx, ny, nz = ct_frame.size
    m = np.zeros((nz, ny, nx), np.uint8)
    abi.check(lib.cccs_contour_rasterize(abi.fptr(p), c.ctypes.data_as(C.POINTER(C.c_uint32)), len(c),
                                         _u3(ct_frame.size), _f3(ct_frame.start), _f3(ct_frame.spacing),
                                         m.ctypes.data), "cccs_contour_rasterize")
    return m


def resample_mask(mask: np.ndarray, src: Frame, dst: Frame, device: int = 0) -> np.ndarray:
    lib = abi.load()
    m = np.ascontiguousarray(mask, np.uint8)
    out = np.zeros((dst.size[2], dst.size[1], dst.size[0]), np.uint8)
    abi.check(lib.cccs_mask_resample(device, m.ctypes.data, _u3(src.size), _f3(src.start), _f3(src.spacing),
                                     _u3(dst.size), _f3(dst.start), _f3(dst.spacing), out.ctypes.data),
              "cccs_mask_resample")
    return out


def contour_mask(counts, points_mm, ct_frame: Frame, dens_frame: Frame, device: int = 0) -> np.ndarray:
    """ROI mask on the calculation grid (generateContourMask, cudaMemFunctions.cu:218-295)."""
    return resample_mask(rasterize(counts, points_mm, ct_frame), ct_frame, dens_frame, device)


def roi_extents(points_mm: np.ndarray, frame: Frame):
    lib = abi.load()
    p = np.ascontiguousarray(points_mm, np.float32)
    cs, cz = (C.c_uint32 * 3)(), (C.c_uint32 * 3)()
    abi.check(lib.cccs_roi_extents(abi.fptr(p), p.size // 3, _u3(frame.size), _f3(frame.start), _f3(frame.spacing),
                                   cs, cz), "cccs_roi_extents")
    return tuple(cs), tuple(cz)


def roi_centroid(mask: np.ndarray, frame: Frame) -> np.ndarray:
    lib = abi.load()
    m = np.ascontiguousarray(mask, np.uint8)
    c = (C.c_float * 3)()
    abi.check(lib.cccs_roi_centroid(m.ctypes.data, _u3(frame.size), _f3(frame.start), _f3(frame.spacing), c),
              "cccs_roi_centroid")
    return np.array(c[:], np.float32)


def preprocess(dicom_dir: str, ptv: str, iso_voxel_cm: float, bbox_roi: Optional[str] = "Body",
               lut=SIEMENS_LUT, exact_ptv: bool = False, device: int = 0) -> dict:
    """The DICOM branch of dosecalc-preprocess (main.cu:139-320): density on the isotropic grid,
    the calc_bbox from the bbox ROI (exact name; the whole volume when absent) with the reference's
    one-voxel margin from the edges, the PTV mask on the density grid and its centroid."""
    ct = read_ct(dicom_dir)
    dens, frame = ct_to_density(ct, iso_voxel_cm, lut, device)
    with RTStruct(dicom_dir) as rt:
        ip = rt.find(ptv, exact_ptv)
        if ip < 0:
            raise KeyError(f"no structure matches {ptv!r}")
        pc, pp = rt.contours(ip)
        ptv_mask = contour_mask(pc, pp, ct.frame, frame, device)
        ib = rt.find(bbox_roi, True) if bbox_roi else -1
        if ib < 0:
            start, size = (0, 0, 0), frame.size
        else:
            bc, bp = (pc, pp) if ib == ip else rt.contours(ib)
            start, size = roi_extents(bp, frame)
        names = list(rt.names)
    # main.cu:284-290: keep the box off the volume's first voxel and size it below the volume's by 2
    start = tuple(max(int(s), 1) for s in start)
    size = tuple(min(int(z), int(n) - 2) for z, n in zip(size, frame.size))
    return {"density": dens, "frame": frame, "bbox_start": start, "bbox_size": size, "ptv_mask": ptv_mask,
            "isocenter": roi_centroid(ptv_mask, frame), "roi_names": names, "ct": ct}
