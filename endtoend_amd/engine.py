"""Python-side handle on the HIP CCCS engine (a thin layer over include/cccs.h).

`ProblemSpec` carries the problem description (the reference's CONSTANTS + SHM_DATA),
`Engine` owns one resident device context (cccs_ctx), and `make_beam` builds a beam from
gantry/couch/collimator angles exactly as the reference's beam-list loader does
(DoseCalcIO/beam.cpp:441-443 then BEAM::reconfigure).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from typing import Optional, Sequence

import numpy as np

from . import abi

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
PI_F = np.float32(3.14159265358979323846)


@dataclasses.dataclass
class KernelTables:
    """Spectrum, radial boundaries and the 48x24 poly-energetic kernel of one beam quality."""
    spectrum: np.ndarray  # (4, nspec): fluence, energy, mu_en, mu
    radii: np.ndarray     # (24,)
    poly: np.ndarray      # (48, 24)


def load_tables(name: str = "spec_6mv") -> KernelTables:
    path = os.path.join(DATA_DIR, name + ".txt")
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines() if ln and not ln.startswith("#")]
    ns = int(lines[0].split()[1])
    spec = np.array([[float(v) for v in ln.split()] for ln in lines[1:1 + ns]], np.float32).T.copy()
    radii = np.array(lines[2 + ns].split(), np.float32)
    poly = np.array([[float(v) for v in ln.split()] for ln in lines[4 + ns:4 + ns + 48]], np.float32)
    return KernelTables(spec, radii, poly)


def cumulative_kernel(t: KernelTables, nphi: int, ntheta: int):
    """make_cumulative_kernel (kernel.cpp:260-431) through the library: (kernel, theta, phi)."""
    lib = abi.load()
    nr = t.poly.shape[1]
    kern = np.zeros((ntheta, nr), np.float32)
    nrays = nphi * ntheta // 2
    theta = np.zeros(nrays, np.float32)
    phi = np.zeros(nrays, np.float32)
    poly = np.ascontiguousarray(t.poly, np.float32)
    abi.check(lib.cccs_cumulative_kernel(abi.fptr(poly), nphi, ntheta, abi.fptr(kern), abi.fptr(theta),
                                         abi.fptr(phi)), "cccs_cumulative_kernel")
    return kern, theta, phi


@dataclasses.dataclass
class ProblemSpec:
    density: np.ndarray            # (nz, ny, nx) float32
    voxel: tuple
    start: tuple                   # centre of voxel 0 [cm]
    kernel: np.ndarray             # (ntheta, nradii) cumulative
    radii: np.ndarray
    theta: np.ndarray
    phi: np.ndarray
    spectrum: np.ndarray           # (4, nspec)
    nphi: int
    ntheta: int
    rev_latspacing: float
    rev_longspacing: float
    penumbra: float = 1.0
    kernel_extent: float = 2.0
    ss_factor: int = 1
    bbox_start: tuple = (0, 0, 0)
    bbox_size: tuple | None = None
    max_rev_size: tuple = (0, 0, 0)
    beamhard_correct: float = 1.0

    def __post_init__(self):
        self.density = np.ascontiguousarray(self.density, np.float32)
        nz, ny, nx = self.density.shape
        if self.bbox_size is None:
            self.bbox_size = (nx, ny, nz)
        for name in ("kernel", "radii", "theta", "phi", "spectrum"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name), np.float32))
        self._keep = []

    @property
    def size(self):
        nz, ny, nx = self.density.shape
        return (nx, ny, nz)

    @property
    def bbox_shape(self):
        """(z, y, x) shape of calc_bbox arrays."""
        bx, by, bz = self.bbox_size
        return (bz, by, bx)

    def struct(self) -> abi.Problem:
        p = abi.Problem()
        p.size[:] = self.size
        p.voxel[:] = [float(v) for v in self.voxel]
        p.start[:] = [float(v) for v in self.start]
        p.bbox_start[:] = [int(v) for v in self.bbox_start]
        p.bbox_size[:] = [int(v) for v in self.bbox_size]
        p.max_rev_size[:] = [int(v) for v in self.max_rev_size]
        p.rev_latspacing = self.rev_latspacing
        p.rev_longspacing = self.rev_longspacing
        p.kernel_extent = self.kernel_extent
        p.penumbra = self.penumbra
        p.ss_factor = self.ss_factor
        p.nphi = self.nphi
        p.ntheta = self.ntheta
        p.nradii = int(self.radii.shape[0])
        p.beamhard_correct = self.beamhard_correct
        spec = [np.ascontiguousarray(self.spectrum[k]) for k in range(4)]
        self._keep = spec
        p.density = abi.fptr(self.density)
        p.kernel = abi.fptr(self.kernel)
        p.radii = abi.fptr(self.radii)
        p.conv_theta_deg = abi.fptr(self.theta)
        p.conv_phi_deg = abi.fptr(self.phi)
        p.nspec = int(self.spectrum.shape[1])
        p.spec_fluence = abi.fptr(spec[0])
        p.spec_energy = abi.fptr(spec[1])
        p.spec_mu_en = abi.fptr(spec[2])
        p.spec_mu = abi.fptr(spec[3])
        return p


def deg2rad_f(deg: float) -> np.float32:
    """(float(deg) * PI) / 180.0f as the beam-list loader computes it (beam.cpp:441-442)."""
    return np.float32(np.float32(np.float32(deg) * PI_F) / np.float32(180.0))


def make_beam(gantry_deg: float, couch_deg: float = 0.0, coll_deg: float = 0.0, iso=(0.0, 0.0, 0.0),
              sad: float = 100.0, fmap_size=(20, 20), beamlet_size=(0.5, 0.5)) -> abi.Beam:
    b = abi.Beam()
    b.azimuth = deg2rad_f(gantry_deg)
    b.zenith = deg2rad_f(couch_deg)
    b.coll = deg2rad_f(coll_deg)
    b.isocenter[:] = [float(v) for v in iso]
    b.sad = sad
    b.fmap_size[:] = [int(v) for v in fmap_size]
    b.beamlet_size[:] = [float(v) for v in beamlet_size]
    abi.check(abi.load().cccs_beam_reconfigure(C.byref(b), 0), "cccs_beam_reconfigure")
    return b


def retarget_beam(beam: abi.Beam, iso, sad: float, fmap_size, beamlet_size) -> abi.Beam:
    """A copy of `beam` (same gantry, couch and collimator angles) with a new isocenter, SAD and
    fluence-map geometry, reconfigured (BEAM::reconfigure, beam.cpp:300-331)."""
    b = beam.copy()
    b.isocenter[:] = [float(v) for v in iso]
    b.sad = sad
    b.fmap_size[:] = [int(v) for v in fmap_size]
    b.beamlet_size[:] = [float(v) for v in beamlet_size]
    abi.check(abi.load().cccs_beam_reconfigure(C.byref(b), 0), "cccs_beam_reconfigure")
    return b


def beams_array(beams: Sequence[abi.Beam]):
    arr = (abi.Beam * len(beams))()
    for i, b in enumerate(beams):
        C.memmove(C.addressof(arr[i]), C.addressof(b), C.sizeof(abi.Beam))
    return arr


class AngleCache:
    """An angle cache (cccs_angle_cache_*): unit-fluence TERMA per beam source, resident in HBM."""

    def __init__(self, lib, h):
        self.lib, self.h = lib, h

    def info(self):
        """(distinct sources cached, device bytes)."""
        n, b = C.c_int(0), C.c_uint64(0)
        abi.check(self.lib.cccs_angle_cache_info(self.h, C.byref(n), C.byref(b)), "cccs_angle_cache_info")
        return n.value, b.value

    def close(self):
        if self.h:
            abi.check(self.lib.cccs_angle_cache_destroy(self.h), "cccs_angle_cache_destroy")
            self.h = None


class Engine:
    """One resident cccs_ctx on a HIP device."""

    def __init__(self, spec: ProblemSpec, device: int = 0):
        self.spec = spec
        self.lib = abi.load()
        self._p = spec.struct()
        h = C.c_void_p()
        abi.check(self.lib.cccs_ctx_create(C.byref(self._p), device, C.byref(h)), "cccs_ctx_create")
        self.ctx = h

    def close(self):
        if self.ctx:
            self.lib.cccs_ctx_destroy(self.ctx)
            self.ctx = None
            self._angle_cache = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def memory(self):
        r, s = C.c_uint64(), C.c_uint64()
        abi.check(self.lib.cccs_ctx_memory(self.ctx, C.byref(r), C.byref(s)), "cccs_ctx_memory")
        return r.value, s.value

    def fluence_projection(self, mask: np.ndarray, beams) -> list:
        """Structure mask (z, y, x; non-zero inside) projected onto each beam's fluence map
        (cccs_fluence_projection); returns one (fy, fx) float map per beam."""
        m = np.ascontiguousarray(mask, np.uint8)
        if m.shape != self.spec.density.shape:
            raise ValueError("mask must have the density's shape")
        arr = (abi.Beam * len(beams))(*beams)
        sizes = [(int(b.fmap_size[1]), int(b.fmap_size[0])) for b in beams]
        out = np.zeros(sum(a * b for a, b in sizes), np.float32)
        abi.check(self.lib.cccs_fluence_projection(self.ctx, m.ctypes.data, arr, len(beams), abi.fptr(out)),
                  "cccs_fluence_projection")
        maps, o = [], 0
        for fy, fx in sizes:
            maps.append(out[o:o + fy * fx].reshape(fy, fx).copy())
            o += fy * fx
        return maps

    def fluence_convolve(self, fmaps: np.ndarray, kernel: np.ndarray) -> np.ndarray:
        """Fluence maps (n, h, w) or (h, w) convolved with an odd (kh, kw) kernel, same size, zero
        outside (cccs_fluence_convolve)."""
        f = np.ascontiguousarray(fmaps, np.float32)
        one = f.ndim == 2
        if one:
            f = f[None]
        k = np.ascontiguousarray(kernel, np.float32)
        out = np.zeros_like(f)
        abi.check(self.lib.cccs_fluence_convolve(self.ctx, abi.fptr(f), f.shape[0], f.shape[2], f.shape[1],
                                                 abi.fptr(k), k.shape[1], k.shape[0], abi.fptr(out)),
                  "cccs_fluence_convolve")
        return out[0] if one else out

    def set_roi_masks(self, rois) -> None:
        """ROI reduction of later beamlet columns (cccs_set_roi_masks).  rois: list of
        (mask over the crop box as (z, y, x) bytes, crop_start (x, y, z), crop_size (x, y, z));
        an empty list turns it off."""
        n = len(rois)
        masks = [np.ascontiguousarray(m, np.uint8) for m, _, _ in rois]
        for m, (_, _, size) in zip(masks, rois):
            if m.shape != (size[2], size[1], size[0]):
                raise ValueError("mask shape must be (crop_size z, y, x)")
        ptrs = (C.c_void_p * max(1, n))(*[m.ctypes.data for m in masks])
        st = (C.c_uint32 * max(3, 3 * n))(*[int(v) for _, s0, _ in rois for v in s0])
        sz = (C.c_uint32 * max(3, 3 * n))(*[int(v) for _, _, s1 in rois for v in s1])
        abi.check(self.lib.cccs_set_roi_masks(self.ctx, n, C.cast(ptrs, C.c_void_p), st, sz), "cccs_set_roi_masks")

    def get_schedule(self, key: str) -> int:
        v = C.c_int(0)
        abi.check(self.lib.cccs_ctx_get_schedule(self.ctx, key.encode(), C.byref(v)), "cccs_ctx_get_schedule")
        return v.value

    def autotune(self, beams: Sequence[abi.Beam], fmaps: Sequence[np.ndarray], kind: int = abi.CCCS_TUNE_BEAMLET) -> dict:
        """Coordinate-descent autotuning of the launch-schedule knobs on this workload
        (cccs_autotune); the chosen values stay on the context."""
        fm = [np.ascontiguousarray(f, np.float32) for f in fmaps]
        ptrs = (abi._fp * len(fm))(*[abi.fptr(f) for f in fm])
        rep = abi.TuneReport()
        abi.check(self.lib.cccs_autotune(self.ctx, kind, beams_array(beams), len(beams), ptrs, C.byref(rep)),
                  "cccs_autotune")
        return {"knobs": {rep.key[i].value.decode(): int(rep.value[i]) for i in range(rep.nkeys)},
                "default_ms": rep.default_ms, "best_ms": rep.best_ms, "trials": rep.trials}

    def set_schedule(self, key: str, value: int):
        """Launch-schedule knob (cccs_ctx_set_schedule); e.g. "cone_ramp_min"."""
        abi.check(self.lib.cccs_ctx_set_schedule(self.ctx, key.encode(), int(value)), "cccs_ctx_set_schedule")

    def angle_cache(self, beams: Sequence[abi.Beam], max_bytes: int = 64 << 30) -> "AngleCache":
        """Resident unit-TERMA grids of the distinct beam sources in `beams` (cccs_angle_cache_create)."""
        h = C.c_void_p()
        abi.check(self.lib.cccs_angle_cache_create(self.ctx, beams_array(beams), len(beams), max_bytes, C.byref(h)),
                  "cccs_angle_cache_create")
        return AngleCache(self.lib, h)

    def use_angle_cache(self, cache: Optional["AngleCache"]):
        """Full-beam doses of this context take the TERMA of cached sources from `cache` (None: off)."""
        abi.check(self.lib.cccs_ctx_use_angle_cache(self.ctx, cache.h if cache is not None else None),
                  "cccs_ctx_use_angle_cache")
        self._angle_cache = cache  # kept alive while the context uses it

    def beam_dose(self, beams: Sequence[abi.Beam], fmaps: Sequence[np.ndarray]) -> np.ndarray:
        dose = np.zeros(self.spec.density.shape, np.float32)
        fm = [np.ascontiguousarray(f, np.float32) for f in fmaps]
        ptrs = (abi._fp * len(fm))(*[abi.fptr(f) for f in fm])
        arr = beams_array(beams)
        abi.check(self.lib.cccs_beam_dose(self.ctx, arr, len(beams), ptrs, abi.fptr(dose)), "cccs_beam_dose")
        return dose

    # ---- device-resident inputs (for throughput measurements) ----
    def dev_alloc(self, nbytes: int) -> int:
        ptr = C.c_void_p()
        abi.check(self.lib.cccs_dev_alloc(self.ctx, nbytes, C.byref(ptr)), "cccs_dev_alloc")
        return ptr.value

    def dev_free(self, ptr: int):
        abi.check(self.lib.cccs_dev_free(self.ctx, C.c_void_p(ptr)), "cccs_dev_free")

    def dev_upload(self, ptr: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr)
        abi.check(self.lib.cccs_dev_upload(self.ctx, C.c_void_p(ptr), a.ctypes.data_as(C.c_void_p), a.nbytes),
                  "cccs_dev_upload")

    def dev_download(self, ptr: int, out: np.ndarray):
        abi.check(self.lib.cccs_dev_download(self.ctx, out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes),
                  "cccs_dev_download")

    def dev_zero(self, ptr: int, nbytes: int):
        abi.check(self.lib.cccs_dev_zero(self.ctx, C.c_void_p(ptr), nbytes), "cccs_dev_zero")

    def dev_copy(self, dst: int, src: int, nbytes: int):
        abi.check(self.lib.cccs_dev_copy(self.ctx, C.c_void_p(dst), C.c_void_p(src), nbytes), "cccs_dev_copy")

    def sync(self):
        abi.check(self.lib.cccs_ctx_sync(self.ctx), "cccs_ctx_sync")

    def beam_dose_device(self, beams: Sequence[abi.Beam], d_fluence: int, d_dose: int):
        arr = beams_array(beams)
        abi.check(self.lib.cccs_beam_dose_device(self.ctx, arr, len(beams), C.c_void_p(d_fluence),
                                                 C.c_void_p(d_dose)), "cccs_beam_dose_device")

    def profile(self, enable: bool = True):
        abi.check(self.lib.cccs_ctx_profile(self.ctx, int(enable)), "cccs_ctx_profile")

    def profile_reset(self):
        abi.check(self.lib.cccs_ctx_profile_reset(self.ctx), "cccs_ctx_profile_reset")

    def profile_read(self) -> dict:
        rows = (abi.KernelStats * 16)()
        n = C.c_int(0)
        abi.check(self.lib.cccs_ctx_profile_read(self.ctx, rows, 16, C.byref(n)), "cccs_ctx_profile_read")
        return {rows[i].name.decode(): dict(launches=rows[i].launches, ms=rows[i].ms_total,
                                            bytes=rows[i].alg_bytes, flops=rows[i].alg_flops,
                                            items=rows[i].work_items) for i in range(n.value)}

    def beam_terma(self, beam: abi.Beam, fmap: np.ndarray, beamletnum: int = -1) -> np.ndarray:
        out = np.zeros(self.spec.bbox_shape, np.float32)
        f = np.ascontiguousarray(fmap, np.float32)
        abi.check(self.lib.cccs_beam_terma(self.ctx, C.byref(beam), abi.fptr(f), beamletnum, abi.fptr(out)),
                  "cccs_beam_terma")
        return out

    def geometry(self, beam: abi.Beam, fmap: np.ndarray) -> abi.BeamGeometry:
        g = abi.BeamGeometry()
        f = np.ascontiguousarray(fmap, np.float32)
        abi.check(self.lib.cccs_beam_geometry_of(self.ctx, C.byref(beam), abi.fptr(f), C.byref(g)),
                  "cccs_beam_geometry_of")
        return g

    def siddon_paths(self, source, points: np.ndarray, max_segments: int = 1024):
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        n = pts.shape[0]
        src = np.ascontiguousarray(source, np.float32)
        ijk = np.zeros((n, max_segments, 3), np.int32)
        mid = np.zeros((n, max_segments, 3), np.float32)
        ln = np.zeros((n, max_segments), np.float32)
        cnt = np.zeros(n, np.int32)
        rpl = np.zeros(n, np.float32)
        abi.check(self.lib.cccs_siddon_paths(self.ctx, abi.fptr(src), abi.fptr(pts), n, max_segments,
                                             ijk.ctypes.data_as(C.POINTER(C.c_int32)), abi.fptr(mid), abi.fptr(ln),
                                             cnt.ctypes.data_as(C.POINTER(C.c_int32)), abi.fptr(rpl)),
                  "cccs_siddon_paths")
        return ijk, mid, ln, cnt, rpl

    def beamlet_dense(self, beam: abi.Beam, fmap: np.ndarray, mode: int = abi.CCCS_BEAMLET_FAITHFUL,
                      max_beamlets: int | None = None) -> np.ndarray:
        f = np.ascontiguousarray(fmap, np.float32)
        nact = int((f > 0).sum())
        cap = max_beamlets or max(1, nact)
        out = np.zeros((cap,) + self.spec.bbox_shape, np.float32)
        n = C.c_int(0)
        abi.check(self.lib.cccs_beamlet_dense(self.ctx, C.byref(beam), abi.fptr(f), mode, abi.fptr(out), cap,
                                              C.byref(n)), "cccs_beamlet_dense")
        return out[:n.value]

    def beamlet_dose(self, beams: Sequence[abi.Beam], fmaps: Sequence[np.ndarray], threshold: float = 1e-4,
                     mode: int = abi.CCCS_BEAMLET_FAITHFUL, keep: bool = True, percs: bool = False):
        """Sparse columns of every active beamlet: a list of (beam, beamlet, keys, values), plus the
        row-block sizes of ROI-reduced columns, plus (perc_nonzero, perc_dropped, perc_reducedrop)
        as the last element when percs=True; with keep=False only the totals (the library's native
        counting sink, no Python per column); with keep=None the columns are built on the device and
        not copied out."""
        fm = [np.ascontiguousarray(f, np.float32) for f in fmaps]
        ptrs = (abi._fp * len(fm))(*[abi.fptr(f) for f in fm])
        arr = beams_array(beams)
        if keep is None:  # device-resident columns, nothing copied to the host
            abi.check(self.lib.cccs_beamlet_dose(self.ctx, arr, len(beams), ptrs, threshold, mode,
                                                 abi.SPARSE_SINK(), None), "cccs_beamlet_dose")
            return None
        if not keep:
            tot = abi.SinkTotals()
            sink = abi.SPARSE_SINK(C.cast(self.lib.cccs_sink_count, C.c_void_p).value)
            abi.check(self.lib.cccs_beamlet_dose(self.ctx, arr, len(beams), ptrs, threshold, mode, sink,
                                                 C.cast(C.pointer(tot), C.c_void_p)), "cccs_beamlet_dose")
            return {"columns": int(tot.columns), "nnz": int(tot.nnz), "value_sum": float(tot.value_sum)}
        cols = []

        def sink_fn(user, colp):
            c = colp.contents
            k = np.ctypeslib.as_array(c.keys, shape=(c.nnz,)).copy() if c.nnz else np.zeros(0, np.uint64)
            v = np.ctypeslib.as_array(c.values, shape=(c.nnz,)).copy() if c.nnz else np.zeros(0, np.float32)
            col = (c.beam_index, c.beamlet_index, k, v)
            if c.nblocks:  # ROI-reduced column: its row-block sizes as a fifth element
                col += (np.ctypeslib.as_array(c.row_block_sizes, shape=(c.nblocks,)).copy(),)
            if percs:
                col += ((float(c.perc_nonzero), float(c.perc_dropped), float(c.perc_reducedrop)),)
            cols.append(col)
            return 0

        cb = abi.SPARSE_SINK(sink_fn)
        abi.check(self.lib.cccs_beamlet_dose(self.ctx, arr, len(beams), ptrs, threshold, mode, cb, None),
                  "cccs_beamlet_dose")
        return cols

    def beamlet_dose_h5(self, beams: Sequence[abi.Beam], fmaps: Sequence[np.ndarray], coeff_file, uid_base: int = 0,
                        labels=None, threshold: float = 1e-4, mode: int = abi.CCCS_BEAMLET_FAITHFUL) -> None:
        """Columns of every active beamlet written straight into an open h5.CoeffFile (the native
        HDF5 sink; no Python per column): beam i of this request becomes beam uid uid_base + i."""
        coeff_file.add_beams(beams, fmaps, uid_base=uid_base, labels=labels)
        fm = [np.ascontiguousarray(f, np.float32) for f in fmaps]
        ptrs = (abi._fp * len(fm))(*[abi.fptr(f) for f in fm])
        rc = self.lib.cccs_beamlet_dose(self.ctx, beams_array(beams), len(beams), ptrs, threshold, mode,
                                        coeff_file.sink, C.c_void_p(coeff_file.handle))
        if rc != abi.CCCS_OK:
            from . import h5
            detail = h5.load().cccs_h5_last_error().decode(errors="replace")
            msg = self.lib.cccs_last_error().decode(errors="replace")
            raise abi.CCCSError(rc, "cccs_beamlet_dose", f"{msg}{'; ' + detail if detail else ''}")

# ---- the reference's file formats (host functions of the library; no device needed) ----------

def load_beam_list(path: str, requested: int = 0) -> list:
    """Beams of a beam list file (load_beam_list semantics), reconfigured."""
    lib = abi.load()
    n = C.c_int(0)
    cap = 4096
    while True:
        arr = (abi.Beam * cap)()
        rc = lib.cccs_load_beam_list(path.encode(), requested, arr, cap, C.byref(n))
        if rc == abi.CCCS_EINVAL and n.value > cap:
            cap = n.value
            continue
        abi.check(rc, "cccs_load_beam_list")
        return [arr[i] for i in range(n.value)]


def load_omni_beam_list(path: str, beam_count: int):
    """(beams, fluence-map file names) of an omni_beam_list.txt."""
    lib = abi.load()
    arr = (abi.Beam * max(1, beam_count))()
    names = C.create_string_buffer(1024 * max(1, beam_count))
    n = C.c_int(0)
    abi.check(lib.cccs_load_omni_beam_list(path.encode(), beam_count, arr, names, 1024, C.byref(n)),
              "cccs_load_omni_beam_list")
    raw = names.raw
    fn = [raw[i * 1024:(i + 1) * 1024].split(b"\0", 1)[0].decode() for i in range(n.value)]
    return [arr[i] for i in range(n.value)], fn


def load_omni_header(path: str) -> dict:
    lib = abi.load()
    p = abi.Problem()
    bc, red = C.c_int(0), C.c_int(0)
    spec, target = C.create_string_buffer(100), C.create_string_buffer(100)
    abi.check(lib.cccs_load_omni_header(path.encode(), C.byref(p), C.byref(bc), spec, target, C.byref(red)),
              "cccs_load_omni_header")
    return {"problem": p, "beam_count": bc.value, "beam_spec": spec.value.decode(),
            "target_structure": target.value.decode(), "reduce": bool(red.value)}


def write_omni_header(path: str, p, beam_count: int, beam_spec: str, target: str, reduce: bool = False) -> None:
    abi.check(abi.load().cccs_write_omni_header(path.encode(), C.byref(p), beam_count, beam_spec.encode(),
                                                target.encode(), int(reduce)), "cccs_write_omni_header")


def write_omni_beam_list(path: str, beams, iso_type=None, iso_loc=None, orient_type=None, fmap_names=None) -> None:
    """omni_beam_list.txt (write_omni_beam_list semantics); string lists default per beam."""
    n = len(beams)
    arr = (abi.Beam * max(1, n))(*beams)

    def strs(v):
        if v is None:
            return None
        a = (C.c_char_p * n)(*[x.encode() for x in v])
        return C.cast(a, C.c_void_p), a

    keep = [strs(v) for v in (iso_type, iso_loc, orient_type, fmap_names)]
    ptrs = [k[0] if k else None for k in keep]
    abi.check(abi.load().cccs_write_omni_beam_list(path.encode(), arr, n, *ptrs), "cccs_write_omni_beam_list")


def fmap_aperture_fill(fmap: np.ndarray, fill_val: float = 1.0) -> np.ndarray:
    """Aperture-ready fill (fmap_post_apertureready semantics) of a (fy, fx) map; returns a copy."""
    f = np.ascontiguousarray(fmap, np.float32).copy()
    abi.check(abi.load().cccs_fmap_aperture_fill(abi.fptr(f), f.shape[1], f.shape[0], fill_val),
              "cccs_fmap_aperture_fill")
    return f


def fmap_is_aperture_ready(fmap: np.ndarray) -> bool:
    f = np.ascontiguousarray(fmap, np.float32)
    r = C.c_int(0)
    abi.check(abi.load().cccs_fmap_is_aperture_ready(abi.fptr(f), f.shape[1], f.shape[0], C.byref(r)),
              "cccs_fmap_is_aperture_ready")
    return bool(r.value)


def read_f32(path: str, shape) -> np.ndarray:
    out = np.zeros(shape, np.float32)
    abi.check(abi.load().cccs_read_f32(path.encode(), abi.fptr(out), out.size), "cccs_read_f32")
    return out


def write_f32(path: str, data: np.ndarray) -> None:
    d = np.ascontiguousarray(data, np.float32)
    abi.check(abi.load().cccs_write_f32(path.encode(), abi.fptr(d), d.size), "cccs_write_f32")

class Server:
    """Request loop over device-resident contexts (cccs_server_*): submit single-beam dose or
    beamlet-column requests as they arrive, wait for each; p50/p99 from stats()."""

    def __init__(self, spec: "ProblemSpec", slots: int = 2, hbm_budget: int = 64 << 30, max_batch: int = 64,
                 device: int = 0):
        self.lib = abi.load()
        self.spec = spec
        self._p = spec.struct()
        h = C.c_void_p()
        abi.check(self.lib.cccs_server_create(C.byref(self._p), device, slots, hbm_budget, max_batch, C.byref(h)),
                  "cccs_server_create")
        self.h = h

    def submit(self, beam: abi.Beam, fmap: np.ndarray, kind: int = abi.CCCS_REQ_BEAM_DOSE, threshold: float = 1e-4,
               mode: int = abi.CCCS_BEAMLET_FAITHFUL) -> int:
        f = np.ascontiguousarray(fmap, np.float32)
        rid = C.c_uint64(0)
        abi.check(self.lib.cccs_server_submit(self.h, kind, C.byref(beam), abi.fptr(f), threshold, mode, C.byref(rid)),
                  "cccs_server_submit")
        return rid.value

    def wait(self, rid: int, kind: int = abi.CCCS_REQ_BEAM_DOSE, counts_only: bool = False):
        """(result, latency_ms, batch_size): a dose volume, or a list of (beamlet, keys, values); with
        counts_only a beamlet request's columns go to the library's native counting sink and the
        result is their totals."""
        lat, bs = C.c_double(0), C.c_int(0)
        if kind == abi.CCCS_REQ_BEAMLET and counts_only:
            tot = abi.SinkTotals()
            sink = abi.SPARSE_SINK(C.cast(self.lib.cccs_sink_count, C.c_void_p).value)
            abi.check(self.lib.cccs_server_wait(self.h, rid, None, sink, C.cast(C.pointer(tot), C.c_void_p),
                                                C.byref(lat), C.byref(bs)), "cccs_server_wait")
            return ({"columns": int(tot.columns), "nnz": int(tot.nnz), "value_sum": float(tot.value_sum)}, lat.value,
                    bs.value)
        if kind == abi.CCCS_REQ_BEAM_DOSE:
            dose = np.zeros(self.spec.density.shape, np.float32)
            abi.check(self.lib.cccs_server_wait(self.h, rid, abi.fptr(dose), abi.SPARSE_SINK(), None, C.byref(lat),
                                                C.byref(bs)), "cccs_server_wait")
            return dose, lat.value, bs.value
        cols = []

        def sink_fn(user, colp):
            c = colp.contents
            k = np.ctypeslib.as_array(c.keys, shape=(c.nnz,)).copy() if c.nnz else np.zeros(0, np.uint64)
            v = np.ctypeslib.as_array(c.values, shape=(c.nnz,)).copy() if c.nnz else np.zeros(0, np.float32)
            cols.append((c.beamlet_index, k, v))
            return 0

        cb = abi.SPARSE_SINK(sink_fn)
        abi.check(self.lib.cccs_server_wait(self.h, rid, None, cb, None, C.byref(lat), C.byref(bs)), "cccs_server_wait")
        return cols, lat.value, bs.value

    def preload_angles(self, beams: Sequence[abi.Beam], max_bytes: int = 64 << 30) -> int:
        """Angle cache of `beams` shared by every slot (cccs_server_preload_angles); returns the number
        of distinct sources cached."""
        n = C.c_int(0)
        abi.check(self.lib.cccs_server_preload_angles(self.h, beams_array(beams), len(beams), max_bytes, C.byref(n)),
                  "cccs_server_preload_angles")
        return n.value

    def stats(self) -> dict:
        s = abi.ServerInfo()
        abi.check(self.lib.cccs_server_stats(self.h, C.byref(s)), "cccs_server_stats")
        return {k: getattr(s, k) for k, _ in abi.ServerInfo._fields_}

    def close(self):
        if self.h:
            abi.check(self.lib.cccs_server_destroy(self.h), "cccs_server_destroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
