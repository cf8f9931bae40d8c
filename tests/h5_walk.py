"""A small HDF5 reader for the tests, written against the HDF5 C library with ctypes and
independent of the library under test: it walks a file and returns every group, attribute and
dataset with the file's own datatype (byte order, sign, width, array/compound layout), so tests
can check the written files type for type against the reference's writers, and can rebuild the
dose-coefficient matrix the way the reference's Python reader does
(CCCS/extern/python/read_dose_data.py:34-133).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_hid = C.c_int64
_lib = None

H5I_GROUP, H5I_DATASET = 2, 5
H5T_INTEGER, H5T_FLOAT, H5T_STRING, H5T_COMPOUND, H5T_ARRAY = 0, 1, 3, 6, 10
H5S_SCALAR = 0
H5D_CONTIGUOUS, H5D_CHUNKED = 1, 2


def library_path() -> str | None:
    from endtoend_amd import build
    root = build.hdf5_root()
    return os.path.join(root, "lib", "libhdf5.so") if root else None


def _load():
    global _lib
    if _lib is None:
        lib = C.CDLL(library_path())
        lib.H5open()
        lib.H5Eset_auto2(_hid(0), None, None)
        for n in ("H5Fopen", "H5Oopen", "H5Aopen", "H5Aget_type", "H5Aget_space", "H5Dopen2", "H5Dget_type",
                  "H5Dget_space", "H5Dget_create_plist", "H5Tget_super", "H5Tget_member_type", "H5Tcopy"):
            getattr(lib, n).restype = _hid
        lib.H5Tget_member_name.restype = C.c_void_p
        lib.H5Tget_member_offset.restype = C.c_size_t
        lib.H5Tget_size.restype = C.c_size_t
        lib.H5Sget_simple_extent_npoints.restype = C.c_int64
        _lib = lib
    return _lib


def _native(name):
    return _hid.in_dll(_load(), name + "_g").value


@dataclass
class Attr:
    value: object
    dtype: str          # numpy-style description of the FILE type ('<u2', '<f4', 'S4', 'vlen-str', ...)
    scalar: bool        # scalar dataspace (else simple)
    shape: tuple


@dataclass
class Dataset:
    data: np.ndarray
    dtype: str
    layout: str         # 'contiguous' | 'chunked' | 'compact'
    chunk: tuple
    nfilters: int
    attrs: dict = field(default_factory=dict)


@dataclass
class Group:
    attrs: dict = field(default_factory=dict)
    groups: dict = field(default_factory=dict)
    datasets: dict = field(default_factory=dict)


def _np_dtype(t):
    """numpy dtype of HDF5 datatype t (file layout: offsets and byte order kept)."""
    lib = _load()
    cls = lib.H5Tget_class(_hid(t))
    size = lib.H5Tget_size(_hid(t))
    if cls == H5T_INTEGER:
        order = "<" if lib.H5Tget_order(_hid(t)) == 0 else ">"
        return np.dtype(f"{order}{'i' if lib.H5Tget_sign(_hid(t)) == 1 else 'u'}{size}")
    if cls == H5T_FLOAT:
        order = "<" if lib.H5Tget_order(_hid(t)) == 0 else ">"
        return np.dtype(f"{order}f{size}")
    if cls == H5T_ARRAY:
        nd = lib.H5Tget_array_ndims(_hid(t))
        dims = (C.c_uint64 * nd)()
        lib.H5Tget_array_dims2(_hid(t), dims)
        base = lib.H5Tget_super(_hid(t))
        bt = _np_dtype(base)
        lib.H5Tclose(_hid(base))
        return np.dtype((bt, tuple(dims)))
    if cls == H5T_COMPOUND:
        names, formats, offsets = [], [], []
        for i in range(lib.H5Tget_nmembers(_hid(t))):
            p = lib.H5Tget_member_name(_hid(t), i)
            names.append(C.string_at(p).decode())
            lib.H5free_memory(C.c_void_p(p))
            offsets.append(lib.H5Tget_member_offset(_hid(t), i))
            mt = lib.H5Tget_member_type(_hid(t), i)
            formats.append(_np_dtype(mt))
            lib.H5Tclose(_hid(mt))
        return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": size})
    if cls == H5T_STRING:
        return np.dtype(f"S{size}")
    raise NotImplementedError(f"HDF5 type class {cls}")


def _describe(dt: np.dtype) -> str:
    if dt.names:
        return "compound"
    if dt.subdtype is not None:
        return f"{dt.subdtype[0].str}{list(dt.subdtype[1])}"
    return dt.str


def _read_obj(obj, t, space, reader):
    """Read attribute/dataset `obj` of type t and dataspace `space` with `reader`(obj, memtype, buf)."""
    lib = _load()
    npts = lib.H5Sget_simple_extent_npoints(_hid(space))
    scalar = lib.H5Sget_simple_extent_type(_hid(space)) == H5S_SCALAR
    nd = lib.H5Sget_simple_extent_ndims(_hid(space))
    dims = (C.c_uint64 * max(1, nd))()
    lib.H5Sget_simple_extent_dims(_hid(space), dims, None)
    shape = tuple(dims[:nd]) if nd > 0 else ()
    if lib.H5Tget_class(_hid(t)) == H5T_STRING and lib.H5Tis_variable_str(_hid(t)) > 0:
        mt = lib.H5Tcopy(_hid(_native("H5T_C_S1")))
        lib.H5Tset_size(_hid(mt), C.c_size_t(-1))  # H5T_VARIABLE
        ptrs = (C.c_char_p * max(1, npts))()
        reader(obj, mt, ptrs)
        lib.H5Tclose(_hid(mt))
        vals = [p.decode() for p in ptrs[:npts]]
        return (vals[0] if scalar else vals), "vlen-str", scalar, shape
    dt = _np_dtype(t)
    buf = np.zeros(max(1, npts) * dt.itemsize, np.uint8)
    reader(obj, t, buf.ctypes.data_as(C.c_void_p))  # the file type as memory type: no conversion
    arr = np.frombuffer(buf.tobytes(), dt, count=npts)
    if dt.kind == "S":
        vals = [bytes(v).split(b"\0")[0].decode() for v in arr]
        strpad = lib.H5Tget_strpad(_hid(t))
        return (vals[0] if scalar else vals), f"S{dt.itemsize}{'z' if strpad == 0 else ''}", scalar, shape
    val = arr[0] if scalar else arr.reshape(shape)
    return val, _describe(dt), scalar, shape


def _attrs(loc):
    lib = _load()
    names = []
    CB = C.CFUNCTYPE(C.c_int, _hid, C.c_char_p, C.c_void_p, C.c_void_p)
    cb = CB(lambda _l, n, _i, _d: names.append(n.decode()) or 0)
    idx = C.c_uint64(0)
    lib.H5Aiterate2(_hid(loc), 0, 0, C.byref(idx), cb, None)
    out = {}
    for n in names:
        a = lib.H5Aopen(_hid(loc), n.encode(), _hid(0))
        t, s = lib.H5Aget_type(_hid(a)), lib.H5Aget_space(_hid(a))
        out[n] = Attr(*_read_obj(a, t, s, lambda o, m, b: lib.H5Aread(_hid(o), _hid(m), b)))
        lib.H5Tclose(_hid(t)), lib.H5Sclose(_hid(s)), lib.H5Aclose(_hid(a))
    return out


def _group(g) -> Group:
    lib = _load()
    node = Group(attrs=_attrs(g))
    names = []
    CB = C.CFUNCTYPE(C.c_int, _hid, C.c_char_p, C.c_void_p, C.c_void_p)
    cb = CB(lambda _l, n, _i, _d: names.append(n.decode()) or 0)
    idx = C.c_uint64(0)
    lib.H5Literate(_hid(g), 0, 0, C.byref(idx), cb, None)
    for n in names:
        o = lib.H5Oopen(_hid(g), n.encode(), _hid(0))
        kind = lib.H5Iget_type(_hid(o))
        if kind == H5I_GROUP:
            node.groups[n] = _group(o)
        elif kind == H5I_DATASET:
            t, s, p = lib.H5Dget_type(_hid(o)), lib.H5Dget_space(_hid(o)), lib.H5Dget_create_plist(_hid(o))
            val, desc, _, _ = _read_obj(o, t, s, lambda d, m, b: lib.H5Dread(_hid(d), _hid(m), _hid(0), _hid(0),
                                                                              _hid(0), b))
            layout = lib.H5Pget_layout(_hid(p))
            chunk = ()
            if layout == H5D_CHUNKED:
                cd = (C.c_uint64 * 8)()
                nd = lib.H5Pget_chunk(_hid(p), 8, cd)
                chunk = tuple(cd[:nd])
            node.datasets[n] = Dataset(np.asarray(val), desc,
                                       {0: "compact", H5D_CONTIGUOUS: "contiguous", H5D_CHUNKED: "chunked"}[layout],
                                       chunk, lib.H5Pget_nfilters(_hid(p)), _attrs(o))
            lib.H5Pclose(_hid(p)), lib.H5Tclose(_hid(t)), lib.H5Sclose(_hid(s))
        lib.H5Oclose(_hid(o))
    return node


def walk(path: str) -> Group:
    lib = _load()
    f = lib.H5Fopen(path.encode(), 0, _hid(0))
    if f < 0:
        raise OSError(f"cannot open {path}")
    try:
        return _group(f)
    finally:
        lib.H5Fclose(_hid(f))


def coefficient_matrix(root: Group):
    """(dense A [num_vox x num_beamlets], col_idx) built as read_dose_data.data_mat_parser does:
    columns in increasing beam then beamlet number, rows = lindex; num_vox from the ROI
    capacities when the file is reduced, else from full_dicom_size."""
    specs = root.groups["calc_specs"].attrs
    if "roi_order" in specs:
        nvox = int(np.sum(specs["row_block_capacities"].value))
    else:
        nvox = int(np.prod(specs["full_dicom_size"].value))
    cols = []
    for bname in sorted(root.groups["beams"].groups["data"].groups, key=lambda s: int(s.split("_")[1])):
        bg = root.groups["beams"].groups["data"].groups[bname]
        for lname in sorted(bg.groups, key=lambda s: int(s.split("_")[1])):
            cols.append((int(bname.split("_")[1]), int(lname.split("_")[1]), bg.groups[lname]))
    A = np.zeros((nvox, len(cols)), np.float32)
    for j, (_, _, g) in enumerate(cols):
        A[g.datasets["lindex"].data.astype(np.int64), j] = g.datasets["coeffs"].data
    return A, [(b, l) for b, l, _ in cols]


def external_link(path: str, link: str):
    """(file name, object path) of the external link `link` (a path inside the file at `path`), or
    None when `link` is not an external link."""
    lib = _load()
    f = lib.H5Fopen(path.encode(), 0, _hid(0))
    if f < 0:
        raise OSError(f"cannot open {path}")
    try:
        info = (C.c_ubyte * 64)()
        if lib.H5Lget_info(_hid(f), link.encode(), info, _hid(0)) < 0:
            raise KeyError(link)
        # H5L_info_t: H5L_type_t type (int) first; H5L_TYPE_EXTERNAL = 64
        if C.cast(info, C.POINTER(C.c_int))[0] != 64:
            return None
        buf = (C.c_char * 4096)()
        if lib.H5Lget_val(_hid(f), link.encode(), buf, C.c_size_t(4096), _hid(0)) < 0:
            raise OSError("H5Lget_val failed")
        fn, on = C.c_char_p(), C.c_char_p()
        flags = C.c_uint(0)
        if lib.H5Lunpack_elink_val(buf, C.c_size_t(4096), C.byref(flags), C.byref(fn), C.byref(on)) < 0:
            raise OSError("H5Lunpack_elink_val failed")
        return fn.value.decode(), on.value.decode()
    finally:
        lib.H5Fclose(_hid(f))
