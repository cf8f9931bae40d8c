"""A DICOM Part 10 writer for the tests (test infrastructure), written from the published encoding
(DICOM PS3.5 sections 6-7, PS3.10 section 7): CT image slices and an RT Structure Set in explicit
or implicit VR little endian, with sequences and items of defined or undefined length.  Only what
the ingestion tests need; values are encoded as the standard's VRs prescribe (text padded to even
length, UI with NUL, US as little-endian uint16, IS/DS as backslash-separated decimal strings)."""
from __future__ import annotations

import struct

import numpy as np

EXPLICIT, IMPLICIT = "1.2.840.10008.1.2.1", "1.2.840.10008.1.2"
CT_SOP, RTSTRUCT_SOP = "1.2.840.10008.5.1.4.1.1.2", "1.2.840.10008.5.1.4.1.1.481.3"
_LONG = {"OB", "OD", "OF", "OL", "OW", "SQ", "UC", "UR", "UT", "UN", "OV", "SV", "UV"}


def _pad(b: bytes, vr: str) -> bytes:
    if len(b) % 2:
        b += b"\0" if vr in ("UI", "OB") else b" "
    return b


def value(vr: str, v) -> bytes:
    if vr == "US":
        return b"".join(struct.pack("<H", int(x)) for x in np.atleast_1d(v))
    if vr in ("IS", "DS"):
        vals = np.atleast_1d(v)
        if vr == "IS":
            s = "\\".join(str(int(x)) for x in vals)
        else:
            s = "\\".join(ds(x) for x in vals)
        return _pad(s.encode(), vr)
    if vr == "OW":
        return bytes(v)
    return _pad(str(v).encode(), vr)


def ds(x: float) -> str:
    """A decimal string of at most 16 characters."""
    for prec in (10, 9, 8, 7, 6):
        s = f"{float(x):.{prec}g}"
        if len(s) <= 16:
            return s
    return f"{float(x):.5g}"


def element(tag: int, vr: str, payload: bytes, explicit: bool) -> bytes:
    g, e = tag >> 16, tag & 0xFFFF
    head = struct.pack("<HH", g, e)
    if explicit or g == 0x0002:
        if vr in _LONG:
            return head + vr.encode() + b"\0\0" + struct.pack("<I", len(payload)) + payload
        return head + vr.encode() + struct.pack("<H", len(payload)) + payload
    return head + struct.pack("<I", len(payload)) + payload


def dataset(elems, explicit: bool) -> bytes:
    """elems: list of (tag, vr, value) or (tag, 'SQ', [items], undefined_seq, undefined_items)."""
    out = b""
    for el in sorted(elems, key=lambda x: x[0]):
        tag, vr = el[0], el[1]
        if vr == "SQ":
            items, und_seq, und_item = el[2], el[3], el[4]
            body = b""
            for it in items:
                ib = dataset(it, explicit)
                if und_item:
                    body += struct.pack("<HHI", 0xFFFE, 0xE000, 0xFFFFFFFF) + ib + struct.pack("<HHI", 0xFFFE, 0xE00D, 0)
                else:
                    body += struct.pack("<HHI", 0xFFFE, 0xE000, len(ib)) + ib
            g, e = tag >> 16, tag & 0xFFFF
            if und_seq:
                length, tail = 0xFFFFFFFF, struct.pack("<HHI", 0xFFFE, 0xE0DD, 0)
            else:
                length, tail = len(body), b""
            if explicit:
                out += struct.pack("<HH", g, e) + b"SQ\0\0" + struct.pack("<I", length) + body + tail
            else:
                out += struct.pack("<HHI", g, e, length) + body + tail
        else:
            out += element(tag, vr, value(vr, el[2]), explicit)
    return out


def write(path: str, sop_class: str, sop_instance: str, elems, transfer: str = EXPLICIT) -> None:
    explicit = transfer == EXPLICIT
    meta = (element(0x00020001, "OB", b"\0\1", True) + element(0x00020002, "UI", value("UI", sop_class), True) +
            element(0x00020003, "UI", value("UI", sop_instance), True) +
            element(0x00020010, "UI", value("UI", transfer), True) +
            element(0x00020012, "UI", value("UI", "1.2.3.4.5.6"), True))
    meta = element(0x00020000, "UL", struct.pack("<I", len(meta)), True) + meta
    body = dataset(list(elems) + [(0x00080016, "UI", sop_class), (0x00080018, "UI", sop_instance)], explicit)
    with open(path, "wb") as f:
        f.write(b"\0" * 128 + b"DICM" + meta + body)


def write_ct_series(d: str, pixels: np.ndarray, spacing_mm, z_mm, origin_xy_mm, slope=1.0, intercept=-1024.0,
                    signed=True, transfer: str = EXPLICIT, names=None, thickness=None):
    """pixels [z, y, x] stored values (int16 when signed, else uint16), one file per slice."""
    nz, ny, nx = pixels.shape
    for k in range(nz):
        name = names[k] if names else f"ct_{k:03d}.dcm"
        px = pixels[k].astype(np.int16 if signed else np.uint16)
        elems = [(0x00080060, "CS", "CT"), (0x00200013, "IS", k + 1),
                 (0x00200032, "DS", [origin_xy_mm[0], origin_xy_mm[1], z_mm[k]]),
                 (0x00200037, "DS", [1, 0, 0, 0, 1, 0]), (0x00280002, "US", 1), (0x00280010, "US", ny),
                 (0x00280011, "US", nx), (0x00280030, "DS", list(spacing_mm)), (0x00280100, "US", 16),
                 (0x00280101, "US", 16), (0x00280102, "US", 15), (0x00280103, "US", 1 if signed else 0),
                 (0x00281052, "DS", intercept), (0x00281053, "DS", slope),
                 (0x7FE00010, "OW", px.tobytes())]
        if thickness is not None:
            elems.append((0x00180050, "DS", thickness))
        write(f"{d}/{name}", CT_SOP, f"1.2.826.0.1.{k + 1}", elems, transfer)


def write_rtstruct(path: str, rois, transfer: str = EXPLICIT, undefined: bool = False):
    """rois: list of (number, name, [contour points [n, 3] in mm, ...])."""
    roiseq, cseq = [], []
    for num, name, contours in rois:
        roiseq.append([(0x30060022, "IS", num), (0x30060026, "LO", name)])
        items = []
        for c in contours:
            c = np.asarray(c, np.float64).reshape(-1, 3)
            items.append([(0x30060042, "CS", "CLOSED_PLANAR"), (0x30060046, "IS", len(c)),
                          (0x30060050, "DS", c.reshape(-1))])
        cseq.append([(0x3006002A, "IS", [255, 0, 0]), (0x30060084, "IS", num),
                     (0x30060040, "SQ", items, undefined, undefined)])
    elems = [(0x00080060, "CS", "RTSTRUCT"), (0x30060020, "SQ", roiseq, undefined, undefined),
             (0x30060039, "SQ", cseq, undefined, not undefined)]
    write(path, RTSTRUCT_SOP, "1.2.826.0.2.1", elems, transfer)
