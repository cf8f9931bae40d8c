"""The C-ABI library loads and exports every entry point declared in include/cccs.h; the host-only
entry points work without a GPU; device entry points fail cleanly when no device is present."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from endtoend_amd import abi
from endtoend_amd import engine as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "cccs.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cccs_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(abi.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    # the Python mirror covers the whole header too
    mirrored = {n for n, _, _ in abi.EXPORTS}
    assert set(names) <= mirrored, sorted(set(names) - mirrored)


def test_struct_layouts_match_header():
    # sizes of the POD structs as laid out by the C compiler (x86-64 SysV)
    assert C.sizeof(abi.Beam) == 4 * (3 + 3 + 3 + 3 + 1 + 2 + 2)
    assert C.sizeof(abi.BeamGeometry) == 4 * (3 + 3 + 3 + 3 + 1 + 48 * 3 + 48 * 3)
    assert C.sizeof(abi.Problem) % 8 == 0


def test_reconfigure_gantry0_known_answer():
    b = E.make_beam(0.0, 0.0, 0.0, iso=(1.0, 2.0, 3.0), sad=100.0)
    # gantry 0: the source sits SAD above the isocentre along -y and the beam points along +y
    np.testing.assert_allclose(list(b.source), [1.0, -98.0, 3.0], atol=1e-5)
    np.testing.assert_allclose(list(b.direction), [0.0, 1.0, 0.0], atol=1e-6)


def test_cumulative_kernel_host_entry():
    t = E.load_tables()
    k, th, ph = E.cumulative_kernel(t, 8, 8)
    assert k.shape == (8, 24) and th.shape == (32,) and ph.shape == (32,)
    assert np.all(np.diff(k, axis=1) >= 0)


def test_bad_arguments_are_rejected_without_a_device():
    lib = abi.load()
    t = E.load_tables()
    k = np.zeros(8 * 24, np.float32)
    # ntheta must divide 48 and be even
    assert lib.cccs_cumulative_kernel(abi.fptr(t.poly.ravel().copy()), 8, 7, abi.fptr(k), None, None) == abi.CCCS_EINVAL
    assert b"48" in lib.cccs_last_error()
    assert lib.cccs_beam_reconfigure(None, 0) == abi.CCCS_EINVAL
    h = C.c_void_p()
    assert lib.cccs_ctx_create(None, 0, C.byref(h)) == abi.CCCS_EINVAL


def test_ctx_create_without_device_reports_edevice():
    from endtoend_amd import workloads as W
    w = W.small(n=8, voxel=1.0, fmap=(4, 4))
    p = w.spec.struct()
    lib = abi.load()
    h = C.c_void_p()
    code = lib.cccs_ctx_create(C.byref(p), 0, C.byref(h))
    if code == abi.CCCS_OK:
        lib.cccs_ctx_destroy(h)
        pytest.skip("a HIP device is present")
    assert code == abi.CCCS_EDEVICE
    assert not h.value


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(FileNotFoundError):
        abi.load(str(tmp_path / "libcccs.so"))
