"""The reference-named drop-in shim (endtoend_amd/shim): both libraries build and export the
reference's in-process entry points by their own C++ names and signatures
(dosecalc-beam/nvbbRayConvolve_host.cu:158-169,296-308,745-781;
dosecalc-beamlet/nvbbRayConvolve_host.cu:283-294,444-456,1412-1413; the SparsifyManager calls of
dosecalc-beamlet/main.cu:462-519); the re-declared reference types keep the reference's data
members in order; and the test program fails cleanly without a GPU."""
import os
import re
import subprocess

import pytest

from endtoend_amd import build as B
from endtoend_amd import workloads as W

from tests.shim_case import DRIVERS, write_case

REF = "/root/reference/CCCS"
TYPES = os.path.join(B.SHIM_DIR, "reference_types.hpp")

BEAM_SIGS = [
    "initCudaConstandTex(SHM_DATA*, MONO_KERNELS*, CONSTANTS*, int, int, int, bool, bool, bool)",
    "radconvolveTexture(MONO_KERNELS*, CONSTANTS*, BEAM*, int, int, int, int, bool, bool, bool)",
    "freeCudaTexture(int, int*)",
    "cuda_sum_device_float_dose(float*, int, CONSTANTS*, bool, bool)",
]
BEAMLET_SIGS = [
    "initCudaConstandTex(SHM_DATA*, MONO_KERNELS*, CONSTANTS*, int, int, int, bool, bool, bool)",
    "radconvolveTexture(MONO_KERNELS*, CONSTANTS*, BEAM*, int, int, int, int, int, bool, bool, bool)",
    "freeCudaTexture(int, int*, bool)",
    "SparsifyManager::activate()",
    "SparsifyManager::deactivate()",
]


def _exports(lib):
    r = subprocess.run(["nm", "-D", "-C", "--defined-only", lib], capture_output=True, text=True, check=True)
    return {line.split(" ", 2)[2] for line in r.stdout.splitlines() if line.count(" ") >= 2}


@pytest.mark.parametrize("kind,sigs", [("beam", BEAM_SIGS), ("beamlet", BEAMLET_SIGS)])
def test_shim_exports_the_reference_entry_points(kind, sigs):
    lib = B.SHIM_LIBS[kind]
    if kind == "beamlet" and B.hdf5_root() is None:
        pytest.skip("no HDF5 C library")
    assert os.path.exists(lib), "run the build first"
    ex = _exports(lib)
    missing = [s for s in sigs if s not in ex]
    assert not missing, missing
    if kind == "beamlet":
        assert any(s.startswith("SparsifyManager::init(") and "HEADER_PATIENT&" in s and "WRITE_STRATEGY" in s
                   for s in ex)


def _members(text, name):
    """Data member names of struct/class `name`, in order (declarations of the form `type name;`
    or `type name[...];` or `type name = ...;` / `type name{...};`, functions and enums skipped)."""
    m = re.search(r"(?:struct|class)\s+" + name + r"\b[^;{]*\{", text)
    assert m, name
    i, depth = m.end(), 1
    body = []
    while depth:
        c = text[i]
        depth += c == "{"
        depth -= c == "}"
        body.append(c)
        i += 1
    body = "".join(body[:-1])
    body = re.sub(r"//[^\n]*|/\*.*?\*/", "", body, flags=re.S)
    body = re.sub(r"\b(public|private|protected)\s*:", ";", body)
    # nested braces (inline function bodies, enum lists, default initialisers) end a statement
    while True:
        b2 = re.sub(r"\{[^{}]*\}", ";", body)
        if b2 == body:
            break
        body = b2
    out = []
    for st in body.split(";"):
        st = " ".join(st.split())
        if not st or "(" in st or st.startswith(("enum", "friend", "static", "typedef", "struct", "class")):
            continue
        st = re.sub(r"=.*$", "", st).strip()
        st = re.sub(r"\[.*?\]", "", st).strip()
        out.append(st.split()[-1].lstrip("*&"))
    return out


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
@pytest.mark.parametrize("name,header", [("CONSTANTS", "include/server/brain_defs.h"),
                                         ("SHM_DATA", "include/server/brain_defs.h"),
                                         ("BEAM", "DoseCalcIO/beam.h"),
                                         ("KERNEL", "DoseCalcIO/kernel.h"),
                                         ("MONO_KERNELS", "DoseCalcIO/kernel.h"),
                                         ("HEADER_PATIENT", "DoseCalcIO/io_data_structs.h")])
def test_redeclared_types_keep_the_reference_members_in_order(name, header):
    ref = _members(open(os.path.join(REF, header)).read(), name)
    ours = _members(open(TYPES).read(), name)
    assert ours == ref


def test_driver_fails_cleanly_without_a_device(tmp_path):
    w = W.small(n=16, voxel=0.5, gantry=30.0, fmap=(4, 4))
    write_case(str(tmp_path), w.spec, w.beams, w.fmaps)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([DRIVERS["beam"], str(tmp_path), str(tmp_path / "dose.raw")], capture_output=True, text=True,
                       env=env, timeout=120)
    assert r.returncode == 1
    assert "initCudaConstandTex" in r.stderr
