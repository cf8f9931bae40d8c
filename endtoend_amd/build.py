"""Build the CCCS HIP library (endtoend_amd/lib/libcccs.so) for gfx950 with hipcc, the host-only
HDF5 file layer (endtoend_amd/lib/libcccs_h5.so) with g++ where HDF5 is installed, and the
reference-named drop-in shims over them (endtoend_amd/lib/libcccs_shim_{beam,beamlet}.so).

Every source is compiled with -ffp-contract=off (the Siddon segment lists must match the host
oracle bit for bit) and with packed FP32 disabled.  Objects go to endtoend_amd/_build/, the
shared library to endtoend_amd/lib/.  Usage: python -m endtoend_amd.build [--force]
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libcccs.so")

SOURCES = ["host.cpp", "io.cpp", "server.cpp", "tune.cpp", "dicom.cpp", "context.hip", "beam.hip", "beamlet.hip",
           "ingest.hip"]
HEADERS = ["cccs_internal.hpp", "context.hpp", "device_common.hpp", "errors.hpp"]

ARCH = os.environ.get("CCCS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]


def _stale(target: str, deps: list[str]) -> bool:
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def _compile(src: str, force: bool) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    deps = [path, os.path.join(ROOT, "include", "cccs.h")] + [os.path.join(CSRC, h) for h in HEADERS]
    if force or _stale(obj, deps):
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC] + COMMON + lang + ["-c", path, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


H5_LIB = os.path.join(LIBDIR, "libcccs_h5.so")
H5_SOURCE = os.path.join(CSRC, "h5io.cpp")


def hdf5_root() -> str | None:
    """Prefix holding include/hdf5.h and lib/libhdf5.so (CCCS_HDF5_ROOT, else /opt/conda or /usr)."""
    for root in [os.environ.get("CCCS_HDF5_ROOT"), "/opt/conda", "/usr"]:
        if root and os.path.exists(os.path.join(root, "include", "hdf5.h")) and \
                os.path.exists(os.path.join(root, "lib", "libhdf5.so")):
            return root
    return None


def build_h5(force: bool = True, verbose: bool = False) -> str | None:
    """Compile the host-only HDF5 file layer (libcccs_h5.so) against the HDF5 C library.

    libstdc++ is linked statically so that the HDF5 prefix's own C++ runtime is never picked up
    through the run path.  Returns None (and builds nothing) when no HDF5 installation is found."""
    root = hdf5_root()
    if root is None:
        if verbose:
            print("libcccs_h5.so skipped: no HDF5 headers/library (set CCCS_HDF5_ROOT)")
        return None
    deps = [H5_SOURCE, os.path.join(ROOT, "include", "cccs.h"), os.path.join(ROOT, "include", "cccs_h5.h")]
    if not force and not _stale(H5_LIB, deps):
        return H5_LIB
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = H5_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-shared", "-fPIC", "-static-libstdc++", "-static-libgcc",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(root, "include"), H5_SOURCE,
           "-L", os.path.join(root, "lib"), "-lhdf5", f"-Wl,-rpath,{os.path.join(root, 'lib')}", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for h5io.cpp:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, H5_LIB)
    if verbose:
        print(f"built {H5_LIB}")
    return H5_LIB


SHIM_DIR = os.path.join(HERE, "shim")
SHIM_LIBS = {"beam": os.path.join(LIBDIR, "libcccs_shim_beam.so"),
             "beamlet": os.path.join(LIBDIR, "libcccs_shim_beamlet.so")}


def build_shim(force: bool = True, verbose: bool = False) -> dict:
    """The reference-named drop-in (endtoend_amd/shim): libcccs_shim_beam.so (dosecalc-beam's entry
    points) over libcccs.so, and libcccs_shim_beamlet.so (dosecalc-beamlet's, with its
    SparsifyManager) over libcccs.so and libcccs_h5.so.  Host C++ only (g++); the beamlet shim is
    skipped when the HDF5 layer is not built."""
    src = os.path.join(SHIM_DIR, "reference_shim.cpp")
    deps = [src, os.path.join(SHIM_DIR, "reference_shim.hpp"), os.path.join(SHIM_DIR, "reference_types.hpp"),
            os.path.join(ROOT, "include", "cccs.h"), os.path.join(ROOT, "include", "cccs_h5.h"), LIB]
    built = {}
    for kind, out in SHIM_LIBS.items():
        if kind == "beamlet" and not os.path.exists(H5_LIB):
            continue
        if force or _stale(out, deps):
            tmp = out + ".tmp"
            cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-shared", "-fPIC", "-I", os.path.join(ROOT, "include"),
                   "-I", SHIM_DIR] + (["-DCCCS_SHIM_BEAMLET"] if kind == "beamlet" else []) + \
                  [src, "-L", LIBDIR, "-lcccs"] + (["-lcccs_h5"] if kind == "beamlet" else []) + \
                  ["-Wl,-rpath,$ORIGIN", "-Wl,--allow-shlib-undefined", "-o", tmp]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"g++ failed for the {kind} shim:\n{r.stdout}\n{r.stderr}")
            os.replace(tmp, out)
            if verbose:
                print(f"built {out}")
        built[kind] = out
    return built


def build(force: bool = True, verbose: bool = False) -> str:
    """Compile every source and link libcccs.so (and libcccs_h5.so); returns the HIP library path."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_h5(force=force, verbose=verbose)
    build_shim(force=force, verbose=verbose)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
