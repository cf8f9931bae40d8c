"""Regenerate the shipped kernel data file from the reference's spectrum and kernel tables.

Reads <cccs_data>/spectra/<name>.spec, <cccs_data>/dsa/<kernel files> and <cccs_data>/dsa/radii.dat
through the library's own parser (cccs_poly_kernel_from_files) and writes <name>.txt here:
the normalised spectrum, the 24 radial boundaries and the 48 x 24 poly-energetic total kernel,
every float printed with 9 significant digits (exact FP32 round trip).

    python -m endtoend_amd.data.make_tables /path/to/CCCS/data spec_6mv
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

from endtoend_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))


def make(data_dir: str, name: str) -> str:
    lib = abi.load()
    poly = np.zeros(48 * 24, np.float32)
    radii = np.zeros(24, np.float32)
    spec = np.zeros(4 * 30, np.float32)
    n = C.c_int(0)
    code = lib.cccs_poly_kernel_from_files(
        os.path.join(data_dir, "spectra", name + ".spec").encode(), os.path.join(data_dir, "dsa").encode(),
        os.path.join(data_dir, "dsa", "radii.dat").encode(), abi.fptr(poly), abi.fptr(radii), abi.fptr(spec),
        C.byref(n))
    abi.check(code, "cccs_poly_kernel_from_files")
    ns = n.value
    out = os.path.join(HERE, name + ".txt")
    with open(out, "w") as f:
        f.write(f"# poly-energetic CCCS kernel tables derived from the {name} spectrum\n")
        f.write(f"nspec {ns}\n# fluence(normalised) energy[MeV] mu_en[cm2/g] mu[cm2/g]\n")
        for e in range(ns):
            f.write(" ".join(f"{spec[k * ns + e]:.9g}" for k in range(4)) + "\n")
        f.write("radii 24\n" + " ".join(f"{r:.9g}" for r in radii) + "\n")
        f.write("poly 48 24\n")
        for t in range(48):
            f.write(" ".join(f"{v:.9g}" for v in poly[t * 24:(t + 1) * 24]) + "\n")
    return out


if __name__ == "__main__":
    print(make(sys.argv[1], sys.argv[2] if len(sys.argv) > 2 else "spec_6mv"))
