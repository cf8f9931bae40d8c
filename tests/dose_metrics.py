"""Dose comparison metrics used by the parity tests: max error in % of Dmax and a global gamma
pass rate (dose-difference / distance-to-agreement)."""
from __future__ import annotations

import numpy as np
from scipy.ndimage import map_coordinates


def max_err_pct(ref: np.ndarray, ev: np.ndarray) -> float:
    dmax = float(np.abs(ref).max())
    return 100.0 * float(np.abs(ev.astype(np.float64) - ref.astype(np.float64)).max()) / dmax if dmax > 0 else 0.0


def gamma_pass_rate(ref: np.ndarray, ev: np.ndarray, voxel_mm: float, dd: float = 0.01, dta_mm: float = 1.0,
                    cutoff: float = 0.10, step_mm: float = 0.25) -> float:
    """Fraction of reference voxels (dose >= cutoff*Dmax) with gamma <= 1.  The evaluated dose is
    searched trilinearly on a step_mm lattice inside the DTA sphere (global normalisation)."""
    ref = ref.astype(np.float64)
    ev = ev.astype(np.float64)
    dmax = ref.max()
    if dmax <= 0:
        return 1.0
    idx = np.argwhere(ref >= cutoff * dmax)
    if idx.size == 0:
        return 1.0
    rv = ref[tuple(idx.T)]
    g2 = np.full(rv.shape, np.inf)
    n = int(np.floor(dta_mm / step_mm))
    r = np.arange(-n, n + 1) * step_mm
    for oz in r:
        for oy in r:
            for ox in r:
                d2 = (ox * ox + oy * oy + oz * oz) / (dta_mm * dta_mm)
                if d2 > 1.0 + 1e-9:
                    continue
                coords = idx.T.astype(np.float64) + (np.array([oz, oy, ox]) / voxel_mm)[:, None]
                e = map_coordinates(ev, coords, order=1, mode="nearest")
                g2 = np.minimum(g2, ((e - rv) / (dd * dmax)) ** 2 + d2)
    return float(np.mean(g2 <= 1.0))
