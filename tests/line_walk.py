"""Float64 collapsed-cone line walks written from the method, for pinning the oracle's walks.

The cumulative kernel of a row is piecewise linear in the radius between the table's radii and 0
at r = 0.  Radiological distances come from running sums of 0.5 * delr * density half steps, in
float64.  The walk towards lower sample indices takes kernel row `row_lo`, the walk towards higher
indices row `row_hi`.  In the full-beam walk the destination sample itself contributes nothing
and invalid samples (density < 0) carry no TERMA and add no distance; in the packed (pillar) walk
the destination contributes its first half step in both directions and a wall (density < 0) ends
the walk.
"""
import numpy as np


def kcum(krow, r, radii):
    """Piecewise-linear cumulative kernel of one row at radius r (K(0) = 0)."""
    k = np.asarray(krow, np.float64)
    if r <= radii[0]:
        return r / radii[0] * k[0]
    j = int(np.searchsorted(radii, r))
    j = min(j, len(radii) - 1)
    return k[j - 1] + (r - radii[j - 1]) / (radii[j] - radii[j - 1]) * (k[j] - k[j - 1])


def brute_walk(dens, terma, kern, row_lo, row_hi, d, nphi, radii, packed):
    """Dose of every sample of one line; kern is the (ntheta, nradii) cumulative table."""
    dens = np.asarray(dens, np.float64)
    terma = np.asarray(terma, np.float64)
    radii = np.asarray(radii, np.float64)
    n = len(dens)
    rmax = radii[-1]
    out = np.zeros(n)
    for i in range(n):
        if dens[i] < 0:
            continue
        total = 0.0
        for row, step in ((row_lo, -1), (row_hi, +1)):
            last = 0.0
            ks = range(i if packed else i + step, -1 if step < 0 else n, step)
            acc = 0.0  # radiological distance from the destination's reference point to the far edge
            for k in ks:
                if dens[k] < 0:
                    if packed:
                        break
                    continue
                if packed:
                    r = 0.5 * dens[i] * d if k == i else acc + 0.5 * dens[k] * d
                else:
                    r = acc + 0.5 * dens[k] * d
                if r > rmax:
                    break
                c = kcum(kern[row], r, radii)
                total += (c - last) * terma[k]
                last = c
                if not (packed and k == i):
                    acc = r + 0.5 * dens[k] * d
                else:
                    acc = 0.5 * dens[i] * d
                if r > radii[-2] and np.searchsorted(radii, r) >= len(radii) - 1:
                    break
        out[i] = total / nphi
    return out


def row_of_theta(theta_deg, ntheta):
    """Kernel row whose polar bin [t, t + 1) * 180 / ntheta holds a cone at polar angle theta."""
    return int(np.floor(float(theta_deg) * ntheta / 180.0))


def mean_radii(radii):
    """Midpoints of the radial bins [0, r0], [r0, r1], ...: the radii the pillar walk steps on."""
    r = np.asarray(radii, np.float64)
    return np.concatenate([[0.5 * r[0]], 0.5 * (r[1:] + r[:-1])])
