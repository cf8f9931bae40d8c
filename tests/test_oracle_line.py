"""Pins the collapsed-cone line walks of the oracle (A6 RowConvolve, A12 PackRowConvolve).

On a uniform line the walk telescopes: the sum of (K(r_k) - K(r_k+1)) * T equals T * K(r_last)
with K the piecewise-linear cumulative kernel (0 at r = 0).  On heterogeneous lines a float64
brute-force sum with radiological depths from prefix sums must agree.  A swapped kernel row,
a lost half-step or a miscounted self voxel fails these.
"""
import numpy as np
import pytest

from endtoend_amd import engine as E
from oracle import oracle as O

from line_walk import brute_walk, kcum, mean_radii

T = E.load_tables()
KERN, _, _ = E.cumulative_kernel(T, 8, 8)
R = T.radii.astype(np.float64)
MEAN_R = mean_radii(R)


def K(row, r, radii):
    return kcum(KERN[row], r, radii)


@pytest.mark.parametrize("P", [0, 1, 3])
def test_uniform_line_telescopes(P):
    n, rho, t, d, nphi = 60, 1.1, 0.37, 0.25, 8
    out = O.row_convolve(np.full(n, rho, np.float32), np.full(n, t, np.float32), KERN, T.radii, P, d, nphi)
    for i in range(n):
        want = 0.0
        if i >= 1:
            want += t * K(P, (i - 0.5) * rho * d, R)
        if i <= n - 2:
            want += t * K(8 - 1 - P, (n - 1 - i - 0.5) * rho * d, R)
        np.testing.assert_allclose(out[i], want / nphi, rtol=2e-5, atol=1e-9)


def _brute(dens, terma, P, d, nphi, radii, packed):
    return brute_walk(dens, terma, KERN, P, KERN.shape[0] - 1 - P, d, nphi, radii, packed)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_heterogeneous_line_matches_float64_walk(seed):
    rng = np.random.default_rng(seed)
    n, d, nphi = 90, 0.3, 6
    dens = rng.uniform(0.05, 2.0, n).astype(np.float32)
    terma = rng.uniform(0.0, 1.0, n).astype(np.float32)
    P = int(rng.integers(0, 4))
    out = O.row_convolve(dens, terma, KERN, T.radii, P, d, nphi)
    want = _brute(dens.astype(np.float64), terma.astype(np.float64), P, d, nphi, R, packed=False)
    np.testing.assert_allclose(out, want, rtol=1e-4, atol=1e-6 * np.abs(want).max())


def test_walk_stops_at_kernel_reach():
    # scaled-down radii: the walk must stop once the radiological distance exceeds the last radius
    rng = np.random.default_rng(9)
    n, d, nphi = 70, 0.5, 4
    dens = rng.uniform(0.5, 1.5, n).astype(np.float32)
    terma = rng.uniform(0.0, 1.0, n).astype(np.float32)
    radii = (T.radii * np.float32(0.1)).astype(np.float32)  # reach 6 cm  < line of 35 cm
    out = O.row_convolve(dens, terma, KERN, radii, 2, d, nphi)
    want = _brute(dens.astype(np.float64), terma.astype(np.float64), 2, d, nphi, radii.astype(np.float64), False)
    np.testing.assert_allclose(out, want, rtol=1e-4, atol=1e-6 * np.abs(want).max())


def test_invalid_tail_contributes_nothing():
    # samples outside the crop box / calc_bbox carry density -3/-2 and TERMA 0; walking through
    # them (as the reference does) must leave every valid destination unchanged
    rng = np.random.default_rng(4)
    n, d, nphi = 50, 0.25, 8
    dens = rng.uniform(0.3, 1.7, n).astype(np.float32)
    terma = rng.uniform(0.0, 1.0, n).astype(np.float32)
    core = O.row_convolve(dens, terma, KERN, T.radii, 1, d, nphi)
    pad_d = np.concatenate([np.full(7, -3.0), dens, np.full(9, -2.0)]).astype(np.float32)
    pad_t = np.concatenate([np.zeros(7), terma, np.zeros(9)]).astype(np.float32)
    full = O.row_convolve(pad_d, pad_t, KERN, T.radii, 1, d, nphi)
    np.testing.assert_array_equal(full[7:7 + n], core)
    assert np.all(full[:7] == 0) and np.all(full[7 + n:] == 0)


@pytest.mark.parametrize("P", [0, 2])
def test_packed_uniform_line_telescopes(P):
    # PackRowConvolve counts the destination voxel itself in both directions
    n, rho, t, d, nphi = 40, 0.9, 0.5, 0.25, 8
    out = O.row_convolve(np.full(n, rho, np.float32), np.full(n, t, np.float32), KERN, MEAN_R.astype(np.float32), P,
                         d, nphi, packed=True)
    for i in range(n):
        rf = i * rho * d if i >= 1 else 0.5 * rho * d
        rb = (n - 1 - i) * rho * d if i <= n - 2 else 0.5 * rho * d
        want = t * (K(P, rf, MEAN_R) + K(8 - 1 - P, rb, MEAN_R))
        np.testing.assert_allclose(out[i], want / nphi, rtol=2e-5, atol=1e-9)


@pytest.mark.parametrize("seed", [5, 6])
def test_packed_line_stops_at_walls(seed):
    rng = np.random.default_rng(seed)
    n, d, nphi = 80, 0.25, 8
    dens = rng.uniform(0.2, 1.8, n).astype(np.float32)
    dens[[10, 11, 45, 46]] = -1e-12  # pillar walls
    terma = rng.uniform(0.0, 1.0, n).astype(np.float32)
    terma[dens < 0] = 0.0
    out = O.row_convolve(dens, terma, KERN, MEAN_R.astype(np.float32), 1, d, nphi, packed=True)
    want = _brute(dens.astype(np.float64), terma.astype(np.float64), 1, d, nphi, MEAN_R, packed=True)
    np.testing.assert_allclose(out, want, rtol=1e-4, atol=1e-6 * np.abs(want).max())
    # segments between walls are independent: changing one does not touch the others
    terma2 = terma.copy()
    terma2[20:40] *= 3.0
    out2 = O.row_convolve(dens, terma2, KERN, MEAN_R.astype(np.float32), 1, d, nphi, packed=True)
    np.testing.assert_array_equal(out2[:12], out[:12])
    np.testing.assert_array_equal(out2[45:], out[45:])
