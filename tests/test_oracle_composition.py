"""Pins how the oracle composes its stages, which the per-stage tests leave open.

* Kernel row of a cone: the REV cone dose of every cone of a beam (beam path) and of the beamlet
  path equals a float64 walk of the same REV density and TERMA whose kernel rows come from the
  cone's polar angle (the row whose polar bin holds theta, and its mirror for the opposite
  direction), not from the oracle's ray-index arithmetic.
* Radii of the beamlet walk: that float64 walk steps on the midpoints of the radial bins
  (dosecalc-beamlet/nvbbRayConvolve_host.cu:351-357), computed here from the raw table.
* Placement of the beamlet unpack: packed grids filled with a linear function of the PVCS point
  the raytrace assigns to each node unpack to the same linear function of each voxel centre, so
  the unpack inverts the raytrace's map along every axis, depth included.

Each check is also run against an oracle built with the matching mistake (oracle/Makefile
`mutants`: the phi index as kernel row, raw radii in the beamlet walk, the unpack half a sample
off along the beamlet axis) and must fail there.
"""
import numpy as np
import pytest

from endtoend_amd import workloads as W
from oracle import oracle as O

from line_walk import brute_walk, mean_radii, row_of_theta


def _beam_case():
    # 6 azimuths x 4 polar rows: the theta index (ray % 4) and the phi index (ray // 4) differ on
    # most rays, and every polar row is used
    return W.small(n=20, voxel=0.5, gantry=30.0, fmap=(4, 4), nphi=6, ntheta=8, hetero=True)


def _beamlet_case():
    w = W.small(n=20, voxel=0.5, gantry=40.0, couch=20.0, fmap=(3, 3), nphi=6, ntheta=8, hetero=True)
    w.spec.kernel_extent = 1.0
    f = np.zeros((3, 3), np.float32)
    f[1, 1] = 1.0
    f[0, 2] = 0.7
    return w, f


def _lines(dens, terma, k=5):
    """Up to k (tZ, tY) lines with the most TERMA on valid samples."""
    nz, ny, _ = dens.shape
    score = np.where(dens >= 0, terma, 0.0).sum(axis=2)
    order = np.argsort(score.ravel())[::-1][:k]
    return [(int(i) // ny, int(i) % ny) for i in order if score.ravel()[i] > 0]


def check_beam_rows():
    w = _beam_case()
    spec, b, f = w.spec, w.beams[0], w.fmaps[0]
    terma, _ = O.terma(spec, b, f)
    code, g = O.geometry(spec, b, terma)
    assert code == 0
    ntheta, nphi = spec.ntheta, spec.nphi
    radii = np.asarray(spec.radii, np.float64)
    checked = 0
    for ray in range(g.nrays):
        rs = tuple(int(x) for x in g.rev_size[ray])
        d, t, o = O.cone_stage(spec, b, f, ray, rs)
        row = row_of_theta(spec.theta[ray], ntheta)
        for tz, ty in _lines(d, t, 3):
            want = brute_walk(d[tz, ty], t[tz, ty], spec.kernel, row, ntheta - 1 - row, spec.rev_longspacing,
                              nphi, radii, packed=False)
            np.testing.assert_allclose(o[tz, ty], want, rtol=1e-4, atol=1e-6 * np.abs(want).max())
            checked += 1
    assert checked >= 3 * g.nrays - 3


def check_beamlet_rows_and_radii():
    w, f = _beamlet_case()
    spec, b = w.spec, w.beams[0]
    ntheta, nphi = spec.ntheta, spec.nphi
    mr = mean_radii(spec.radii)
    checked = 0
    for ray in range(nphi * ntheta // 2):
        d, t, o = O.beamlet_cone_stage(spec, b, f, ray)
        row = row_of_theta(spec.theta[ray], ntheta)
        for tz, ty in _lines(d, t, 2):
            want = brute_walk(d[tz, ty], t[tz, ty], spec.kernel, row, ntheta - 1 - row, spec.rev_longspacing,
                              nphi, mr, packed=True)
            # FP32 sums of kernel differences: 2e-5 of the line maximum at worst on this case
            np.testing.assert_allclose(o[tz, ty], want, rtol=1e-3, atol=1e-4 * np.abs(want).max())
            checked += 1
    assert checked >= nphi * ntheta // 2


def check_beamlet_unpack_placement():
    w, f = _beamlet_case()
    spec = w.spec
    coef = np.array([100.0, 0.9, -1.3, 0.6], np.float32)
    cols = O.beamlet_unpack_linear(spec, w.beams[0], f, coef)
    assert cols.shape[0] == int((f > 0).sum())
    bx0, by0, bz0 = spec.bbox_start
    vx, vy, vz = spec.voxel
    sx, sy, sz = spec.start
    z, y, x = np.indices(spec.bbox_shape)
    gx, gy, gz = sx + (x + bx0) * vx, sy + (y + by0) * vy, sz + (z + bz0) * vz
    want = coef[0] + coef[1] * gx + coef[2] * gy + coef[3] * gz
    for c in cols:
        m = c != 0
        # the unpack writes every voxel of the pillar's interior: a slab around the beamlet axis
        assert m.sum() > 50
        np.testing.assert_allclose(c[m], want[m], rtol=0, atol=2e-3)


def test_beam_cone_rows_follow_the_cone_polar_angle():
    check_beam_rows()


def test_beamlet_cone_walks_use_mean_radii_and_theta_rows():
    check_beamlet_rows_and_radii()


def test_beamlet_unpack_inverts_the_raytrace_map():
    check_beamlet_unpack_placement()


@pytest.fixture(scope="module")
def mutants():
    O.build("mutants")
    return {k: O.mutant_path(k) for k in (1, 2, 3)}


@pytest.mark.parametrize("k,check", [(1, check_beam_rows), (1, check_beamlet_rows_and_radii),
                                     (2, check_beamlet_rows_and_radii), (3, check_beamlet_unpack_placement)])
def test_each_check_catches_its_planted_mistake(mutants, k, check):
    with O.using(mutants[k]):
        with pytest.raises(AssertionError):
            check()
