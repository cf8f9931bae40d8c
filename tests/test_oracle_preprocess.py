"""Preprocess inputs (the first 'next' row of the scope table): the structure projection onto
fluence maps (findFluenceProjection + cudaRaycast, dosecalc-preprocess/cudaMemFunctions.cu:308-393,
kernels.cuh:136-224), the aperture-ready fill (fmapProcessing.cpp:7-46) and the omni beam-list
writer (DoseCalcIO/beam.cpp:577-607).

The oracle's projection is pinned by geometry the raycast cannot escape:
* a voxelised ellipsoid projects inside the analytic projection of the ellipsoid grown by the
  sampling reach (trilinear support of one voxel plus the half-voxel fetch shift, under 3 voxels)
  and contains the analytic projection of the ellipsoid shrunk by the same reach;
* a one-voxel rod on the beam axis lights exactly the beamlets whose nine sub-rays reach it: the
  central 2x2 of an even map (through their corner sub-rays) and the central one of an odd map;
* mirrored beams on a mirror-symmetric mask give mirrored maps about the plane the half-voxel
  fetch shift makes the mask symmetric about, and not about the mask's own plane.
"""
import numpy as np
import pytest

from endtoend_amd import engine as E
from endtoend_amd import workloads as W
from oracle import oracle as O


def _spec(ph):
    return W.problem(ph, nphi=4, ntheta=4, max_rev_size=(800, 800, 800))


def _ellipsoid_case(n=40, voxel=0.4, semi=(3.0, 2.0, 1.5), centre=(0.4, -0.3, 0.2)):
    ph = W.water_cube(n, voxel)
    ph.ptv = (centre, semi)
    return ph, W.ptv_mask(ph).astype(np.uint8)


@pytest.mark.parametrize("g,c,coll", [(0.0, 0.0, 0.0), (33.0, 0.0, 0.0), (90.0, 20.0, 0.0), (215.0, -40.0, 15.0),
                                      (300.0, 70.0, -30.0)])
def test_projection_between_grown_and_shrunk_ellipsoid(g, c, coll):
    ph, mask = _ellipsoid_case()
    spec = _spec(ph)
    fmap, bsz = (24, 20), (0.5, 0.5)
    b = E.make_beam(g, c, coll, iso=(0.0, 0.0, 0.0), fmap_size=fmap, beamlet_size=bsz)
    got = O.fluence_projection(spec, mask, [b])[0]
    reach = 3.0 * ph.voxel
    centre, semi = ph.ptv
    inner = W.projection_mask(b, (centre, tuple(s - reach for s in semi)), fmap, bsz)
    outer = W.projection_mask(b, (centre, tuple(s + reach for s in semi)), fmap, bsz)
    assert inner.sum() > 10
    assert np.all(got >= inner), "a beamlet through the shrunk ellipsoid was missed"
    assert np.all(got <= outer), "a beamlet outside the grown ellipsoid was lit"
    assert set(np.unique(got)) <= {0.0, 1.0}


@pytest.mark.parametrize("fmap,expect", [((4, 4), [(1, 1), (1, 2), (2, 1), (2, 2)]), ((5, 5), [(2, 2)])])
def test_axis_rod_lights_the_subray_pattern(fmap, expect):
    """A rod along the beam axis (gantry 0: the y axis) one voxel thick: with 2 cm beamlets only
    sub-rays that pass through the axis reach it."""
    n, voxel = 40, 0.25
    ph = W.water_cube(n, voxel)
    mask = np.zeros((n, n, n), np.uint8)
    x, y, z = ph.coords()
    ix, iz = int(np.argmin(np.abs(x))), int(np.argmin(np.abs(z)))
    mask[iz, :, ix] = 1
    b = E.make_beam(0.0, 0.0, 0.0, iso=(float(x[ix]), 0.0, float(z[iz])), fmap_size=fmap, beamlet_size=(2.0, 2.0))
    got = O.fluence_projection(_spec(ph), mask, [b])[0]
    lit = sorted(zip(*np.nonzero(got)))
    assert lit == expect


def test_mirrored_beams_give_mirrored_maps():
    """The raycast samples the mask without the +0.5 texel shift, so the structure it sees is the
    mask moved by half a voxel along +x, +y, +z.  A mask symmetric under x -> -x is therefore seen
    symmetric about x = +voxel/2: gantry g and 360-g about an isocentre on that plane are mirror
    images and give mirrored maps (up to boundary beamlets flipped by the float rotation).  About
    x = 0 they are not mirrored, which pins the shift."""
    ph, mask = _ellipsoid_case(n=40, voxel=0.4, semi=(3.0, 2.0, 1.5), centre=(0.0, 0.0, 0.3))
    assert np.array_equal(mask, mask[:, :, ::-1])
    spec = _spec(ph)
    shifted_diff = 0
    for g in (25.0, 70.0, 130.0):
        for iso_x, tol in ((0.5 * ph.voxel, 2), (0.0, None)):
            a = E.make_beam(g, 0.0, 0.0, iso=(iso_x, 0.0, 0.0), fmap_size=(16, 12), beamlet_size=(0.5, 0.5))
            b = E.make_beam(360.0 - g, 0.0, 0.0, iso=(iso_x, 0.0, 0.0), fmap_size=(16, 12), beamlet_size=(0.5, 0.5))
            ma, mb = O.fluence_projection(spec, mask, [a, b])
            assert ma.sum() > 10
            d = int(np.abs(ma - mb[:, ::-1]).sum())
            if tol is not None:
                assert d <= tol
            else:
                shifted_diff += d
    assert shifted_diff > 6


def test_maps_of_several_beams_come_back_to_back():
    ph, mask = _ellipsoid_case()
    spec = _spec(ph)
    beams = [E.make_beam(g, 0.0, 0.0, fmap_size=(fx, fy), beamlet_size=(0.5, 0.5))
             for g, fx, fy in ((0.0, 10, 6), (90.0, 7, 9), (45.0, 12, 12))]
    together = O.fluence_projection(spec, mask, beams)
    alone = [O.fluence_projection(spec, mask, [b])[0] for b in beams]
    for t, a, b in zip(together, alone, beams):
        assert t.shape == (b.fmap_size[1], b.fmap_size[0])
        assert np.array_equal(t, a)


# ---- aperture-ready fill (host functions of the library) -------------------------------------------
def test_aperture_fill_reproduces_the_run_bookkeeping():
    f = np.array([[1, 1, 0, 0, 1, 1, 1, 0],
                  [0.5, 0, 0.2, 0.3, 0.4, 0, 0, 0],
                  [0, 0, 0, 2, 0, 0, 0, 0],
                  [1, 0, 1, 0, 1, 0, 0, 1]], np.float32)
    out = E.fmap_aperture_fill(f, fill_val=7.0)
    # a run is recorded by its last non-zero beamlet: the fill covers the gap and all but the last
    # beamlet of the next run
    assert out[0].tolist() == [1, 1, 7, 7, 7, 7, 1, 0]
    assert np.allclose(out[1], [0.5, 7, 7, 7, 0.4, 0, 0, 0])
    assert out[2].tolist() == f[2].tolist()
    assert out[3].tolist() == [1, 7, 1, 7, 1, 7, 7, 1]
    assert E.fmap_is_aperture_ready(out[2:3])
    assert not E.fmap_is_aperture_ready(f)
    assert E.fmap_is_aperture_ready(E.fmap_aperture_fill(f))
    assert E.fmap_is_aperture_ready(np.zeros((3, 4), np.float32))


# ---- omni beam list writer ----------------------------------------------------------------------------
def test_omni_beam_list_writer_round_trip(tmp_path):
    beams = [E.make_beam(g, c, coll, iso=(1.0, -2.0, 0.5), fmap_size=(fx, fy), beamlet_size=(0.5, 0.25))
             for g, c, coll, fx, fy in ((0.0, 0.0, 0.0, 40, 40), (123.5, 17.0, 5.0, 20, 32), (270.0, -45.0, 0.0, 8, 8))]
    path = str(tmp_path / "omni_beam_list.txt")
    E.write_omni_beam_list(path, beams, iso_type=["ptv", "manual", "unspec"], iso_loc=["in", "out", "unspec"])
    text = open(path).read().splitlines()
    assert text[0] == "3"
    assert text[1].split()[21] == "fmap-000000.raw" and text[3].split()[21] == "fmap-000002.raw"
    assert text[2].split()[5:7] == ["manual", "out"] and text[2].split()[13] == "auto"
    assert text[-1].startswith("#beam_id")
    got, names = E.load_omni_beam_list(path, 3)
    assert names == ["fmap-000000.raw", "fmap-000001.raw", "fmap-000002.raw"]
    for a, b in zip(beams, got):
        for f in ("azimuth", "zenith", "coll"):
            assert abs(getattr(a, f) - getattr(b, f)) < 1e-5
        assert list(a.fmap_size) == list(b.fmap_size)
        np.testing.assert_allclose(a.beamlet_size[:], b.beamlet_size[:])
        np.testing.assert_allclose(a.isocenter[:], b.isocenter[:], atol=1e-6)
        np.testing.assert_allclose(a.source[:], b.source[:], atol=1e-4)
