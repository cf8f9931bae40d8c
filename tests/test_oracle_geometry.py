"""Pins the oracle's beam geometry: crop box (A3), REV grids (A4), PVCS->REV resampling (A5) and
REV->PVCS resampling (A7), with rotations recomputed here in float64 (A16).

Trilinear interpolation reproduces linear fields exactly, so a linear density (or REV dose)
resampled through the oracle must equal the linear function evaluated at the sample position
computed independently here; a transposed axis, a missing +0.5 or a wrong rotation order fails.
"""
import math

import numpy as np
import pytest

from endtoend_amd import engine as E
from endtoend_amd import workloads as W
from oracle import oracle as O


def _rot(axis, t):
    x, y, z = axis
    s, c = math.sin(t), math.cos(t)
    k = 1 - c
    return np.array([[c + k * x * x, k * x * y - s * z, k * x * z + s * y],
                     [k * y * x + s * z, c + k * y * y, k * y * z - s * x],
                     [k * z * x - s * y, k * z * y + s * x, c + k * z * z]])


def beam_fwd(th, ph, coll):  # CudaUtilities/geometry.cu:69-75
    return _rot((0, 1, 0), ph + coll) @ _rot((math.sin(-ph), 0, math.cos(-ph)), -th)


def kernel_fwd(th, ph):      # geometry.cu:126-131
    return _rot((0, 1, 0), -ph) @ _rot((math.cos(ph), 0, -math.sin(ph)), -th)


def _spec_with(n, v, dens=None, theta=None, phi=None, penumbra=1.0):
    ph = W.Phantom(np.full((n, n, n), 1.0, np.float32) if dens is None else dens.astype(np.float32), v)
    spec = W.problem(ph, nphi=1, ntheta=2, penumbra=penumbra)
    if theta is not None:
        spec.theta = np.asarray(theta, np.float32)
        spec.phi = np.asarray(phi, np.float32)
        spec.nphi = len(theta)
    return spec


@pytest.mark.parametrize("gantry", [0.0, 90.0])
def test_crop_box_single_voxel(gantry):
    n, v, pen = 16, 0.5, 1.0
    spec = _spec_with(n, v, penumbra=pen)
    b = E.make_beam(gantry, 0.0, 0.0)
    terma = np.zeros(spec.bbox_shape, np.float32)
    x, y, z = 5, 9, 3
    terma[z, y, x] = 1.0
    code, g = O.geometry(spec, b, terma)
    assert code == 0
    bev = beam_fwd(float(b.azimuth), 0.0, 0.0) @ np.array([x, y, z], float)
    m = pen / v
    # the running maximum starts at FLT_MIN, the smallest positive float
    # (nvbbRayConvolve_host.cu:436), so a negative BEV coordinate is clamped to ~0
    top = np.maximum(bev, np.finfo(np.float32).tiny)
    np.testing.assert_allclose(list(g.bev_min), bev - m, atol=1e-4)
    np.testing.assert_allclose(list(g.bev_max), top + m, atol=1e-4)
    assert list(g.bev_size) == [int(np.ceil(v_ - (bev_ - 2 * m) - 1e-6)) for v_, bev_ in zip(top, bev)]
    # lim_min = start + (inverse beam rotation of bev_min + bbox_start) * voxel
    want = np.array(spec.start) + (beam_fwd(float(b.azimuth), 0.0, 0.0).T @ (bev - m)) * v
    np.testing.assert_allclose(list(g.lim_min), want, atol=1e-4)


def test_crop_box_empty_terma_is_a_geometry_error():
    spec = _spec_with(8, 1.0)
    code, g = O.geometry(spec, E.make_beam(0.0), np.zeros(spec.bbox_shape, np.float32))
    assert code == -2


def test_rev_grid_sizes_and_origins():
    n, v = 20, 0.5
    theta = [0.0, 30.0, 67.5, 112.5]
    phi = [0.0, 45.0, 202.5, 300.0]
    spec = _spec_with(n, v, theta=theta, phi=phi)
    b = E.make_beam(0.0)
    terma = np.zeros(spec.bbox_shape, np.float32)
    terma[4:15, 2:18, 6:13] = 1.0
    code, g = O.geometry(spec, b, terma)
    assert code == 0
    L = np.array(g.bev_size[:], float) * v
    for r in range(len(theta)):
        Rk = kernel_fwd(math.radians(theta[r]), math.radians(phi[r]))
        corners = np.array([[L[0] * i, L[1] * j, L[2] * k] for i in (0, 1) for j in (0, 1) for k in (0, 1)]) @ Rk.T
        mn, mx = corners.min(0), np.maximum(corners.max(0), np.finfo(np.float32).tiny)
        size = np.ceil(np.abs(mx - mn)[[1, 2, 0]] / v - 1e-9)
        got = np.array(g.rev_size[r][:], float)
        assert np.all(np.abs(got - size) <= 1), (r, got, size)
        np.testing.assert_allclose(g.rev_min[r][:], mn, atol=2e-5)
    if theta[0] == 0.0:
        assert list(g.rev_size[0]) == [g.bev_size[1], g.bev_size[2], g.bev_size[0]]


@pytest.mark.parametrize("gantry,ray", [(0.0, 1), (35.0, 2), (250.0, 5)])
def test_rev_density_resample_is_exact_on_linear_fields(gantry, ray):
    n, v = 28, 0.5
    idx = np.indices((n, n, n)).astype(np.float64)  # (k, j, i)
    a = np.array([0.6, 0.011, -0.007, 0.013])
    lin = a[0] + a[1] * idx[2] + a[2] * idx[1] + a[3] * idx[0]
    ph = W.Phantom(lin.astype(np.float32), v)
    spec = W.problem(ph, nphi=4, ntheta=4)
    b = E.make_beam(gantry, 0.0, 0.0, fmap_size=(6, 6), beamlet_size=(1.0, 1.0))
    t, _ = O.terma(spec, b, np.ones((6, 6), np.float32))
    code, g = O.geometry(spec, b, t)
    assert code == 0
    rs = g.rev_size[ray][:]
    dens, _, _ = O.cone_stage(spec, b, np.ones((6, 6), np.float32), ray, rs)
    th = float(spec.theta[ray]) * math.pi / 180
    phv = float(spec.phi[ray]) * math.pi / 180
    Kinv = kernel_fwd(th, phv).T
    Binv = beam_fwd(float(b.azimuth), float(b.zenith), float(b.coll)).T
    rstart = np.array(g.rev_min[ray][:], float) / v
    off = (np.array(g.lim_min[:], float) - np.array(spec.start)) / v
    tz, ty, tx = np.indices(dens.shape)
    rev = np.stack([tz, tx, ty], -1).reshape(-1, 3) + rstart   # REV coords (x=tZ, y=tX, z=tY)
    bev = rev @ Kinv.T
    gi = bev @ Binv.T + off                                    # PVCS index of each sample
    want = a[0] + gi @ a[1:]
    got = dens.reshape(-1)
    interior = np.all((gi > 1) & (gi < n - 2), 1) & (got >= 0)
    assert interior.sum() > 500
    np.testing.assert_allclose(got[interior], want[interior], atol=2e-4)
    # samples outside the BEV crop box are marked -3, outside the calc_bbox -2
    bs = np.array(g.bev_size[:], float)
    outside = np.any((bev < -1e-3) | (bev >= bs + 1e-3), 1)
    assert np.all(got[outside] == -3.0)


@pytest.mark.parametrize("gantry,ray", [(0.0, 0), (35.0, 3), (160.0, 6)])
def test_rev_to_pvcs_is_exact_on_linear_fields(gantry, ray):
    n, v = 24, 0.5
    spec = W.problem(W.water_cube(n, v), nphi=4, ntheta=4)
    b = E.make_beam(gantry, 0.0, 0.0, fmap_size=(6, 6), beamlet_size=(1.0, 1.0))
    t, _ = O.terma(spec, b, np.ones((6, 6), np.float32))
    code, g = O.geometry(spec, b, t)
    rs = g.rev_size[ray][:]
    tz, ty, tx = np.indices((rs[2], rs[1], rs[0])).astype(np.float64)
    c = np.array([2.0, 0.013, -0.021, 0.017])
    rev = (c[0] + c[1] * tx + c[2] * ty + c[3] * tz).astype(np.float32)  # linear in REV index
    dose = O.rev_to_pvcs(spec, b, g, ray, rev)
    th = float(spec.theta[ray]) * math.pi / 180
    phv = float(spec.phi[ray]) * math.pi / 180
    M = kernel_fwd(th, phv) @ beam_fwd(float(b.azimuth), float(b.zenith), float(b.coll))
    k, j, i = np.indices(dose.shape).astype(np.float64)
    gc = np.stack([i, j, k], -1).reshape(-1, 3) * v + np.array(spec.start)
    r = ((gc - np.array(g.lim_min[:], float)) / v) @ M.T - np.array(g.rev_min[ray][:], float) / v
    # REV coordinate (x, y, z) indexes the grid as (tZ, tX, tY)
    want = c[0] + c[1] * r[:, 1] + c[2] * r[:, 2] + c[3] * r[:, 0]
    got = dose.reshape(-1)
    inside = (r[:, 1] >= 0) & (r[:, 1] < rs[0] - 1) & (r[:, 2] >= 0) & (r[:, 2] < rs[1] - 1) & \
             (r[:, 0] >= 0) & (r[:, 0] < rs[2] - 1)
    margin = np.all(np.abs(r - np.round(r)) > 1e-3, 1)
    sel = inside & margin
    assert sel.sum() > 200
    np.testing.assert_allclose(got[sel], want[sel], rtol=1e-4)
    outside = (r[:, 1] < -1e-3) | (r[:, 2] < -1e-3) | (r[:, 0] < -1e-3) | (r[:, 1] > rs[0] + 1e-3)
    assert np.all(got[outside] == 0)
