"""ROI reduction of sparse columns (_reduce_by_roi, DoseCalcIO/sparse_io.cpp:62-158), the third
'next' row of the scope table.  The oracle is pinned by forms the reduction must take:
* one ROI covering the whole grid re-keys voxel (x, y, z) to (x*ny + y)*nz + z and orders the
  column by that key (the x-outer, z-inner walk);
* a voxel inside two ROIs appears in both blocks, keyed past the first ROI's voxel count;
* voxels outside every ROI vanish, and the result does not depend on the input order.
"""
import numpy as np

from oracle import oracle as O

SIZE = (7, 5, 6)  # nx, ny, nz


def _column(rng, n=60):
    nx, ny, nz = SIZE
    keys = rng.choice(nx * ny * nz, n, replace=False).astype(np.uint64)
    vals = rng.uniform(0.1, 1.0, n).astype(np.float32)
    return keys, vals


def _xyz(keys):
    nx, ny, _ = SIZE
    k = keys.astype(np.int64)
    return k % nx, (k // nx) % ny, k // (nx * ny)


def test_full_grid_roi_is_a_transposed_linear_index():
    rng = np.random.default_rng(1)
    keys, vals = _column(rng)
    nx, ny, nz = SIZE
    full = np.ones((nz, ny, nx), np.uint8)
    rk, rv, blocks = O.reduce_by_roi(keys, vals, SIZE, [(full, (0, 0, 0), SIZE)])
    x, y, z = _xyz(keys)
    expect = (x * ny + y) * nz + z
    order = np.argsort(expect)
    assert np.array_equal(rk, expect[order].astype(np.uint64))
    assert np.array_equal(rv, vals[order])
    assert blocks.tolist() == [len(keys)]


def test_overlapping_rois_duplicate_and_offset():
    rng = np.random.default_rng(2)
    keys, vals = _column(rng, 80)
    nx, ny, nz = SIZE
    # ROI A: box x in [1,5), y in [0,3), z in [2,6); ROI B: a sphere-ish blob over the whole grid
    a_start, a_size = (1, 0, 2), (4, 3, 4)
    a = np.ones((a_size[2], a_size[1], a_size[0]), np.uint8)
    Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    b = (((X - 3) ** 2 + (Y - 2) ** 2 + (Z - 3) ** 2) <= 5).astype(np.uint8)
    rk, rv, blocks = O.reduce_by_roi(keys, vals, SIZE, [(a, a_start, a_size), (b, (0, 0, 0), SIZE)])
    x, y, z = _xyz(keys)
    in_a = (x >= 1) & (x < 5) & (y < 3) & (z >= 2)
    in_b = b[z, y, x].astype(bool)
    assert blocks.tolist() == [int(in_a.sum()), int(in_b.sum())]
    na = int(a.sum())
    # block A: rank inside the box walked x outer, z inner
    la = ((x - 1) * 3 + y) * 4 + (z - 2)
    sel = np.argsort(la[in_a])
    assert np.array_equal(rk[:blocks[0]], la[in_a][sel].astype(np.uint64))
    assert np.array_equal(rv[:blocks[0]], vals[in_a][sel])
    # block B keys start after all of A's voxels and rank B's voxels in the same walk order
    walk = np.transpose(b, (2, 1, 0)).ravel()          # x outer, y, z inner
    rank = np.cumsum(walk) - 1
    idx = (x * ny + y) * nz + z
    lb = na + rank[idx[in_b]]
    sel = np.argsort(lb)
    assert np.array_equal(rk[blocks[0]:], lb[sel].astype(np.uint64))
    assert np.array_equal(rv[blocks[0]:], vals[in_b][sel])
    assert (in_a & in_b).sum() > 0          # the duplicated voxels were exercised


def test_outside_voxels_vanish_and_input_order_is_irrelevant():
    rng = np.random.default_rng(3)
    keys, vals = _column(rng, 90)
    nx, ny, nz = SIZE
    m = np.zeros((2, 2, 3), np.uint8)
    m[0, 1, 2] = m[1, 0, 0] = 1
    rois = [(m, (2, 1, 3), (3, 2, 2))]
    rk, rv, blocks = O.reduce_by_roi(keys, vals, SIZE, rois)
    perm = rng.permutation(len(keys))
    rk2, rv2, blocks2 = O.reduce_by_roi(keys[perm], vals[perm], SIZE, rois)
    assert np.array_equal(rk, rk2) and np.array_equal(rv, rv2) and np.array_equal(blocks, blocks2)
    assert set(rk.tolist()) <= {0, 1}
    # an empty column reduces to nothing
    ek, ev, eb = O.reduce_by_roi(np.zeros(0, np.uint64), np.zeros(0, np.float32), SIZE, rois)
    assert len(ek) == 0 and eb.tolist() == [0]
