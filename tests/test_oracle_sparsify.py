"""Pins _sparsify (A15, DoseCalcIO/sparse_io.cpp:22-58) on hand-built columns: keep v > t*max,
keys are full-volume linear indices emitted with i outermost and k innermost."""
import numpy as np

from oracle import oracle as O


def test_hand_built_column():
    size = (4, 3, 2)           # full volume Nx, Ny, Nz
    crop_start = (1, 1, 0)
    crop_size = (2, 2, 2)      # crop is (z, y, x) = (2, 2, 2)
    dense = np.zeros((2, 2, 2), np.float32)
    dense[0, 0, 0] = 10.0      # (i=0, j=0, k=0) -> full (1,1,0) -> key 0*12 + 1*4 + 1 = 5
    dense[1, 0, 0] = 4.0       # (i=0, j=0, k=1) -> (1,1,1) -> 12 + 4 + 1 = 17
    dense[0, 1, 1] = 0.5       # (i=1, j=1, k=0) -> (2,2,0) -> 8 + 2 = 10 ; dropped at t = 0.1
    dense[1, 1, 0] = 7.0       # (i=0, j=1, k=1) -> (1,2,1) -> 12 + 8 + 1 = 21
    dense[0, 0, 1] = 1.0       # (i=1, j=0, k=0) -> (2,1,0) -> 4 + 2 = 6 ; exactly 0.1*max: dropped
    keys, vals, pnz, pdrop = O.sparsify(dense, size, crop_start, crop_size, 0.1)
    # emission order: i outer, then j, then k inner
    assert keys.tolist() == [5, 17, 21]
    assert vals.tolist() == [10.0, 4.0, 7.0]
    np.testing.assert_allclose(pnz, 100.0 * 3 / 8, rtol=1e-6)
    np.testing.assert_allclose(pdrop, 100.0 * 2 / 8, rtol=1e-6)


def test_threshold_zero_keeps_every_positive_value_in_loop_order():
    rng = np.random.default_rng(0)
    crop = (5, 4, 3)
    dense = rng.uniform(-1, 1, (crop[2], crop[1], crop[0])).astype(np.float32)
    keys, vals, _, _ = O.sparsify(dense, (9, 8, 7), (2, 3, 1), crop, 0.0)
    exp_k, exp_v = [], []
    for i in range(crop[0]):
        for j in range(crop[1]):
            for k in range(crop[2]):
                if dense[k, j, i] > 0:
                    exp_k.append(((k + 1) * 8 + (j + 3)) * 9 + (i + 2))
                    exp_v.append(dense[k, j, i])
    assert keys.tolist() == exp_k
    assert np.array_equal(vals, np.array(exp_v, np.float32))
