"""A0, FCBB part: the fluence-map convolution the north star names (the reference's README.md:15-20
describes it; no source exists).  Definition: zero outside the map, same size,
out[y][x] = sum_j sum_i K[j][i] F[y - (j - kh/2)][x - (i - kw/2)].  The oracle is pinned against
scipy.signal.convolve2d (mode 'same', zero fill) and closed forms: a delta kernel is the
identity, a shifted delta shifts with zero fill, the map is linear, and with a zero border the
total is the map total times the kernel total."""
import numpy as np
import pytest
from scipy.signal import convolve2d

from oracle import oracle as O


@pytest.mark.parametrize("h,w,kh,kw", [(20, 20, 3, 3), (13, 29, 5, 7), (40, 17, 31, 1), (7, 9, 9, 11)])
def test_matches_scipy_convolve2d(h, w, kh, kw):
    rng = np.random.default_rng(h * 100 + w)
    f = rng.uniform(0, 1, (h, w)).astype(np.float32)
    k = rng.normal(0, 1, (kh, kw)).astype(np.float32)
    got = O.fluence_convolve(f, k)
    ref = convolve2d(f.astype(np.float64), k.astype(np.float64), mode="same", boundary="fill", fillvalue=0.0)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())


def test_delta_kernels_and_linearity():
    rng = np.random.default_rng(7)
    f = rng.uniform(0, 1, (3, 16, 24)).astype(np.float32)
    d = np.zeros((5, 5), np.float32)
    d[2, 2] = 1.0
    assert np.array_equal(O.fluence_convolve(f, d), f)
    s = np.zeros((5, 5), np.float32)
    s[2, 4] = 1.0          # i - kw/2 = 2: out[y][x] = F[y][x - 2]
    out = O.fluence_convolve(f, s)
    assert np.array_equal(out[:, :, 2:], f[:, :, :-2]) and not out[:, :, :2].any()
    g = rng.uniform(0, 1, f.shape).astype(np.float32)
    k = rng.uniform(0, 1, (3, 5)).astype(np.float32)
    lhs = O.fluence_convolve(2.0 * f + 3.0 * g, k)
    rhs = 2.0 * O.fluence_convolve(f, k) + 3.0 * O.fluence_convolve(g, k)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-5, atol=1e-5)


def test_total_is_preserved_with_a_zero_border():
    rng = np.random.default_rng(9)
    f = np.zeros((30, 30), np.float32)
    f[6:24, 6:24] = rng.uniform(0, 1, (18, 18))
    k = rng.uniform(0, 1, (7, 9)).astype(np.float32)
    out = O.fluence_convolve(f, k)
    assert abs(float(out.sum(dtype=np.float64)) - float(f.sum(dtype=np.float64)) * float(k.sum(dtype=np.float64))) \
        < 1e-4 * float(out.sum())
