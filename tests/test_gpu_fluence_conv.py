"""GPU fluence-map convolution (cccs_fluence_convolve) against the oracle: bit-identical (the same
FP32 terms added in the same order)."""
import numpy as np
import pytest

from endtoend_amd import abi
from endtoend_amd import workloads as W
from endtoend_amd.engine import Engine
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    w = W.small(n=16)
    e = Engine(w.spec)
    yield e
    e.close()


@pytest.mark.parametrize("n,h,w,kh,kw", [(1, 20, 20, 1, 1), (3, 40, 40, 3, 5), (2, 37, 53, 31, 31), (5, 64, 64, 9, 7),
                                          (1, 1, 1, 3, 3)])
def test_convolve_matches_oracle(eng, n, h, w, kh, kw):
    print(f"[case] conv n{n} {h}x{w} k{kh}x{kw}", flush=True)
    rng = np.random.default_rng(n * 1000 + h)
    f = rng.uniform(0, 1, (n, h, w)).astype(np.float32)
    f[f < 0.3] = 0.0
    k = rng.uniform(-0.2, 1, (kh, kw)).astype(np.float32)
    got = eng.fluence_convolve(f, k)
    ref = O.fluence_convolve(f, k)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_convolve_rejects_even_or_large_kernels(eng):
    f = np.ones((8, 8), np.float32)
    for k in (np.ones((2, 3), np.float32), np.ones((33, 1), np.float32)):
        with pytest.raises(abi.CCCSError) as e:
            eng.fluence_convolve(f, k)
        assert e.value.code == abi.CCCS_EINVAL
