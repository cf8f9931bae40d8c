"""A request larger than the job budget is split into jobs that run one after another, each with
its own tables, packed grids and compaction.  Splitting a request into many jobs (a 1 GiB job
budget) must give the same columns, bit for bit and in the same order, as one job: through a Python
sink, the native counting sink, with device-resident columns between calls, and with ROI
reduction (whose compaction takes a second host round trip)."""
import numpy as np
import pytest

from endtoend_amd import abi
from endtoend_amd import workloads as W
from endtoend_amd.engine import Engine

pytestmark = pytest.mark.gpu

FA = abi.CCCS_BEAMLET_FAITHFUL


def _columns(eng, w, n):
    return eng.beamlet_dose(w.beams[:n], w.fmaps[:n], threshold=1e-4, mode=FA, percs=True)


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x[0] == y[0] and x[1] == y[1]
        assert np.array_equal(x[2], y[2]) and np.array_equal(x[3].view(np.uint32), y[3].view(np.uint32))
        assert x[-1] == y[-1]


def test_many_jobs_give_the_columns_of_one_job():
    w = W.cfg3()
    n = 4
    with Engine(w.spec) as eng:
        eng.set_schedule("bl_job_gb", 256)
        print("[case] one job", flush=True)
        one = _columns(eng, w, n)
        tot_one = eng.beamlet_dose(w.beams[:n], w.fmaps[:n], threshold=1e-4, mode=FA, keep=False)
        eng.set_schedule("bl_job_gb", 1)
        print("[case] 1 GiB jobs", flush=True)
        many = _columns(eng, w, n)
        _same(one, many)
        assert len(many) == sum(int((f > 0).sum()) for f in w.fmaps[:n])
        tot_many = eng.beamlet_dose(w.beams[:n], w.fmaps[:n], threshold=1e-4, mode=FA, keep=False)
        assert tot_many == tot_one
        # device-resident columns, then a call that copies them out: the second call's jobs must
        # follow the first call's compactions
        eng.beamlet_dose(w.beams[:n], w.fmaps[:n], threshold=1e-4, mode=FA, keep=None)
        again = _columns(eng, w, n)
        _same(one, again)


def test_many_jobs_with_roi_reduction():
    w = W.cfg3()
    n = 3
    nz, ny, nx = w.spec.density.shape
    slab = np.zeros((nz, ny, nx), np.uint8)
    slab[nz // 2 - 4:nz // 2 + 4] = 1
    box = np.ones((20, 30, 40), np.uint8)
    box[:, ::2, :] = 0
    rois = [(box, (nx // 2 - 20, ny // 2 - 15, nz // 2 - 10), (40, 30, 20)), (slab, (0, 0, 0), (nx, ny, nz))]
    with Engine(w.spec) as eng:
        eng.set_roi_masks(rois)
        eng.set_schedule("bl_job_gb", 256)
        print("[case] roi one job", flush=True)
        one = _columns(eng, w, n)
        eng.set_schedule("bl_job_gb", 1)
        print("[case] roi 1 GiB jobs", flush=True)
        many = _columns(eng, w, n)
        _same(one, many)
        for x, y in zip(one, many):
            assert np.array_equal(x[4], y[4])  # row-block sizes
