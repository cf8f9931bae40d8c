"""The schedule autotuner (cccs_autotune): it picks values from its candidate sets, never reports a
slower choice than the one it started from, and the results after tuning are those before it
(bit-identical for the beamlet wave counts, FP32 rounding for the cone-kernel schedule)."""
import numpy as np
import pytest

from endtoend_amd import abi
from endtoend_amd import workloads as W
from endtoend_amd.engine import Engine
from tests.dose_metrics import max_err_pct

pytestmark = pytest.mark.gpu


def test_autotune_beam_path_keeps_the_dose():
    w = W.cfg1()
    print("[case] autotune beam path on the 128^3 water beam", flush=True)
    with Engine(w.spec) as eng:
        before = eng.beam_dose(w.beams, w.fmaps)
        rep = eng.autotune(w.beams, w.fmaps, kind=abi.CCCS_TUNE_BEAM)
        after = eng.beam_dose(w.beams, w.fmaps)
    print(rep, flush=True)
    k = rep["knobs"]
    assert k["cone_ramp_min"] in (32, 48, 64, 96, 128)
    assert k["cone_class_threads2"] in (64, 128, 256)
    assert rep["best_ms"] <= rep["default_ms"] and rep["trials"] >= 10
    assert max_err_pct(before, after) <= 1e-3


def test_autotune_beamlet_path_keeps_the_columns():
    w = W.small(n=32, voxel=0.5, gantry=30.0, hetero=True, fmap=(8, 8), beamlet=(1.0, 1.0), weights="rand")
    w.spec.kernel_extent = 1.0
    print("[case] autotune beamlet path", flush=True)
    with Engine(w.spec) as eng:
        before = eng.beamlet_dose(w.beams, w.fmaps, threshold=1e-4)
        rep = eng.autotune(w.beams, w.fmaps, kind=abi.CCCS_TUNE_BEAMLET)
        after = eng.beamlet_dose(w.beams, w.fmaps, threshold=1e-4)
    print(rep, flush=True)
    assert rep["knobs"]["bl_walk_waves"] in (16384, 32768, 65536)
    assert rep["knobs"]["bl_streams"] in (1, 2, 3, 4)
    assert rep["knobs"]["bl_walk_split"] in (0, 128, 256, 384)
    assert len(before) == len(after) > 10
    for a, b in zip(before, after):
        assert a[:2] == b[:2] and np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])


def test_two_ray_pipeline_gives_the_same_columns():
    """bl_streams = n overlaps consecutive cone directions on n streams; the columns must be
    identical to the one-stream pipeline, on a multi-beam job with every mode that uses it."""
    w = W.small(n=32, voxel=0.5, gantry=215.0, couch=-45.0, hetero=True, fmap=(8, 8), beamlet=(0.8, 0.8),
                weights="rand", nphi=8, ntheta=8)
    w.spec.kernel_extent = 1.5
    b2 = W.small(n=32, voxel=0.5, gantry=30.0, hetero=True, fmap=(8, 8), beamlet=(1.0, 1.0), weights="rand").beams[0]
    beams, fmaps = [w.beams[0], b2], [w.fmaps[0], w.fmaps[0]]
    print("[case] one vs two streams", flush=True)
    with Engine(w.spec) as eng:
        out = {}
        for ns in (1, 2, 3, 4):
            eng.set_schedule("bl_streams", ns)
            out[ns] = [eng.beamlet_dose(beams, fmaps, threshold=1e-4, mode=abi.CCCS_BEAMLET_FAITHFUL)]
    for ns in (2, 3, 4):
        for c1, c2 in zip(out[1], out[ns]):
            assert len(c1) == len(c2) > 10
            for a, b in zip(c1, c2):
                assert a[:2] == b[:2] and np.array_equal(a[2], b[2])
                assert np.array_equal(a[3].view(np.uint32), b[3].view(np.uint32))
