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
    assert len(before) == len(after) > 10
    for a, b in zip(before, after):
        assert a[:2] == b[:2] and np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
