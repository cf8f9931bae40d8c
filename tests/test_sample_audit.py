"""CPU audit of the beamlet sample stage's indexing (tests/audit/sample_audit.hip): the library's own
planning builds the job tables of a workload, and the sample kernel's passes are replayed on the
host with the kernel's own helpers, checking every LDS, queue, beamlet-record and REV-store index
against its bound, and that no line queues more TERMA requests than it reserved.  The pillar
mapping, whose quotients skip the IEEE division away from multiples, is compared bit for bit with
its division form on every call, and that quotient floor with the IEEE one on an adversarial
sweep.  The pass A.1 cull (footprint half-width and the depth-linear bounds of footprint_cull)
must keep every footprint sample: every sample of every third audited line is classified exactly,
and each footprint sample the cull drops is a violation.  Run on the 50-beam batched workload
(every 13th REV line of every direction) and on two beams of the 4pi precompute (2.5 mm
beamlets, 96 cones, every 5th line)."""
import json
import os
import subprocess

import pytest

from endtoend_amd import workloads as W

from tests.shim_case import write_case

AUDIT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "audit", "_build", "sample_audit")


def _audit(tmp_path, w, stride):
    write_case(str(tmp_path), w.spec, w.beams, w.fmaps)
    r = subprocess.run([AUDIT, str(tmp_path), str(stride)], capture_output=True, text=True, timeout=600)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert r.returncode == 0, out
    return out


def test_batched_workload_indices_stay_in_bounds(tmp_path):
    out = _audit(tmp_path, W.cfg3(), 13)
    assert out["violations"] == 0 and out["footprint_lines"] > 1000
    assert out["max_len"] <= 1024 and out["max_pillar_word"] < 128 and out["max_footprint_pillars"] <= 64
    assert out["shortfall"] == 0
    assert out["locates"] > 100000 and out["floor_quot_checked"] > 10000000 and out["floor_quot_mismatch"] == 0
    assert out["cull_checked"] > 1000000


def test_4pi_sample_indices_stay_in_bounds(tmp_path):
    out = _audit(tmp_path, W.cfg5(indices=[0, 581]), 5)
    assert out["violations"] == 0 and out["footprint_lines"] > 1000 and out["directions"] == 48
    assert out["locates"] > 100000 and out["floor_quot_mismatch"] == 0
    assert out["cull_checked"] > 50000
