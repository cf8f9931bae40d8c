"""CPU check of the cone kernel's segment-sum walks (tests/audit/walk_check.hip): on random
heterogeneous lines (water, lung, bone and air runs, TERMA with zero runs) the segment-sum walks
agree with the sequential per-sample walk of the reference's RowConvolve loop to 2e-3 of the line's
largest value (the reference accumulates radiological distance in FP32; lines with long air runs
drift by up to about 1e-3), and the kernel's scatter-built lookup table equals its definition entry
by entry."""
import json
import os
import subprocess

CHECK = os.path.join(os.path.dirname(os.path.abspath(__file__)), "audit", "_build", "walk_check")


def test_walk_forms_agree():
    r = subprocess.run([CHECK, "1500", "11"], capture_output=True, text=True, timeout=300)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert r.returncode == 0, out
    assert out["walk_mismatch"] == 0 and out["table_errors"] == 0
    assert out["destinations"] > 200000 and out["iters"] > 0
