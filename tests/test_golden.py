"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the oracle).

CPU: the oracle still reproduces every fixture bit for bit (a regression pin on the checker).
GPU: the HIP path against the committed vectors, at the parity bars of the other GPU tests:
Siddon segment lists bit-exact, TERMA rtol 2e-5, beam dose <= 0.5 % Dmax, faithful beamlet
columns <= 0.05 % of the column Dmax, sparse columns equal to _sparsify of the same dense column.
"""
import importlib.util
import json
import os

import numpy as np
import pytest

from endtoend_amd import abi
from tests.dose_metrics import max_err_pct

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
G = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(G)


def _load(name):
    z = np.load(os.path.join(HERE, name))  # allow_pickle=False (default)
    return json.loads(str(z["case"])), z


def _bits_equal(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "f":
        return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.astype(a.dtype).view(np.uint32))
    return np.array_equal(a, b)


def test_oracle_reproduces_beam_fixture():
    case, z = _load("beam_small.npz")
    out = G.beam_outputs(G.case_workload(case))
    for k, v in out.items():
        assert _bits_equal(v, z[k]), k


def test_oracle_reproduces_beamlet_fixture():
    case, z = _load("beamlet_small.npz")
    out = G.beamlet_outputs(G.case_workload(case), case)
    for k, v in out.items():
        assert _bits_equal(v, z[k]), k


@pytest.mark.gpu
def test_gpu_beam_path_against_fixture():
    from endtoend_amd.engine import Engine
    case, z = _load("beam_small.npz")
    w = G.case_workload(case)
    b, f = w.beams[0], w.fmaps[0]
    eng = Engine(w.spec)
    try:
        pts = G.siddon_points(w)
        g = eng.siddon_paths(b.source[:], pts, 256)
        assert np.array_equal(g[3], z["counts"])
        for p in range(len(pts)):
            n = int(z["counts"][p])
            assert np.array_equal(g[0][p, :n], z["ijk"][p, :n])
            assert _bits_equal(g[1][p, :n], z["mid"][p, :n])
            assert _bits_equal(g[2][p, :n], z["len"][p, :n])
        t = eng.beam_terma(b, f)
        np.testing.assert_allclose(t, z["terma"], rtol=2e-5, atol=1e-7 * float(z["terma"].max()))
        d = eng.beam_dose([b], [f])
        err = max_err_pct(z["dose"], d)
        print(f"golden beam: max err {err:.5f} % Dmax", flush=True)
        assert err <= 0.5
    finally:
        eng.close()


@pytest.mark.gpu
def test_gpu_beamlet_path_against_fixture():
    from endtoend_amd.engine import Engine
    case, z = _load("beamlet_small.npz")
    w = G.case_workload(case)
    f = G.beamlet_fmap(w, case)
    eng = Engine(w.spec)
    try:
        dense = eng.beamlet_dense(w.beams[0], f, abi.CCCS_BEAMLET_FAITHFUL)
        ref = z["columns"]
        assert dense.shape == ref.shape
        errs = [max_err_pct(r, d) for r, d in zip(ref, dense)]
        print(f"golden beamlet: max err {max(errs):.6f} % Dmax", flush=True)
        assert max(errs) <= 0.05
        cols = eng.beamlet_dose([w.beams[0]], [f], threshold=case["threshold"], mode=abi.CCCS_BEAMLET_FAITHFUL)
        assert [c[1] for c in cols] == case["active"]
        from oracle import oracle as O
        for i, (bi, bl, k, v) in enumerate(cols):
            ek, ev, _, _ = O.sparsify(dense[i], w.spec.size, w.spec.bbox_start, w.spec.bbox_size, case["threshold"])
            assert np.array_equal(k, ek) and _bits_equal(v, ev)
            # and against the fixture's keys: any difference sits at the threshold
            sym = np.setxor1d(k, z[f"keys{i}"])
            assert sym.size <= max(2, k.size // 1000)
    finally:
        eng.close()
