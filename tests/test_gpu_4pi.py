"""GPU parity at the full 4pi precompute settings: the 128^3 water phantom with its ellipsoidal
PTV, beams of the 1162-beam 4pi candidate list (the reference's beamlist_4pi.txt), 64x64 maps of
0.25 cm beamlets open over the PTV projection and 96 cones (ntheta 12 x nphi 8, the largest cone
set the library takes: 48 bidirectional lines).

Bars:
  columns of a few beamlets per beam against the oracle's: <= 0.05 % of the column Dmax (dense
      columns of the batched path) and <= 0.05 % + the threshold (sparse columns scattered back);
  the whole beams, batched together: every active beamlet has a column, its keys are unique and
      inside the volume, every kept value is above threshold x column max, and each beam's columns
      are bit-identical to those of the same beam computed alone;
  beam-dose requests served while the precompute's beamlet requests run on the same resident
      state: bit-identical to the direct calls.
"""
import numpy as np
import pytest

from endtoend_amd import abi
from endtoend_amd import workloads as W
from endtoend_amd.engine import Engine, Server
from oracle import oracle as O
from tests.dose_metrics import max_err_pct

pytestmark = pytest.mark.gpu

TOL_PCT = 0.05
THRESH = 1e-4
INDICES = (97, 873)  # gantry 30 / couch -54 and gantry 210 / couch 30


@pytest.fixture(scope="module")
def w4():
    w = W.cfg5(indices=INDICES)
    assert w.spec.ntheta == 12 and w.spec.nphi == 8
    print(f"[case] {w.name}: {[int((f > 0).sum()) for f in w.fmaps]} active beamlets", flush=True)
    eng = Engine(w.spec)
    yield w, eng
    eng.close()


def _pick(f, n=3):
    """The first, middle and last active beamlets of a map, alone."""
    act = np.flatnonzero(f.ravel() > 0)
    g = np.zeros_like(f)
    for i in np.unique(act[np.linspace(0, len(act) - 1, n).astype(int)]):
        g.ravel()[i] = f.ravel()[i]
    return g


@pytest.mark.parametrize("bi", [0, 1])
def test_4pi_columns_match_oracle(w4, bi):
    w, eng = w4
    f = _pick(w.fmaps[bi])
    ref, _ = O.beamlet_dense(w.spec, w.beams[bi], f)
    dense = eng.beamlet_dense(w.beams[bi], f, abi.CCCS_BEAMLET_FAITHFUL)
    assert dense.shape == ref.shape and ref.shape[0] == int((f > 0).sum())
    errs = [max_err_pct(r, g) for r, g in zip(ref, dense)]
    print(f"beam {INDICES[bi]}: dense max err {max(errs):.5f} % Dmax over {len(errs)} columns", flush=True)
    assert max(errs) <= TOL_PCT, errs
    cols = eng.beamlet_dose([w.beams[bi]], [f], threshold=THRESH, mode=abi.CCCS_BEAMLET_FAITHFUL)
    nx, ny = w.spec.size[0], w.spec.size[1]
    bx, by, bz = (int(c) for c in w.spec.bbox_start)
    serr = []
    for (b, bl, k, v), r in zip(cols, ref):
        kk = k.astype(np.int64)
        x, y, z = kk % nx, (kk // nx) % ny, kk // (nx * ny)
        d = np.zeros(r.shape, np.float32)
        d[z - bz, y - by, x - bx] = v
        serr.append(max_err_pct(r, d))
    assert max(serr) <= TOL_PCT + 100.0 * THRESH, serr


def test_4pi_whole_beams_batched(w4):
    w, eng = w4
    nvox = int(np.prod(w.spec.size))
    both = eng.beamlet_dose(w.beams, w.fmaps, threshold=THRESH, mode=abi.CCCS_BEAMLET_FAITHFUL)
    want = [(b, int(i)) for b, f in enumerate(w.fmaps) for i in np.flatnonzero(f.ravel() > 0)]
    assert [(c[0], c[1]) for c in both] == want
    for b, bl, k, v in both:
        assert k.size > 0 and np.unique(k).size == k.size
        assert int(k.max()) < nvox
        assert float(v.min()) > THRESH * float(v.max())
    for bi in range(len(w.beams)):
        alone = eng.beamlet_dose([w.beams[bi]], [w.fmaps[bi]], threshold=THRESH, mode=abi.CCCS_BEAMLET_FAITHFUL)
        mine = [c for c in both if c[0] == bi]
        assert len(alone) == len(mine)
        for (_, bl, k, v), (_, bl2, k2, v2) in zip(mine, alone):
            assert bl == bl2 and np.array_equal(k, k2)
            assert np.array_equal(v.view(np.uint32), v2.view(np.uint32))


def test_4pi_beam_requests_during_precompute(w4):
    w, eng = w4
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(4):
        b = w.beams[i % len(w.beams)].copy()
        n = int(rng.integers(20, 65))
        b.fmap_size[:] = (n, n)
        b.beamlet_size[:] = (0.5, 0.5)
        reqs.append((b, rng.uniform(0.0, 1.0, (n, n)).astype(np.float32)))
    ref_dose = [eng.beam_dose([b], [f]) for b, f in reqs]
    ref_tot = [eng.beamlet_dose([b], [f], threshold=THRESH, keep=False) for b, f in zip(w.beams, w.fmaps)]
    with Server(w.spec, slots=2) as srv:
        pre = [srv.submit(b, f, abi.CCCS_REQ_BEAMLET, THRESH) for b, f in zip(w.beams, w.fmaps)]
        ids = [srv.submit(b, f, abi.CCCS_REQ_BEAM_DOSE) for b, f in reqs]
        got = [srv.wait(r) for r in ids]
        tot = [srv.wait(r, abi.CCCS_REQ_BEAMLET, counts_only=True)[0] for r in pre]
    for (dose, lat, _), ref in zip(got, ref_dose):
        assert np.array_equal(dose.view(np.uint32), ref.view(np.uint32))
        assert lat > 0
    for t, r in zip(tot, ref_tot):
        assert t["columns"] == r["columns"] and t["nnz"] == r["nnz"] and t["value_sum"] == r["value_sum"]
