"""Edge cases through the C-ABI on the GPU: empty and single-beamlet fluence maps, threshold
extremes, bad arguments, a REV grid over the limit (the reference's ArraySizeError), and a context
that keeps working after an error."""
import numpy as np
import pytest

from endtoend_amd import abi
from endtoend_amd import workloads as W
from endtoend_amd.engine import Engine
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    w = W.small(n=28, voxel=0.5, gantry=40.0, hetero=True, fmap=(6, 6), beamlet=(1.0, 1.0), weights="rand")
    w.spec.kernel_extent = 1.0
    print("[case] edges small28", flush=True)
    eng = Engine(w.spec)
    yield w, eng
    eng.close()


def test_zero_fluence(small):
    w, eng = small
    z = np.zeros_like(w.fmaps[0])
    # full beam: no TERMA inside the calc_bbox is an error in both the library and the oracle
    with pytest.raises(abi.CCCSError) as e:
        eng.beam_dose(w.beams, [z])
    assert e.value.code == abi.CCCS_EGEOM
    with pytest.raises(Exception):
        O.beam_dose(w.spec, w.beams, [z])
    # beamlets: no active beamlet, no column
    assert eng.beamlet_dose(w.beams, [z], threshold=1e-4) == []
    assert eng.beamlet_dose([], [], threshold=1e-4) == []


def test_single_corner_beamlet_and_threshold_extremes(small):
    w, eng = small
    f = np.zeros_like(w.fmaps[0])
    f[0, 5] = 0.7
    cols0 = eng.beamlet_dose(w.beams, [f], threshold=0.0)
    assert len(cols0) == 1 and cols0[0][1] == 5
    dense = eng.beamlet_dense(w.beams[0], f)
    assert len(cols0[0][2]) == int(np.count_nonzero(dense[0]))   # threshold 0 keeps every v > 0
    top = eng.beamlet_dose(w.beams, [f], threshold=0.999999)
    vmax = float(dense[0].max())
    assert len(top[0][2]) >= 1 and np.all(top[0][3] > np.float32(0.999999) * vmax)
    assert eng.beamlet_dose(w.beams, [f], threshold=1.0)[0][2].size == 0   # v > max: none


def test_bad_arguments_leave_the_context_usable(small):
    w, eng = small
    f = w.fmaps[0]
    with pytest.raises(abi.CCCSError) as e:
        eng.beamlet_dose(w.beams, [f], threshold=-1.0)
    assert e.value.code == abi.CCCS_EINVAL
    with pytest.raises(abi.CCCSError) as e:
        eng.beamlet_dose(w.beams, [f], mode=7)
    assert e.value.code == abi.CCCS_EINVAL
    bad = w.beams[0].copy()
    bad.beamlet_size[:] = (0.0, 1.0)
    with pytest.raises(abi.CCCSError) as e:
        eng.beam_dose([bad], [f])
    assert e.value.code == abi.CCCS_EINVAL
    # the context still computes the same dose afterwards
    d1 = eng.beam_dose(w.beams, [f])
    d2 = eng.beam_dose(w.beams, [f])
    assert np.array_equal(d1, d2) and d1.max() > 0


def test_rev_grid_over_the_limit_is_array_size_error():
    """Request 135 of the 300-request serving stream (a 32 cm field at gantry 245 deg, couch 24 deg on the
    256x256x180 thorax) has REV grids beyond the beam path's derived max_rev_size.  The reference
    throws ArraySizeError (nvbbRayConvolve_host.cu:150-153); the oracle's geometry and the library
    both report CCCS_EGEOM, and the next request on the same context succeeds."""
    w = W.cfg4(nreq=300)   # the stream draws its arrival times first, so the length fixes the beams
    b, f = w.beams[135], w.fmaps[135]
    import ctypes as C
    t, _ = O.terma(w.spec, b, f)
    g = abi.BeamGeometry()
    p = w.spec.struct()
    assert O.load().orc_beam_geometry(C.byref(p), C.byref(b), O.fp(t), C.byref(g)) == abi.CCCS_EGEOM
    print("[case] serving request 135 (REV over the limit)", flush=True)
    with Engine(w.spec) as eng:
        with pytest.raises(abi.CCCSError) as e:
            eng.beam_dose([b], [f])
        assert e.value.code == abi.CCCS_EGEOM
        d = eng.beam_dose([w.beams[1]], [w.fmaps[1]])
        assert d.max() > 0
