"""The reference-named shim on the GPU, through the test program that calls the reference's entry
points in main.cu's order (tests/shim/shim_driver.cpp):

* dosecalc-beam: initCudaConstandTex -> radconvolveTexture -> cuda_sum_device_float_dose ->
  freeCudaTexture gives a dose bit-identical to cccs_beam_dose of the same beams;
* dosecalc-beamlet: SparsifyManager::init/activate -> initCudaConstandTex -> radconvolveTexture ->
  deactivate -> freeCudaTexture writes a coefficient file whose columns (keys, values,
  percentages) are bit-identical to cccs_beamlet_dose's, filed under each BEAM's uid, with the
  CENTRAL write strategy and with PER_BEAMLET (each column in a sidecar file behind external links)."""
import os
import subprocess

import numpy as np
import pytest

from endtoend_amd import abi
from endtoend_amd import build as B
from endtoend_amd import workloads as W
from endtoend_amd.engine import Engine

from tests.shim_case import DRIVERS, write_case

pytestmark = pytest.mark.gpu


def _run(kind, case, out, *extra):
    print(f"[case] shim driver {kind} {' '.join(extra)}", flush=True)
    r = subprocess.run([DRIVERS[kind], case, out, *extra], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_beam_entry_points_match_cccs_beam_dose(tmp_path):
    ws = [W.small(n=28, voxel=0.5, gantry=g, couch=c, hetero=True, fmap=(6, 6), weights="rand", seed=s)
          for g, c, s in ((30.0, 0.0, 3), (215.0, -40.0, 4))]
    spec = ws[0].spec
    beams = [w.beams[0] for w in ws]
    fmaps = [w.fmaps[0] for w in ws]
    write_case(str(tmp_path), spec, beams, fmaps)
    _run("beam", str(tmp_path), str(tmp_path / "dose.raw"))
    got = np.fromfile(str(tmp_path / "dose.raw"), np.float32).reshape(spec.density.shape)
    with Engine(spec) as eng:
        want = eng.beam_dose(beams, fmaps)
    assert want.max() > 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.skipif(B.hdf5_root() is None, reason="no HDF5 C library")
@pytest.mark.parametrize("strategy", ["central", "per_beamlet"])
def test_beamlet_entry_points_write_the_library_columns(tmp_path, strategy):
    from tests import h5_walk as HW
    ws = [W.small(n=28, voxel=0.5, gantry=g, couch=c, hetero=True, fmap=(6, 6)) for g, c in
          ((30.0, 0.0), (120.0, 20.0), (300.0, -30.0))]
    spec = ws[0].spec
    spec.kernel_extent = 1.0
    beams = [w.beams[0] for w in ws]
    fmaps = []
    for i, w in enumerate(ws):
        f = w.fmaps[0].copy()
        f.ravel()[i::3] = 0.0
        fmaps.append(f)
    uids = [4, 5, 9]  # two consecutive uids share a request, the third goes alone
    write_case(str(tmp_path), spec, beams, fmaps, uids=uids)
    out = str(tmp_path / "coeffs.h5")
    _run("beamlet", str(tmp_path), out, strategy)
    with Engine(spec) as eng:
        cols = eng.beamlet_dose(beams, fmaps, threshold=1e-4, mode=abi.CCCS_BEAMLET_FAITHFUL, percs=True)
    data = HW.walk(out).groups["beams"].groups["data"]
    n = 0
    for bi, bl, k, v, pc in cols:
        g = data.groups[f"beam_{uids[bi]:05d}"]
        if len(k) == 0:
            continue
        c = g.groups[f"beamlet_{bl:05d}"]
        assert np.array_equal(c.datasets["lindex"].data, k)
        assert np.array_equal(c.datasets["coeffs"].data.view(np.uint32), v.view(np.uint32))
        assert c.attrs["perc_nonzero"].value == np.float32(pc[0])
        assert c.attrs["perc_dropped"].value == np.float32(pc[1])
        if strategy == "per_beamlet":   # read above through the link into the column's own file
            side = f"coeffs.h5_external/beam_{uids[bi]:05d}/beam{uids[bi]:05d}_beamlet{bl:05d}.bmlt"
            assert HW.external_link(out, f"/beams/data/beam_{uids[bi]:05d}/beamlet_{bl:05d}/coeffs") == \
                (side, "/coeffs")
            assert os.path.exists(os.path.join(str(tmp_path), side))
        n += 1
    assert n == sum(int((f > 0).sum()) for f in fmaps) > 20
