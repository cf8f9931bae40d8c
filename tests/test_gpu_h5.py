"""Beamlet columns streamed from the GPU pipeline into Dose_Coefficients.h5 through the native
HDF5 sink (cccs_h5_coeff_sink) are the columns the pipeline delivers to any other sink: keys,
values (bit for bit), perc_nonzero / perc_dropped and, for ROI-reduced columns, the row-block
sizes; and a full-beam dose written as Dose.h5 reads back bit for bit."""
import ctypes as C

import numpy as np
import pytest

from endtoend_amd import abi
from endtoend_amd import build as B
from endtoend_amd import workloads as W
from endtoend_amd.engine import Engine, beams_array

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(B.hdf5_root() is None, reason="no HDF5 C library")]

if B.hdf5_root() is not None:
    from endtoend_amd import h5 as H
    from tests import h5_walk as HW


def _all_columns(eng, beams, fmaps):
    """Every column with its percentages, through a Python sink."""
    out = {}

    def sink(user, colp):
        c = colp.contents
        k = np.ctypeslib.as_array(c.keys, shape=(c.nnz,)).copy() if c.nnz else np.zeros(0, np.uint64)
        v = np.ctypeslib.as_array(c.values, shape=(c.nnz,)).copy() if c.nnz else np.zeros(0, np.float32)
        blocks = np.ctypeslib.as_array(c.row_block_sizes, shape=(c.nblocks,)).copy() if c.nblocks else None
        out[(c.beam_index, c.beamlet_index)] = (k, v, c.perc_nonzero, c.perc_dropped, blocks, c.perc_reducedrop)
        return 0

    cb = abi.SPARSE_SINK(sink)
    fm = [np.ascontiguousarray(f, np.float32) for f in fmaps]
    ptrs = (abi._fp * len(fm))(*[abi.fptr(f) for f in fm])
    abi.check(eng.lib.cccs_beamlet_dose(eng.ctx, beams_array(beams), len(beams), ptrs, 1e-4,
                                        abi.CCCS_BEAMLET_FAITHFUL, cb, None), "cccs_beamlet_dose")
    return out


def _check_file(path, cols, uid_base, reduced):
    root = HW.walk(path)
    data = root.groups["beams"].groups["data"]
    nwritten = 0
    for (bi, bl), (k, v, pnz, pdr, blocks, prd) in cols.items():
        bg = data.groups[f"beam_{uid_base + bi:05d}"]
        name = f"beamlet_{bl:05d}"
        if len(k) == 0:
            assert name not in bg.groups
            continue
        g = bg.groups[name]
        nwritten += 1
        assert np.array_equal(g.datasets["lindex"].data, k)
        assert np.array_equal(g.datasets["coeffs"].data.view(np.uint32), v.view(np.uint32))
        assert g.attrs["perc_nonzero"].value == np.float32(pnz) and g.attrs["perc_dropped"].value == np.float32(pdr)
        if reduced:
            assert np.array_equal(g.attrs["row_block_sizes"].value, blocks)
            assert g.attrs["perc_reducedrop"].value == np.float32(prd)
        else:
            assert "row_block_sizes" not in g.attrs
    assert nwritten > 10
    return root


def test_beamlet_columns_stream_into_the_coefficient_file(tmp_path):
    w = W.small(n=32, voxel=0.5, gantry=30.0, hetero=True, fmap=(8, 8), beamlet=(1.0, 1.0), weights="rand")
    w.spec.kernel_extent = 1.0
    b2 = W.small(n=32, voxel=0.5, gantry=250.0, couch=20.0, hetero=True, fmap=(6, 6), beamlet=(1.0, 1.0))
    beams, fmaps = [w.beams[0], b2.beams[0]], [w.fmaps[0], b2.fmaps[0]]
    fmaps[0] = fmaps[0].copy()
    fmaps[0].ravel()[::4] = 0.0
    path = str(tmp_path / "Dose_Coefficients.h5")
    print("[case] h5 sink, 2 beams", flush=True)
    with Engine(w.spec) as eng:
        cols = _all_columns(eng, beams, fmaps)
        with H.CoeffFile(path, w.spec.struct(), nbeams=2) as f:
            eng.beamlet_dose_h5(beams, fmaps, f, uid_base=0)
            ncols, ncoef = f.counts()
    assert ncols == sum(1 for c in cols.values() if len(c[0])) and ncoef == sum(len(c[0]) for c in cols.values())
    root = _check_file(path, cols, 0, reduced=False)
    meta = root.groups["beams"].groups["metadata"].groups
    for i, f in enumerate(fmaps):
        assert int(meta[f"beam_{i:05d}"].attrs["N_beamlets"].value) == int((f > 0).sum())
    got, uids, maps = H.read_fmaps(path)
    assert uids == [0, 1] and all(np.array_equal(m, f) for m, f in zip(maps, fmaps))


def test_reduced_columns_stream_into_the_coefficient_file(tmp_path):
    from tests.test_gpu_roi import _rois
    w = W.small(n=32, voxel=0.5, gantry=215.0, couch=-45.0, hetero=True, fmap=(8, 8), beamlet=(1.0, 1.0),
                weights="rand")
    w.spec.kernel_extent = 1.0
    rois = _rois(w)
    caps = [int((m == 1).sum()) for m, _, _ in rois]
    path = str(tmp_path / "reduced.h5")
    print("[case] h5 sink, ROI-reduced", flush=True)
    with Engine(w.spec) as eng:
        eng.set_roi_masks(rois)
        cols = _all_columns(eng, w.beams, w.fmaps)
        with H.CoeffFile(path, w.spec.struct(), nbeams=1, rois=(["PTV", "OAR", "SLAB"], caps)) as f:
            eng.beamlet_dose_h5(w.beams, w.fmaps, f, uid_base=5)
    root = _check_file(path, cols, 5, reduced=True)
    A, col_idx = HW.coefficient_matrix(root)
    assert A.shape[0] == sum(caps) and all(b == 5 for b, _ in col_idx)


def test_beam_dose_volume_file(tmp_path):
    w = W.small(n=32, voxel=0.5, gantry=30.0, hetero=True)
    print("[case] Dose.h5", flush=True)
    with Engine(w.spec) as eng:
        d = eng.beam_dose(w.beams, w.fmaps)
    path = str(tmp_path / "Dose.h5")
    H.write_volume(path, d, start=w.spec.struct().start[:], voxel=w.spec.struct().voxel[:])
    got, start, voxel = H.read_volume(path, "volume")
    assert np.array_equal(got.view(np.uint32), d.view(np.uint32)) and d.max() > 0
