"""The reference's file formats through the library's host functions (no device needed): beam
lists (DoseCalcIO/beam.cpp:369-576), the omni header (calc_header.cpp:12-161) and raw float32
volumes (fluence.cpp:19-44, binary_io).  The 1162-beam 4pi list is the reference's own test
input (CCCS/tests/input_files/beamlist_4pi.txt, kept as tests/golden/beamlist_4pi.txt)."""
import os

import numpy as np
import pytest

from endtoend_amd import abi
from endtoend_amd import engine as E
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PI = np.float32(3.141592654)


def _rad(deg):
    return np.float32(np.float32(deg) * PI) / np.float32(180.0)


def test_4pi_beam_list_parses_like_the_reference():
    path = os.path.join(GOLDEN, "beamlist_4pi.txt")
    beams = E.load_beam_list(path)
    rows = []
    for line in open(path):
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        rows.append([float(t) for t in s.split()[:2]])
    assert len(beams) == len(rows) == 1162
    for b, (g, c) in zip(beams, rows):
        assert np.float32(b.azimuth) == _rad(g)
        # BEAM::reconfigure drops the couch angle at gantry 0 and 180 (beam.cpp:300-331)
        assert np.float32(b.zenith) == (np.float32(0) if g in (0.0, 180.0) else _rad(c))
        assert b.sad == 100.0 and b.coll == 0.0 and list(b.isocenter) == [0.0, 0.0, 0.0]
        # parse + reconfigure equals the oracle's reconfigure of the raw angles, bit for bit
        raw = abi.Beam()
        raw.azimuth, raw.zenith, raw.sad = _rad(g), _rad(c), 100.0
        o = O.reconfigure(raw, 0)
        assert (o.azimuth, o.zenith) == (b.azimuth, b.zenith)
        assert list(o.source) == list(b.source) and list(o.direction) == list(b.direction)


def test_beam_list_optional_fields_comments_and_missing_last_newline(tmp_path):
    p = tmp_path / "beams.txt"
    p.write_text("# comment line\n"
                 "\n"
                 "  90  0\n"
                 "0 0 15 iso: 1 2 3 sad: 80\n"
                 "   # indented comment\n"
                 "30 10 bogus 5\n"
                 "0 0 dir: 0 1 0 iso: 0 0 0\n"
                 "45 45")                                   # no trailing newline: dropped (beam.cpp:395)
    beams = E.load_beam_list(str(p))
    assert len(beams) == 4
    b0, b1, b2, b3 = beams
    assert np.float32(b0.azimuth) == _rad(90) and b0.sad == 100.0
    np.testing.assert_allclose(b0.source[:], [100.0, 0.0, 0.0], atol=1e-4)   # gantry 90: source at +x
    assert np.float32(b1.coll) == _rad(15) and list(b1.isocenter) == [1.0, 2.0, 3.0] and b1.sad == 80.0
    np.testing.assert_allclose(b1.source[:], [1.0, 2.0 - 80.0, 3.0], atol=1e-4)
    assert np.float32(b2.coll) == _rad(5)                   # an unknown token is skipped
    np.testing.assert_allclose(b3.direction[:], [0.0, 1.0, 0.0], atol=1e-7)
    np.testing.assert_allclose(b3.source[:], [0.0, -100.0, 0.0], atol=1e-4)
    # requested count stops early
    assert len(E.load_beam_list(str(p), requested=2)) == 2


def test_beam_list_rejects_a_short_line(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("10\n")
    with pytest.raises(abi.CCCSError) as e:
        E.load_beam_list(str(p))
    assert e.value.code == abi.CCCS_EINVAL


def test_missing_files_report_eio(tmp_path):
    with pytest.raises(abi.CCCSError) as e:
        E.load_beam_list(str(tmp_path / "nope.txt"))
    assert e.value.code == abi.CCCS_EIO
    with pytest.raises(abi.CCCSError) as e:
        E.read_f32(str(tmp_path / "nope.raw"), (4,))
    assert e.value.code == abi.CCCS_EIO


def test_omni_header_round_trip_and_bbox_clamp(tmp_path):
    p = abi.Problem()
    p.size[:] = (40, 30, 20)
    p.voxel[:] = (0.25, 0.25, 0.3)
    p.start[:] = (-4.875, -3.625, -2.85)
    p.bbox_start[:] = (5, 40, 2)            # y start beyond the volume: reset to 0
    p.bbox_size[:] = (50, 10, 10)           # x size beyond the volume: clamped to 35
    p.max_rev_size[:] = (500, 500, 500)
    p.rev_latspacing, p.rev_longspacing, p.kernel_extent = 0.25, 0.25, 2.0
    p.ss_factor, p.nphi, p.ntheta, p.nradii, p.penumbra = 3, 8, 8, 24, 1.0
    path = str(tmp_path / "omni-header.txt")
    E.write_omni_header(path, p, 7, "spec_6mv", "PTV", True)
    h = E.load_omni_header(path)
    q = h["problem"]
    assert list(q.size) == [40, 30, 20] and list(q.max_rev_size) == [500, 500, 500]
    np.testing.assert_allclose(q.voxel[:], [0.25, 0.25, 0.3])
    np.testing.assert_allclose(q.start[:], [-4.875, -3.625, -2.85])
    assert list(q.bbox_start) == [5, 0, 2] and list(q.bbox_size) == [35, 10, 10]
    assert (q.ss_factor, q.nphi, q.ntheta, q.nradii) == (3, 8, 8, 24)
    assert (h["beam_count"], h["beam_spec"], h["target_structure"], h["reduce"]) == (7, "spec_6mv", "PTV", True)


def test_omni_beam_list_in_the_writer_format(tmp_path):
    path = tmp_path / "omni_beam_list.txt"
    path.write_text(
        "2\n"
        "0 30.000000 0.000000 0.000000 100.000000 ptv unspec 1.000000 2.000000 3.000000 51.0 -84.6 3.0 "
        "angle 0.5 0.866 0.0 0.500000 0.500000 40 40 fmap-000000.raw\n"
        "1 90.000000 10.000000 5.000000 90.000000 manual in 0.0 0.0 0.0 90.0 0.0 0.0 direction -1.0 0.0 0.0 "
        "1.000000 0.250000 20 32 fmap-000001.raw\n"
        "\n#beam_id  azimuth ...\n")
    beams, names = E.load_omni_beam_list(str(path), 5)
    assert len(beams) == 2 and names == ["fmap-000000.raw", "fmap-000001.raw"]
    assert np.float32(beams[1].azimuth) == _rad(90) and np.float32(beams[1].coll) == _rad(5)
    assert list(beams[0].fmap_size) == [40, 40] and list(beams[1].fmap_size) == [20, 32]
    assert list(beams[1].beamlet_size) == [1.0, 0.25] and beams[1].sad == 90.0
    assert list(beams[0].isocenter) == [1.0, 2.0, 3.0]


def test_raw_volumes_round_trip_and_size_check(tmp_path):
    a = np.random.default_rng(1).random((5, 6, 7), dtype=np.float32)
    path = str(tmp_path / "dose.raw")
    E.write_f32(path, a)
    assert os.path.getsize(path) == a.nbytes
    np.testing.assert_array_equal(E.read_f32(path, a.shape), a)
    with pytest.raises(abi.CCCSError) as e:
        E.read_f32(path, (6, 6, 7))              # the fluence loader's exact-size check
    assert e.value.code == abi.CCCS_EIO
