"""Host-side checks of the synthetic workloads the GPU tests and the bench use (no device work):
the full 4pi precompute settings take their beams from the reference's 1162-beam list unchanged
in angle, aim them at the PTV centre, and open exactly the beamlets whose sub-rays hit the PTV."""
import math
import os

import numpy as np

from endtoend_amd import engine
from endtoend_amd import workloads as W


def test_4pi_settings_and_beams():
    w = W.cfg5(indices=(0, 97, 873))
    assert (w.spec.nphi, w.spec.ntheta) == (8, 12)          # 96 cones, 48 bidirectional lines
    assert w.spec.nphi * w.spec.ntheta // 2 == 48
    cands = engine.load_beam_list(os.path.join(W.DATA, "beamlist_4pi.txt"))
    assert len(cands) == 1162
    for b, i in zip(w.beams, (0, 97, 873)):
        c = cands[i]
        assert (b.azimuth, b.zenith, b.coll) == (c.azimuth, c.zenith, c.coll)
        assert np.allclose(list(b.isocenter), W.water_with_ptv().ptv[0])
        assert b.sad == 100.0 and list(b.fmap_size) == [64, 64]
        assert np.allclose(list(b.beamlet_size), [0.25, 0.25])
        # source at SAD from the isocentre, opposite the beam direction
        src, iso = np.array(list(b.source)), np.array(list(b.isocenter))
        assert math.isclose(float(np.linalg.norm(iso - src)), 100.0, rel_tol=1e-5)


def test_4pi_masks_are_ptv_projections():
    w = W.cfg5(indices=(0, 97))
    ph = W.water_with_ptv()
    for b, f in zip(w.beams, w.fmaps):
        assert f.shape == (64, 64) and f.dtype == np.float32
        assert set(np.unique(f)) <= {0.0, 1.0}
        n = int(f.sum())
        # the PTV (7.0 x 4.8 x 1.0 cm semi-axes) seen through 2.5 mm beamlets: between its
        # smallest and largest silhouette, plus a rim of partly covered beamlets
        assert math.pi * 1.0 * 4.8 / 0.0625 * 0.8 < n < math.pi * 7.0 * 4.8 / 0.0625 * 1.6
        assert not (f[0].any() or f[-1].any() or f[:, 0].any() or f[:, -1].any())  # field covers it
        # the same mask as the 0.5 cm map's, refined: every open 2.5 mm beamlet lies in an open 5 mm one
        coarse = W.projection_mask(b, ph.ptv, (32, 32), (0.5, 0.5))
        assert np.all(coarse.repeat(2, 0).repeat(2, 1)[f > 0] > 0)
