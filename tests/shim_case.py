"""Writes a case directory for the shim test program (tests/shim/shim_driver.cpp): case.txt with
the scalars and one `beam` line per beam, and raw float32 tables and fluence maps."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVERS = {k: os.path.join(HERE, "shim", "_build", f"shim_driver_{k}") for k in ("beam", "beamlet")}


def write_case(d, spec, beams, fmaps, uids=None, threshold=1e-4):
    os.makedirs(d, exist_ok=True)
    uids = list(range(len(beams))) if uids is None else list(uids)

    def raw(name, a):
        np.ascontiguousarray(a, np.float32).tofile(os.path.join(d, name))

    raw("density", spec.density)
    raw("kernel", spec.kernel)
    raw("radii", spec.radii)
    raw("theta", spec.theta)
    raw("phi", spec.phi)
    raw("spectrum", spec.spectrum)
    lines = [f"size {spec.size[0]} {spec.size[1]} {spec.size[2]}",
             "voxel " + " ".join(repr(float(v)) for v in spec.voxel),
             "start " + " ".join(repr(float(v)) for v in spec.start),
             "bbox_start " + " ".join(str(int(v)) for v in spec.bbox_start),
             "bbox_size " + " ".join(str(int(v)) for v in spec.bbox_size),
             "max_rev_size " + " ".join(str(int(v)) for v in spec.max_rev_size),
             f"rev_latspacing {float(spec.rev_latspacing)!r}", f"rev_longspacing {float(spec.rev_longspacing)!r}",
             f"kernel_extent {float(spec.kernel_extent)!r}", f"penumbra {float(spec.penumbra)!r}",
             f"ss_factor {int(spec.ss_factor)}", f"nphi {int(spec.nphi)}", f"ntheta {int(spec.ntheta)}",
             f"nradii {int(spec.radii.shape[0])}", f"nspec {int(spec.spectrum.shape[1])}",
             f"beamhard_correct {float(spec.beamhard_correct)!r}", f"threshold {float(threshold)!r}"]
    for i, (b, f) in enumerate(zip(beams, fmaps)):
        # float32 fields printed exactly (repr of the double holding the float)
        vals = [b.azimuth, b.zenith, b.coll, *b.source, *b.direction, *b.isocenter, b.sad]
        lines.append("beam " + " ".join([str(int(uids[i]))] + [repr(float(np.float32(v))) for v in vals] +
                                        [str(int(b.fmap_size[0])), str(int(b.fmap_size[1]))] +
                                        [repr(float(np.float32(v))) for v in b.beamlet_size]))
        raw(f"fmap_{i}", f)
    with open(os.path.join(d, "case.txt"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
