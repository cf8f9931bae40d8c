"""Deterministic synthetic workloads: phantoms, beam sets and fluence maps.

Every generator is seeded (numpy PCG64) and the seed is part of the returned description, so a
run can be repeated exactly.  Coordinates are in cm with the volume centred on the origin;
`start` is the centre of voxel 0, as in the reference's omni-header (calc_header.cpp:99-101).

PTV projection masks restate the 3x3 sub-ray test of cudaRaycast
(dosecalc-preprocess/kernels.cuh:139-224): a beamlet is open when any of its centre/edge/corner
rays, cast from the source through the iso-plane bixel point and beyond, hits the PTV.  The PTVs
here are analytic ellipsoids, so the hit test is an exact ray/ellipsoid intersection.
"""
from __future__ import annotations

import dataclasses
import math
import os

import numpy as np

from . import engine

SEED_WEIGHTS = 20261202
SEED_ANGLES = 1162
SEED_ARRIVALS = 200
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


# ---- rotations in double (same convention as the library's dm_beam_inv) ----------------------
def _axis(ax, t):
    x, y, z = ax
    s, c = math.sin(t), math.cos(t)
    k = 1 - c
    return np.array([[c + k * x * x, k * x * y - s * z, k * x * z + s * y],
                     [k * y * x + s * z, c + k * y * y, k * y * z - s * x],
                     [k * z * x - s * y, k * z * y + s * x, c + k * z * z]])


def beam_fwd(theta, phi, coll):
    couch = _axis((math.sin(-phi), 0.0, math.cos(-phi)), -theta)
    roll = _axis((0.0, 1.0, 0.0), phi + coll)
    return roll @ couch


def beam_inv(theta, phi, coll):
    return beam_fwd(theta, phi, coll).T


# ---- phantoms ------------------------------------------------------------------------------------
@dataclasses.dataclass
class Phantom:
    density: np.ndarray  # (nz, ny, nx)
    voxel: float
    ptv: tuple | None = None  # (centre xyz, semi-axes xyz) ellipsoid in cm

    @property
    def start(self):
        nz, ny, nx = self.density.shape
        return tuple(-(n - 1) / 2.0 * self.voxel for n in (nx, ny, nz))

    def coords(self):
        nz, ny, nx = self.density.shape
        sx, sy, sz = self.start
        x = sx + np.arange(nx) * self.voxel
        y = sy + np.arange(ny) * self.voxel
        z = sz + np.arange(nz) * self.voxel
        return x, y, z


def water_cube(n: int = 128, voxel: float = 0.25, density: float = 1.0) -> Phantom:
    return Phantom(np.full((n, n, n), density, np.float32), voxel)


def thorax(nx=200, ny=200, nz=160, voxel=0.2) -> Phantom:
    """Water ellipse body (semi-axes 18 x 12 cm in x-y, all z) in air (0.001 g/cc), a lung slab
    of 0.25 g/cc (x in [-13,-3], |y| <= 7), a spine cylinder of 1.85 g/cc (r 1.5 cm at y = +8) and
    a spherical PTV of radius 3 cm at the origin (mask only)."""
    ph = Phantom(np.zeros((nz, ny, nx), np.float32), voxel, ptv=((0.0, 0.0, 0.0), (3.0, 3.0, 3.0)))
    x, y, z = ph.coords()
    X, Y = np.meshgrid(x, y, indexing="xy")  # (ny, nx)
    sl = np.full((ny, nx), 0.001, np.float32)
    body = (X / 18.0) ** 2 + (Y / 12.0) ** 2 <= 1.0
    sl[body] = 1.0
    lung = body & (X >= -13.0) & (X <= -3.0) & (np.abs(Y) <= 7.0)
    sl[lung] = 0.25
    spine = (X ** 2 + (Y - 8.0) ** 2) <= 1.5 ** 2
    sl[spine] = 1.85
    ph.density[:] = sl[None, :, :]
    return ph


def water_with_ptv(n=128, voxel=0.25, semi=(7.0, 4.8, 1.0)) -> Phantom:
    ph = water_cube(n, voxel)
    ph.ptv = ((0.0, 0.0, 0.0), tuple(semi))
    return ph


def ptv_mask(ph: Phantom) -> np.ndarray:
    (cx, cy, cz), (ax, ay, az) = ph.ptv
    x, y, z = ph.coords()
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return (((X - cx) / ax) ** 2 + ((Y - cy) / ay) ** 2 + ((Z - cz) / az) ** 2) <= 1.0


# ---- fluence maps --------------------------------------------------------------------------------
def projection_mask(beam, ptv, fmap_size, beamlet_size) -> np.ndarray:
    """1.0 where any of the beamlet's 3x3 sub-rays intersects the PTV ellipsoid, else 0."""
    (cx, cy, cz), (ax, ay, az) = ptv
    fx, fy = fmap_size
    bx, by = beamlet_size
    R = beam_inv(float(beam.azimuth), float(beam.zenith), float(beam.coll))
    src = np.array(beam.source[:], np.float64)
    iso = np.array(beam.isocenter[:], np.float64)
    ix, iz = np.meshgrid(np.arange(fx), np.arange(fy), indexing="xy")  # (fy, fx)
    out = np.zeros((fy, fx), bool)
    semi = np.array([ax, ay, az])
    c = np.array([cx, cy, cz])
    for sx in (-1, 0, 1):
        for sz in (-1, 0, 1):
            u = bx * (ix + 0.5 * sx - fx * 0.5 + 0.5)
            w = by * (iz + 0.5 * sz - fy * 0.5 + 0.5)
            p = np.stack([u, np.zeros_like(u), w], -1) @ R.T + iso  # bixel centre in PVCS
            d = 3.0 * (p - src)                                      # ray src -> 3x beyond
            o = (src - c) / semi
            dd = d / semi
            A = (dd * dd).sum(-1)
            B = 2 * (o * dd).sum(-1)
            Cq = (o * o).sum() - 1.0
            disc = B * B - 4 * A * Cq
            ok = disc >= 0
            sq = np.sqrt(np.where(ok, disc, 0.0))
            t0 = (-B - sq) / (2 * A)
            t1 = (-B + sq) / (2 * A)
            out |= ok & (t1 >= 0) & (t0 <= 1)
    return out.astype(np.float32)


# ---- beam sets -----------------------------------------------------------------------------------
def coplanar(n: int):
    return [(360.0 * i / n, 0.0) for i in range(n)]


def sphere_angles(n: int, seed: int = SEED_ANGLES):
    """n (gantry, couch) pairs uniform on the sphere, couch restricted to [-90, 90] degrees."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        gantry = math.degrees(math.acos(max(-1.0, min(1.0, -v[1]))))
        couch = math.degrees(math.atan2(v[2], v[0]))
        if couch > 90:
            couch -= 180
            gantry = -gantry
        elif couch < -90:
            couch += 180
            gantry = -gantry
        out.append((round(gantry % 360.0, 3), round(couch, 3)))
    return out


@dataclasses.dataclass
class Workload:
    name: str
    spec: engine.ProblemSpec
    beams: list
    fmaps: list
    seed: int | None = None
    phantom: Phantom | None = None

    @property
    def active_beamlets(self) -> int:
        return int(sum(int((f > 0).sum()) for f in self.fmaps))


def problem(ph: Phantom, nphi=8, ntheta=8, spacing=None, penumbra=1.0, kernel_extent=2.0, ss_factor=1,
            max_rev_size=(0, 0, 0), tables="spec_6mv") -> engine.ProblemSpec:
    t = engine.load_tables(tables)
    kern, th, phi = engine.cumulative_kernel(t, nphi, ntheta)
    sp = spacing or ph.voxel
    return engine.ProblemSpec(density=ph.density, voxel=(ph.voxel,) * 3, start=ph.start, kernel=kern, radii=t.radii,
                              theta=th, phi=phi, spectrum=t.spectrum, nphi=nphi, ntheta=ntheta, rev_latspacing=sp,
                              rev_longspacing=sp, penumbra=penumbra, kernel_extent=kernel_extent, ss_factor=ss_factor,
                              max_rev_size=max_rev_size)


def cfg1() -> Workload:
    """128^3 water at 0.25 cm, one beam at gantry 0, 20x20 x 0.5 cm fluence of 1.0, 64 cones."""
    ph = water_cube(128, 0.25)
    spec = problem(ph, nphi=8, ntheta=8)
    b = engine.make_beam(0.0, 0.0, 0.0, iso=(0, 0, 0), sad=100.0, fmap_size=(20, 20), beamlet_size=(0.5, 0.5))
    return Workload("cfg1-water128-1beam", spec, [b], [np.ones((20, 20), np.float32)], None, ph)


def cfg2(seed: int = SEED_WEIGHTS, nbeams: int = 20) -> Workload:
    """200x200x160 thorax at 0.2 cm, 20 coplanar beams, 40x40 x 0.5 cm maps of U[0,1] weights
    masked to the PTV projection, 64 cones."""
    ph = thorax()
    spec = problem(ph, nphi=8, ntheta=8)
    rng = np.random.default_rng(seed)
    beams, fmaps = [], []
    for g, c in coplanar(nbeams):
        b = engine.make_beam(g, c, 0.0, iso=ph.ptv[0], sad=100.0, fmap_size=(40, 40), beamlet_size=(0.5, 0.5))
        w = rng.uniform(0.0, 1.0, (40, 40)).astype(np.float32)
        m = projection_mask(b, ph.ptv, (40, 40), (0.5, 0.5))
        beams.append(b)
        fmaps.append((w * m).astype(np.float32))
    return Workload("cfg2-thorax200x200x160-20beams", spec, beams, fmaps, seed, ph)


def cfg3(seed: int = SEED_ANGLES, nbeams: int = 50, kernel_extent: float = 1.0) -> Workload:
    """128^3 water at 0.25 cm with an ellipsoidal PTV (7.0 x 4.8 x 1.0 cm semi-axes), 50
    non-coplanar beams on the sphere, 0.5 cm beamlets open over the PTV projection."""
    ph = water_with_ptv()
    spec = problem(ph, nphi=8, ntheta=8, kernel_extent=kernel_extent, max_rev_size=(800, 800, 800))
    beams, fmaps = [], []
    for g, c in sphere_angles(nbeams, seed):
        b = engine.make_beam(g, c, 0.0, iso=ph.ptv[0], sad=100.0, fmap_size=(40, 40), beamlet_size=(0.5, 0.5))
        m = projection_mask(b, ph.ptv, (40, 40), (0.5, 0.5))
        beams.append(b)
        fmaps.append(m)
    return Workload("cfg3-water128-ptv-50beams", spec, beams, fmaps, seed, ph)


def cfg5(nbeams: int | None = None, stride: int = 1, indices=None, kernel_extent: float = 1.0) -> Workload:
    """Full 4pi beamlet precompute: the cfg3 phantom, the 1162 beams of the 4pi candidate list (the
    reference's CCCS/tests/input_files/beamlist_4pi.txt) aimed at the PTV centre, 64x64 maps of
    0.25 cm beamlets open over the PTV projection, 96 cones (12 zenith x 8 azimuth).  `stride`
    and `nbeams` take every stride-th beam of the list, up to nbeams of them, and `indices` names
    list entries instead (bounded samples of the same workload)."""
    ph = water_with_ptv()
    spec = problem(ph, nphi=8, ntheta=12, kernel_extent=kernel_extent, max_rev_size=(800, 800, 800))
    cands = engine.load_beam_list(os.path.join(DATA, "beamlist_4pi.txt"))
    cands = [cands[i] for i in indices] if indices is not None else cands[::stride]
    if nbeams is not None:
        cands = cands[:nbeams]
    beams, fmaps = [], []
    for c in cands:
        b = engine.retarget_beam(c, iso=ph.ptv[0], sad=100.0, fmap_size=(64, 64), beamlet_size=(0.25, 0.25))
        beams.append(b)
        fmaps.append(projection_mask(b, ph.ptv, (64, 64), (0.25, 0.25)))
    return Workload(f"cfg5-water128-ptv-4pi-{len(beams)}beams", spec, beams, fmaps, None, ph)


def small(n=32, voxel=0.5, gantry=0.0, couch=0.0, fmap=(8, 8), beamlet=(1.0, 1.0), nphi=4, ntheta=4,
          hetero=False, ss_factor=1, bbox=None, seed=7, weights="ones") -> Workload:
    """Small cases for parity tests: n^3 water (optionally with a lung block and a bone rod)."""
    ph = water_cube(n, voxel)
    if hetero:
        x, y, z = ph.coords()
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        ph.density[(X < -1.0) & (np.abs(Y) < n * voxel / 4)] = 0.25
        ph.density[((X - 2.0) ** 2 + (Y - 1.0) ** 2) < 1.5] = 1.85
        ph.density[Y > n * voxel / 2 - 2 * voxel] = 0.001
    spec = problem(ph, nphi=nphi, ntheta=ntheta, ss_factor=ss_factor, max_rev_size=(800, 800, 800))
    if bbox is not None:
        spec.bbox_start, spec.bbox_size = bbox
    b = engine.make_beam(gantry, couch, 0.0, iso=(0.0, 0.0, 0.0), sad=100.0, fmap_size=fmap, beamlet_size=beamlet)
    if weights == "ones":
        f = np.ones((fmap[1], fmap[0]), np.float32)
    else:
        f = np.random.default_rng(seed).uniform(0.0, 1.0, (fmap[1], fmap[0])).astype(np.float32)
    return Workload(f"small{n}-g{gantry}-c{couch}", spec, [b], [f], seed, ph)

@dataclasses.dataclass
class ServingWorkload:
    name: str
    spec: engine.ProblemSpec
    beams: list
    fmaps: list
    arrivals: np.ndarray  # seconds from the start of the stream
    rate: float
    seed: int


def cfg4(seed: int = SEED_ARRIVALS, nreq: int = 300, rate: float = 200.0) -> ServingWorkload:
    """Serving: the thorax tissue classes on a 256x256x180 phantom at 2 mm; Poisson arrivals at
    `rate` requests/s; each request is one beam's dose, the beam drawn uniformly from the 1162-beam
    4pi candidate list (the reference's CCCS/tests/input_files/beamlist_4pi.txt), with an n x n
    fluence map of 0.5 cm beamlets, n ~ U{20..64}, weights U[0,1]."""
    ph = thorax(256, 256, 180, 0.2)
    spec = problem(ph, nphi=8, ntheta=8)
    cands = engine.load_beam_list(os.path.join(DATA, "beamlist_4pi.txt"))
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    beams, fmaps = [], []
    for _ in range(nreq):
        b = cands[int(rng.integers(len(cands)))].copy()
        n = int(rng.integers(20, 65))
        b.fmap_size[:] = (n, n)
        b.beamlet_size[:] = (0.5, 0.5)
        beams.append(b)
        fmaps.append(rng.uniform(0.0, 1.0, (n, n)).astype(np.float32))
    return ServingWorkload("cfg4-thorax256x256x180-serving", spec, beams, fmaps, arrivals, rate, seed)
