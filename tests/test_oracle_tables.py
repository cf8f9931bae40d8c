"""Pins the kernel/spectrum tables (A9) and the beam orientation (A1) of the oracle and the product.

A9 is pinned against a float64 numpy re-derivation straight from the reference's text tables
(poly kernel = fluence*energy*mu weighted mean of the mono kernels per category, summed over
categories), against the aggregation / cumulative-sum invariants, and product == oracle bitwise.
A1 restates the round trip of CCCS/unit_tests/beam.cpp:16-67.
"""
import ctypes as C
import os

import numpy as np
import pytest

from endtoend_amd import abi
from endtoend_amd import engine as E
from oracle import oracle as O
from tests.conftest import REFERENCE_DATA

needs_ref = pytest.mark.skipif(not os.path.isdir(REFERENCE_DATA), reason="reference data tree not mounted")
NTHETAS = (2, 4, 6, 8, 12, 16, 24, 48)


def _product_tables(nphi, ntheta):
    lib = abi.load()
    kern = np.zeros(ntheta * 24, np.float32)
    radii = np.zeros(24, np.float32)
    th = np.zeros(nphi * ntheta // 2, np.float32)
    ph = np.zeros(nphi * ntheta // 2, np.float32)
    spec = np.zeros(120, np.float32)
    n = C.c_int()
    d = REFERENCE_DATA
    abi.check(lib.cccs_build_kernel_from_files(os.path.join(d, "spectra", "spec_6mv.spec").encode(),
                                               os.path.join(d, "dsa").encode(),
                                               os.path.join(d, "dsa", "radii.dat").encode(), nphi, ntheta,
                                               abi.fptr(kern), abi.fptr(radii), abi.fptr(th), abi.fptr(ph),
                                               abi.fptr(spec), C.byref(n)), "build_kernel")
    return kern.reshape(ntheta, 24), radii, th, ph, spec[:4 * n.value].reshape(4, n.value)


@needs_ref
@pytest.mark.parametrize("ntheta", NTHETAS)
def test_product_and_oracle_tables_bitwise(ntheta):
    o = O.kernel_tables(REFERENCE_DATA, "spec_6mv", 6, ntheta)
    k, r, th, ph, sp = _product_tables(6, ntheta)
    for a, b in ((k, o["kernel"]), (r, o["radii"]), (th, o["theta"]), (ph, o["phi"]), (sp, o["spectrum"])):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@needs_ref
def test_poly_kernel_against_numpy_from_text_tables():
    d = REFERENCE_DATA
    rows = [ln.split() for ln in open(os.path.join(d, "spectra", "spec_6mv.spec")) if ln.strip()]
    en = np.array([float(r[0]) for r in rows])
    flu = np.array([float(r[1]) for r in rows])
    mu = np.array([float(r[2]) for r in rows])
    flu = flu / flu.sum()
    w = flu * en * mu
    acc = np.zeros((5, 48, 24))
    for e, r in enumerate(rows):
        vals = np.loadtxt(os.path.join(d, "dsa", r[4])).reshape(48, 24, 7, 2)[..., 0]  # (theta, radius, entry)
        acc += w[e] * np.moveaxis(vals[..., :5], -1, 0)
    poly = (acc / w.sum()).sum(0)
    o = O.kernel_tables(d, "spec_6mv", 8, 8)
    np.testing.assert_allclose(o["poly"], poly, rtol=2e-5, atol=1e-9)
    # the shipped table file carries the same poly kernel, radii and spectrum bit for bit
    t = E.load_tables("spec_6mv")
    assert np.array_equal(t.poly.view(np.uint32), o["poly"].view(np.uint32))
    assert np.array_equal(t.radii.view(np.uint32), o["radii"].view(np.uint32))
    assert np.array_equal(t.spectrum.view(np.uint32), o["spectrum"].view(np.uint32))
    np.testing.assert_allclose(o["radii"], np.loadtxt(os.path.join(d, "dsa", "radii.dat")), rtol=1e-7)


@pytest.mark.parametrize("ntheta", NTHETAS)
def test_cumulative_kernel_invariants(ntheta):
    t = E.load_tables()
    poly = t.poly.astype(np.float64)
    nphi = 4
    k, th, ph = O.cumulative(t.poly, nphi, ntheta)
    g = 48 // ntheta
    agg = poly.reshape(ntheta, g, 24).sum(1)
    np.testing.assert_allclose(k[:, 0], agg[:, 0], rtol=1e-6)                 # first radius = aggregated bins
    np.testing.assert_allclose(k, np.cumsum(agg, 1), rtol=2e-5, atol=1e-9)   # cumulative along radius
    assert np.all(np.diff(k, axis=1) >= 0)
    np.testing.assert_allclose(k[:, -1].sum(), poly.sum(), rtol=2e-5)         # nothing lost in aggregation
    # cone angles: theta at bin mid-points, phi uniform; ray = p*ntheta/2 + t
    tt = np.tile(np.arange(ntheta // 2), nphi)
    pp = np.repeat(np.arange(nphi), ntheta // 2)
    np.testing.assert_allclose(th, (tt + 0.5) * 180.0 / ntheta, rtol=1e-6)
    np.testing.assert_allclose(ph, (pp + 0.5) * 360.0 / nphi, rtol=1e-6)
    # product implementation agrees bit for bit
    kp, thp, php = E.cumulative_kernel(t, nphi, ntheta)
    assert np.array_equal(kp.view(np.uint32), k.view(np.uint32))
    assert np.array_equal(thp.view(np.uint32), th.view(np.uint32))
    assert np.array_equal(php.view(np.uint32), ph.view(np.uint32))


def test_spectrum_normalised():
    t = E.load_tables()
    np.testing.assert_allclose(t.spectrum[0].sum(), 1.0, rtol=1e-6)
    assert np.all(np.diff(t.spectrum[1]) > 0) and t.spectrum[1][0] == np.float32(0.2) and t.spectrum[1][-1] == 6.0
    assert np.all(t.spectrum[3] > t.spectrum[2])  # mu > mu_en


def _beam(az_deg, zen_deg):
    b = abi.Beam()
    pi = np.float32(np.pi)
    b.azimuth = np.float32(np.float64(np.float32(az_deg) * pi) / 180.0)
    b.zenith = np.float32(np.float64(np.float32(zen_deg) * pi) / 180.0)
    b.sad = 100.0
    return b


ANGLES = [(45.0 * i, 0.0) for i in range(12)] + [(45.0 * i, 45.0 * j) for j in range(12) for i in range(12)]


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_reference_beam_round_trip(impl):
    """CCCS/unit_tests/beam.cpp:16-67: angles -> direction -> angles -> direction is stable."""
    lib = abi.load()

    def reconf(b, orient):
        if impl == "oracle":
            return O.reconfigure(b, orient)
        c = b.copy()
        abi.check(lib.cccs_beam_reconfigure(C.byref(c), orient), "reconfigure")
        return c

    for az, zen in ANGLES:
        b1 = reconf(_beam(az, zen), 0)
        b2 = abi.Beam()
        b2.direction[:] = list(b1.direction)
        b2.sad = 100.0
        b2 = reconf(b2, 1)
        b3 = _beam(0, 0)
        b3.azimuth, b3.zenith = b2.azimuth, b2.zenith
        b3 = reconf(b3, 0)
        assert np.all(np.abs(np.array(b1.direction[:]) - np.array(b3.direction[:])) <= 1e-6), (az, zen)


def test_product_and_oracle_beam_sources_agree():
    lib = abi.load()
    for az, zen in ANGLES:
        o = O.reconfigure(_beam(az, zen), 0)
        p = _beam(az, zen)
        abi.check(lib.cccs_beam_reconfigure(C.byref(p), 0), "reconfigure")
        # the same float Rodrigues rotation with libm sincosf: bit for bit
        assert list(p.source) == list(o.source) and list(p.direction) == list(o.direction)
        # direction is the unit vector from source to isocentre (BEAM::calc_dir_from_source)
        s = np.array(o.source[:], np.float64)
        np.testing.assert_allclose(-s / np.linalg.norm(s), list(o.direction), atol=1e-6)
