"""Pins the oracle's HEALPix RING primitives (HEALPix C++ 3.00 semantics;
third-party, absent from the reference tree) by closed forms and brute force:
  - ring pixel counts and first pixels, total = 12 nside^2
  - pixel centres: unit vectors, ring z from ring2z, equal phi spacing
  - ang2pix(pix2vec(p)) == p for every pixel (nside <= 64) and samples at 1024
  - equal areas: uniform points land evenly (chi-square)
  - query_disc == {p : angle(centre_p, v) <= R} over all pixels (nside <= 64),
    random centres and radii incl. polar caps, phi wrap and radius >= pi
"""
import math

import numpy as np
import pytest


def _all_vecs(orc, nside):
    return orc.pix2vec(nside, np.arange(12 * nside * nside))


def test_ring_layout(orc):
    import ctypes as C
    L = orc.lib()
    for nside in (1, 2, 4, 8, 64, 1024):
        total = 0
        for ring in range(1, 4 * nside):
            sp, rp, sh = C.c_int64(), C.c_int64(), C.c_int()
            L.orc_ring_info.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int)]
            L.orc_ring_info(nside, ring, C.byref(sp), C.byref(rp), C.byref(sh))
            expect = 4 * min(ring, nside, 4 * nside - ring)
            assert rp.value == expect
            assert sp.value == total
            total += rp.value
        assert total == 12 * nside * nside


@pytest.mark.parametrize("nside", [1, 2, 8, 32])
def test_pixel_centres(orc, nside):
    v = _all_vecs(orc, nside)
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-15)
    p = 0
    for ring in range(1, 4 * nside):
        nr = 4 * min(ring, nside, 4 * nside - ring)
        z = orc.lib().orc_ring2z(nside, ring)
        seg = v[p:p + nr]
        assert np.allclose(seg[:, 2], z, atol=1e-15)
        phi = np.mod(np.arctan2(seg[:, 1], seg[:, 0]), 2 * np.pi)
        d = np.diff(phi)
        assert np.allclose(d, 2 * np.pi / nr, atol=1e-12)
        # first centre at 0 or half a step (shifted rings)
        assert min(abs(phi[0]), abs(phi[0] - np.pi / nr)) < 1e-12
        p += nr


@pytest.mark.parametrize("nside", [1, 4, 16, 64])
def test_ang2pix_round_trip_all_pixels(orc, nside):
    v = _all_vecs(orc, nside)
    th = np.arccos(np.clip(v[:, 2], -1, 1))
    ph = np.mod(np.arctan2(v[:, 1], v[:, 0]), 2 * np.pi)
    got = [orc.ang2pix(nside, t, f) for t, f in zip(th, ph)]
    assert got == list(range(len(v)))


def test_ang2pix_round_trip_samples_high_nside(orc):
    nside = 1024
    rng = np.random.default_rng(3)
    pix = rng.integers(0, 12 * nside * nside, 3000)
    v = orc.pix2vec(nside, pix)
    th = np.arccos(np.clip(v[:, 2], -1, 1))
    ph = np.mod(np.arctan2(v[:, 1], v[:, 0]), 2 * np.pi)
    assert [orc.ang2pix(nside, t, f) for t, f in zip(th, ph)] == pix.tolist()


def test_ang2pix_equal_area(orc):
    nside = 4
    npix = 12 * nside * nside
    rng = np.random.default_rng(11)
    n = 192000
    z = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    th = np.arccos(z)
    counts = np.bincount([orc.ang2pix(nside, t, f) for t, f in zip(th, ph)], minlength=npix)
    exp = n / npix
    chi2 = ((counts - exp) ** 2 / exp).sum()
    # chi-square with 191 dof: mean 191, sd ~19.5
    assert chi2 < 191 + 6 * 19.5


def test_ang2pix_nearest_centre_bound(orc):
    """A point lies within the maximum pixel radius of its pixel's centre."""
    nside = 32
    rng = np.random.default_rng(5)
    z = rng.uniform(-1, 1, 5000)
    ph = rng.uniform(0, 2 * np.pi, 5000)
    th = np.arccos(z)
    pts = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), z], 1)
    pix = np.array([orc.ang2pix(nside, t, f) for t, f in zip(th, ph)])
    c = orc.pix2vec(nside, pix)
    ang = np.arctan2(np.linalg.norm(np.cross(pts, c), axis=1), (pts * c).sum(1))
    # HEALPix max pixel radius is below 1.4 x the mean pixel size sqrt(4pi/npix)
    assert ang.max() < 1.4 * math.sqrt(4 * math.pi / (12 * nside * nside))


def _brute(vecs, th, ph, R):
    v = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    ang = np.arctan2(np.linalg.norm(np.cross(vecs, v), axis=1), vecs @ v)
    return ang, np.nonzero(ang <= R)[0]


@pytest.mark.parametrize("nside", [1, 2, 8, 32, 64])
def test_query_disc_brute_force(orc, nside):
    vecs = _all_vecs(orc, nside)
    rng = np.random.default_rng(nside)
    cases = []
    for _ in range(150):
        cases.append((math.acos(rng.uniform(-1, 1)), rng.uniform(0, 2 * math.pi), rng.uniform(1e-3, 1.5)))
    # polar caps, the phi = 0 seam, huge and tiny radii
    cases += [(1e-3, 0.3, 0.2), (math.pi - 2e-3, 5.0, 0.4), (0.7, 1e-9, 0.3), (1.2, 2 * math.pi - 1e-9, 0.25),
              (2.0, 3.0, 3.2), (0.5, 0.5, 1e-4), (math.pi / 2, math.pi, 2.9), (0.05, 6.0, 0.06)]
    for th, ph, R in cases:
        got = orc.query_disc(nside, th, ph, R)
        ang, want = _brute(vecs, th, ph, R)
        if not np.array_equal(got, want):
            # only pixels within rounding of the edge may differ
            diff = np.setxor1d(got, want)
            assert np.all(np.abs(ang[diff] - R) < 1e-12), (th, ph, R, diff)
        assert np.all(np.diff(got) > 0)  # ascending, no duplicates
