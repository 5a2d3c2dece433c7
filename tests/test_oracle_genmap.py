"""Pins the oracle's map generator (restatement of src/genmap.cpp:22-175).

Checks that do not retype the ring walk:
  - an all-pixel brute force (numpy) of the generator's rules on small
    inputs: footprint = pixel centres within 2h/d by arctan2(|u x v|, u.v),
    weights exp(-0.5 (psi/r)^2/0.333), normalised, times flux; sub-pixel
    rule (2r < theta0, theta0 = acos(1 - dOmega/(2*3.141592))) and the
    fewer-than-2-pixels fallback to ang2pix;
  - the invariant sum(map) * dOmega == sum(flux of kept particles);
  - the mass / negative-density filter (src/genmap.cpp:52-58,82-85);
  - a particle on the observer frame's pole gives a map constant along rings;
  - the OpenMP variant equals the single-threaded one.
"""
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cfg(orc, ini="synth_obs8.ini", nside=None):
    p = orc.params_from_ini(os.path.join(GOLD, ini))
    c = orc.config_from_params(p)
    if nside:
        c.nside = nside
    return c


def _brute_map(orc, recs, cfg):
    nside = cfg.nside
    npix = 12 * nside * nside
    vecs = orc.pix2vec(nside, np.arange(npix))
    R = np.array(list(cfg.rotmatrix)).reshape(3, 3)
    opos = np.array(list(cfg.opos))
    dOmega = 4 * math.pi / npix
    theta0 = math.acos(1.0 - dOmega / (2.0 * 3.141592))
    m = np.zeros(npix)
    tot = 0.0
    ambiguous = 0
    for q in recs:
        if float(q["mass"]) >= cfg.mass_cut or q["dens"] < 0:
            continue
        dv = np.array([q["posx"], q["posy"], q["posz"]], dtype=np.float64) - opos
        d = math.sqrt((dv ** 2).sum())
        flux = cfg.flux_factor * float(q["dens"]) * float(q["mass"]) / (4.0 * math.pi * d * d)
        tot += flux
        t = R.T @ dv
        v = np.array([-t[0], -t[1], t[2]]) / d  # direction after the phi - pi shift
        th = math.acos(max(-1.0, min(1.0, v[2])))
        ph = math.atan2(v[1], v[0]) % (2 * math.pi)
        r = float(q["hsmooth"]) / d
        p0 = orc.ang2pix(nside, th, ph)
        if 2 * r < theta0:
            m[p0] += flux
            continue
        ang = np.arctan2(np.linalg.norm(np.cross(vecs, v), axis=1), vecs @ v)
        if np.any(np.abs(ang - 2 * r) < 1e-11):
            ambiguous += 1
        sel = ang <= 2 * r
        if sel.sum() < 2:
            m[p0] += flux
            continue
        w = np.exp(-0.5 * (ang[sel] / r) ** 2 / 0.333)
        m[sel] += w / w.sum() * flux
    return m / dOmega, tot, ambiguous


def _particles(orc, n, seed, hscale=1.0, spread=60.0):
    rng = np.random.default_rng(seed)
    recs = np.zeros(n, dtype=orc.PARTICLE_DTYPE)
    pos = rng.standard_normal((n, 3)) * spread
    recs["posx"], recs["posy"], recs["posz"] = pos.T.astype(np.float32)
    recs["mass"] = 4100.0
    recs["dens"] = rng.lognormal(0, 1, n).astype(np.float32)
    d = np.linalg.norm(pos - [8, 0, 0], axis=1)
    recs["hsmooth"] = (d * rng.uniform(0.01, 0.3, n) * hscale).astype(np.float32)
    return recs


@pytest.mark.parametrize("nside", [4, 16, 32])
def test_genmap_matches_brute_force(orc, nside):
    cfg = _cfg(orc, nside=nside)
    recs = _particles(orc, 300, nside)
    m, st = orc.genmap(recs, cfg)
    b, tot, amb = _brute_map(orc, recs, cfg)
    assert amb == 0
    assert np.allclose(m, b, rtol=1e-9, atol=1e-12 * np.abs(b).max())
    assert abs(st["total_flux"] - tot) <= 1e-12 * tot


def test_subpixel_and_fallback_rules(orc):
    """Tiny smoothing lengths: 2r < theta0 deposits at ang2pix; 2r just
    above theta0 often catches < 2 pixel centres and falls back."""
    nside = 16
    cfg = _cfg(orc, nside=nside)
    npix = 12 * nside * nside
    theta0 = math.acos(1.0 - 4 * math.pi / npix / (2.0 * 3.141592))
    recs = _particles(orc, 400, 99)
    d = np.linalg.norm(np.stack([recs["posx"], recs["posy"], recs["posz"]], 1).astype(np.float64) - [8, 0, 0], axis=1)
    f = np.random.default_rng(1).uniform(0.2, 1.6, len(recs))
    recs["hsmooth"] = (f * theta0 / 2 * d).astype(np.float32)
    m, st = orc.genmap(recs, cfg)
    b, tot, amb = _brute_map(orc, recs, cfg)
    assert amb == 0
    assert st["subpixel"] > 50 and st["fallback"] > 20 and st["extended"] > 20
    assert np.allclose(m, b, rtol=1e-9, atol=1e-12 * np.abs(b).max())


def test_total_flux_invariant(orc):
    cfg = _cfg(orc, nside=64)
    recs = _particles(orc, 2000, 5)
    m, st = orc.genmap(recs, cfg)
    dOmega = 4 * math.pi / m.size
    assert abs(m.sum() * dOmega - st["total_flux"]) <= 1e-12 * st["total_flux"]


def test_filter(orc):
    cfg = _cfg(orc, "synth_offaxis.ini", nside=16)  # masses 4100, 8200: cut = 4100 * 1.1
    recs = _particles(orc, 50, 3)
    recs["mass"][::3] = 8200.0
    recs["dens"][1::5] = -1.0
    _, st = orc.genmap(recs, cfg)
    keep = (recs["mass"].astype(np.float64) < cfg.mass_cut) & (recs["dens"] >= 0)
    assert st["kept"] == keep.sum()
    b, tot, _ = _brute_map(orc, recs, cfg)
    assert abs(st["total_flux"] - tot) <= 1e-12 * tot


def test_polar_particle_is_ring_symmetric(orc):
    """A particle straight along the observer frame's z axis (theta = 0)
    deposits the same value on every pixel of a ring."""
    cfg = _cfg(orc, nside=32)
    R = np.array(list(cfg.rotmatrix)).reshape(3, 3)
    # world offset whose rotated direction is +z: dv = R @ (0, 0, d)
    dv = R @ np.array([0.0, 0.0, 20.0])
    rec = np.zeros(1, dtype=orc.PARTICLE_DTYPE)
    rec["posx"], rec["posy"], rec["posz"] = (dv + [8, 0, 0]).astype(np.float32)
    rec["mass"], rec["dens"], rec["hsmooth"] = 4100.0, 1.0, 4.0
    m, st = orc.genmap(rec, cfg)
    assert st["extended"] == 1
    p = 0
    for ring in range(1, 4 * 32):
        nr = 4 * min(ring, 32, 128 - ring)
        seg = m[p:p + nr]
        if seg.max() > 0:
            assert seg.max() - seg.min() <= 1e-6 * seg.max()
        p += nr


def test_openmp_equals_serial(orc):
    cfg = _cfg(orc, nside=32)
    recs = _particles(orc, 3000, 8)
    m1, s1 = orc.genmap(recs, cfg, nthreads=1)
    m8, s8 = orc.genmap(recs, cfg, nthreads=4)
    assert {k: v for k, v in s1.items() if k != "total_flux"} == {k: v for k, v in s8.items() if k != "total_flux"}
    assert np.allclose(m1, m8, rtol=1e-12, atol=0)
