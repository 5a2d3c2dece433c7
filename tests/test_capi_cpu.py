"""C ABI checks that need no GPU: the library loads, exports every symbol
include/dmann.h declares, rejects bad arguments before touching a device,
and its host-side footprint enumerator (the kernels' own per-particle and
per-ring code) reproduces the oracle's footprints exactly, with FP32 weights
within 5e-6 of the oracle's FP64 weights after normalisation.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _declared():
    src = open(os.path.join(ROOT, "include", "dmann.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dmann_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol(pw):
    L = pw.load_library()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(pw.capi.EXPORTS)


def test_particle_layout(pw):
    assert pw.PARTICLE_DTYPE.itemsize == 48
    assert pw.PARTICLE_DTYPE.names[:3] == ("mass", "dens", "hsmooth")


def test_create_rejects_bad_arguments_without_a_device(pw):
    L = pw.load_library()
    p = pw.params_for_observer(64, "fp32")
    h = C.c_void_p()
    for nside in (0, -3, 8193):
        q = p.copy()
        q.nside = nside
        assert L.dmann_create(C.byref(h), C.byref(q), 10, 0) == 1
        assert b"nside" in L.dmann_last_error(None)
    q = p.copy()
    q.precision = 7
    assert L.dmann_create(C.byref(h), C.byref(q), 10, 0) == 1
    assert L.dmann_create(C.byref(h), C.byref(p), -1, 0) == 1
    assert L.dmann_create(None, C.byref(p), 1, 0) == 1
    assert L.dmann_render(None, None, 1) == 1
    assert L.dmann_get_map(None, None) == 1


def test_params_from_ini_missing_file(pw):
    with pytest.raises(pw.DmannError):
        pw.params_from_ini("/nonexistent/file.ini")


def test_params_match_oracle(pw, orc):
    for ini in ("synth_obs8.ini", "synth_offaxis.ini", "VL2_info.ini"):
        path = os.path.join(GOLD, ini)
        p, _, _ = pw.params_from_ini(path, "fp32")
        c = orc.config_from_params(orc.params_from_ini(path))
        assert list(p.rotmatrix) == list(c.rotmatrix)
        assert list(p.opos) == list(c.opos)
        assert p.flux_factor == c.flux_factor and p.mass_cut == c.mass_cut and p.nside == c.nside


@pytest.mark.parametrize("workload,nside,m", [("nfw_1e5", 64, 400), ("nfw_sub100_1e7", 512, 300),
                                              ("nfw_sub2000_near_1p1e9", 4096, 60)])
def test_host_footprint_matches_oracle(pw, orc, tmp_path, workload, nside, m):
    ini = pw.write_ini(str(tmp_path / "o.ini"), nside=nside)
    p, _, _ = pw.params_from_ini(ini, "fp32")
    cfg = orc.config_from_params(orc.params_from_ini(ini))
    recs = pw.synth.to_records(pw.synth.make(workload, n=m))
    worst = 0.0
    for q in recs:
        kind, pix, w, fl = pw.capi.host_footprint(p, q)
        opix, ow, ofl = orc.particle_footprint(q, cfg)
        assert fl == ofl
        assert np.array_equal(np.sort(pix), opix)
        if kind == 3:
            o = np.argsort(pix)
            wn = w[o].astype(np.float64)
            wn /= wn.sum()
            worst = max(worst, float(np.max(np.abs(wn - ow) / ow)))
    assert worst <= 5e-6


@pytest.mark.parametrize("nside", [64, 512, 2048, 4096])
def test_ring_lite_matches_ring_setup(pw, tmp_path, nside):
    """The grouped kernel's ring ranges -- an FP32 half-width with an FP64
    recomputation when a column boundary lies within its error bound --
    equal the query_disc construction's ring by ring: column counts, first
    columns and offsets exactly, FP32 terms to a few units of rounding (a
    3e-5 relative error planted in the half-width fails this at nside
    512)."""
    ini = pw.write_ini(str(tmp_path / "o.ini"), nside=nside)
    p, _, _ = pw.params_from_ini(ini, "fp32")
    recs = pw.synth.to_records(pw.synth.make("nfw_sub2000_near_1p1e9", n=1500))
    checked = 0
    for q in recs:
        m = pw.capi.host_ring_check(p, q)
        if m == -2:
            continue
        assert m == 0
        checked += 1
    assert checked > 1000


def _placed(pw, p, cases):
    """Particles at (colatitude, longitude, distance, smoothing length) in
    the map frame, placed by inverting the observer rotation."""
    R = np.array(list(p.rotmatrix)).reshape(3, 3)
    out = []
    for th, ph, d, h in cases:
        v = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        t = np.array([-v[0], -v[1], v[2]]) * d  # undo the phi - pi shift
        pos = R @ t + np.array([8.0, 0, 0])
        rec = np.zeros(1, dtype=pw.PARTICLE_DTYPE)
        rec["posx"], rec["posy"], rec["posz"] = pos.astype(np.float32)
        rec["mass"], rec["dens"], rec["hsmooth"] = 4100.0, 1.0, h
        out.append(rec[0])
    return out


@pytest.mark.parametrize("nside", [64, 512, 1024])
def test_sky_gather_indexing(pw, tmp_path, nside):
    """The sky-covering gather (one wavefront per ring, 2048-column LDS
    segments) visits every pixel of the per-particle ring walk's footprint
    once, with the same FP32 weight, and no other pixel: footprints over a
    pole, across phi = 0, whole-sky ones, and at nside 1024 rings of 4096
    columns (two segments) and polar rings of 2052..4092 (a partial second
    segment)."""
    ini = pw.write_ini(str(tmp_path / "o.ini"), nside=nside)
    p, _, _ = pw.params_from_ini(ini, "fp32")
    rng = np.random.default_rng(11)
    cases = [(0.05, 1.0, 1.0, 0.05), (np.pi - 0.03, 4.0, 1.0, 0.04), (1.2, 1e-4, 1.0, 0.3),
             (2.0, 2 * np.pi - 1e-4, 1.0, 0.35), (1.0, 3.0, 0.2, 0.5), (1.57, 0.5, 0.5, 0.6),
             (0.4, 2.0, 1.0, 0.8)]
    for _ in range(12):
        cases.append((float(np.arccos(rng.uniform(-1, 1))), float(rng.uniform(0, 2 * np.pi)), 1.0,
                      float(rng.uniform(0.26, 0.9))))
    checked = 0
    for q in _placed(pw, p, cases):
        m = pw.capi.host_sky_check(p, q)
        if m == -2:
            continue
        assert m == 0
        checked += 1
    assert checked >= 15


def _ring_centres(nside):
    """Directions (theta, phi) of the centres of all RING pixels at nside
    (the HEALPix RING layout: polar caps of 4 i pixels on ring i, the
    equatorial belt of 4 nside pixels per ring, alternately shifted by half
    a pixel): test inputs."""
    n = nside
    p = np.arange(12 * n * n, dtype=np.int64)
    z = np.empty(p.size)
    phi = np.empty(p.size)
    ncap = 2 * n * (n - 1)
    north = p < ncap
    i = np.floor((1 + np.sqrt(1 + 2 * p[north])) / 2).astype(np.int64)
    j = p[north] - 2 * i * (i - 1) + 1
    z[north] = 1 - i * i / (3.0 * n * n)
    phi[north] = (j - 0.5) * np.pi / (2 * i)
    belt = (p >= ncap) & (p < 12 * n * n - ncap)
    q = p[belt] - ncap
    i = q // (4 * n) + n
    j = q % (4 * n) + 1
    shift = (i - n + 1) % 2
    z[belt] = 4.0 / 3.0 - 2.0 * i / (3.0 * n)
    phi[belt] = (j - 1 + shift / 2.0) * np.pi / (2 * n)
    south = p >= 12 * n * n - ncap
    q = 12 * n * n - p[south]
    i = np.floor((1 + np.sqrt(2 * q - 1)) / 2).astype(np.int64)
    j = 4 * i + 1 - (q - 2 * i * (i - 1))
    z[south] = -1 + i * i / (3.0 * n * n)
    phi[south] = (j - 0.5) * np.pi / (2 * i)
    return np.arccos(z), phi


def test_nest_cells(pw):
    """The NESTED index the grouping's sort keys use (host build of the
    device code): a permutation of 0 .. 12 nside^2 - 1 over the RING pixel
    centres at nside 1..128; at nside 1 the base faces (RING and NESTED
    agree); at nside 2 the first RING pixel is NESTED 3 (the north corner
    of face 0); and each pixel's index at nside / 2 is its index >> 2 (the
    parent cell), which makes consecutive keys compact 2-D patches."""
    for nside in (1, 2, 4, 8, 64, 128):
        th, ph = _ring_centres(nside)
        cells = pw.capi.host_nest_cells(nside, th, ph)
        assert np.array_equal(np.sort(cells), np.arange(12 * nside * nside, dtype=np.uint32)), nside
        if nside == 1:
            assert np.array_equal(cells, np.arange(12, dtype=np.uint32))
        if nside == 2:
            assert cells[0] == 3
        if nside > 1:
            parent = pw.capi.host_nest_cells(nside // 2, th, ph)
            assert np.array_equal(cells >> 2, parent), nside
    with pytest.raises(ValueError):
        pw.capi.host_nest_cells(3, np.zeros(1), np.zeros(1))


@pytest.mark.parametrize("nside", [256, 2048])
def test_polar_series_form_matches_oracle(pw, orc, tmp_path, nside):
    """Footprints of radius <= 0.5 that are too close to a pole for the fast
    form (sin 2r >= 0.43 sin theta: longitude offsets up to pi) take the
    sin(x/2) Taylor series with the fast asin^2 series: their pixel sets
    equal the oracle's and their normalised weights match its FP64 weights
    to 5e-6, over the pole, next to it and across the phi = 0 seam."""
    ini = pw.write_ini(str(tmp_path / "o.ini"), nside=nside)
    p, _, _ = pw.params_from_ini(ini, "fp32")
    cfg = orc.config_from_params(orc.params_from_ini(ini))
    cases = [(0.05, 1.0, 1.0, 0.05), (0.02, 4.0, 1.0, 0.03), (0.15, 1e-4, 1.0, 0.1), (np.pi - 0.05, 2.0, 1.0, 0.08),
             (np.pi - 0.2, 2 * np.pi - 1e-4, 1.0, 0.12), (0.3, 3.0, 1.0, 0.2), (0.001, 0.5, 1.0, 0.24)]
    worst = 0.0
    for q in _placed(pw, p, cases):
        kind, pix, w, fl = pw.capi.host_footprint(p, q)
        opix, ow, ofl = orc.particle_footprint(q, cfg)
        assert np.array_equal(np.sort(pix), opix)
        assert kind == 3
        o = np.argsort(pix)
        wn = w[o].astype(np.float64)
        wn /= wn.sum()
        worst = max(worst, float(np.max(np.abs(wn - ow) / ow)))
    assert worst <= 5e-6


def test_host_footprint_edge_geometries(pw, orc, tmp_path):
    """Footprints over a pole, across the phi = 0 seam, covering the whole
    sky, and tiny ones, placed by inverting the observer rotation."""
    nside = 128
    ini = pw.write_ini(str(tmp_path / "o.ini"), nside=nside)
    p, _, _ = pw.params_from_ini(ini, "fp32")
    cfg = orc.config_from_params(orc.params_from_ini(ini))
    R = np.array(list(p.rotmatrix)).reshape(3, 3)
    cases = []
    for th, ph, d, h in [(0.0005, 1.0, 30, 1.0), (np.pi - 0.002, 4.0, 30, 2.0), (1.0, 1e-4, 25, 3.0),
                         (2.2, 2 * np.pi - 1e-4, 25, 3.0), (1.3, 3.0, 0.2, 0.5), (0.9, 2.0, 0.05, 2.0),
                         (0.6, 0.1, 40, 0.02), (1.57, 6.2, 40, 0.05)]:
        v = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        t = np.array([-v[0], -v[1], v[2]]) * d  # undo the phi - pi shift
        pos = R @ t + np.array([8.0, 0, 0])
        rec = np.zeros(1, dtype=pw.PARTICLE_DTYPE)
        rec["posx"], rec["posy"], rec["posz"] = pos.astype(np.float32)
        rec["mass"], rec["dens"], rec["hsmooth"] = 4100.0, 1.0, h
        cases.append(rec[0])
    for q in cases:
        kind, pix, w, fl = pw.capi.host_footprint(p, q)
        opix, ow, ofl = orc.particle_footprint(q, cfg)
        assert np.array_equal(np.sort(pix), opix)
        assert pix.min() >= 0 and pix.max() < 12 * nside * nside
        if kind == 3:
            o = np.argsort(pix)
            wn = w[o].astype(np.float64)
            wn /= wn.sum()
            assert np.max(np.abs(wn - ow) / ow) <= 5e-6


def test_product_flux_plugins_match_oracle(pw, orc, tmp_path):
    """The product's pre-pass flux (host build of the kernels' code) for every
    plug-in equals the oracle's, which the reference's own plug-ins pin."""
    ini = pw.write_ini(str(tmp_path / "o.ini"), nside=128, centerpos=(0.5, 0.25, -0.75))
    p, _, _ = pw.params_from_ini(ini, "fp64")
    cfg = orc.config_from_params(orc.params_from_ini(ini))
    assert p.decay_factor == cfg.decay_factor and p.vel_to_cgs == cfg.vel_to_cgs
    assert p.c_in_cgs == cfg.c_in_cgs and list(p.cpos) == list(cfg.cpos)
    recs = pw.synth.to_records(pw.synth.make("nfw_1e5", n=300)).copy()
    rng = np.random.default_rng(5)
    recs["sigmav"] = np.geomspace(0.05, 300.0, len(recs)) * 1e5 / cfg.vel_to_cgs * rng.uniform(0.9, 1.1, len(recs))
    for k, kind in enumerate(pw.capi.FLUX_KINDS):
        if kind == "ann_core_nosom":
            continue  # its core table lives in a device context (GPU tests)
        pk = p.with_flux(kind)
        cfg.flux_kind = k
        for q in recs:
            kind_p, _, _, fl = pw.capi.host_footprint(pk, q)
            _, _, ofl = orc.particle_footprint(q, cfg)
            if kind_p:
                assert fl == ofl, kind


def test_host_footprint_cuda_cutoff(pw, orc, tmp_path):
    """With footprint_pad = theta_pix/2 (the reference CUDA render's cutoff
    2r + theta_pix/2, src/chealpixrender.cu:188) the footprint is the set of
    pixel centres within that radius -- equal to the oracle's with the same
    pad, and a superset of the CPU generator's."""
    import math
    nside = 128
    ini = pw.write_ini(str(tmp_path / "o.ini"), nside=nside)
    p, _, _ = pw.params_from_ini(ini, "fp32")
    pc = p.with_cuda_cutoff()
    assert pc.footprint_pad == math.pi / (4 * nside)
    cfg = orc.config_from_params(orc.params_from_ini(ini))
    cfg.footprint_pad = pc.footprint_pad
    cfg0 = orc.config_from_params(orc.params_from_ini(ini))
    recs = pw.synth.to_records(pw.synth.make("nfw_1e5", n=400))
    grew = 0
    for q in recs:
        kind, pix, w, fl = pw.capi.host_footprint(pc, q)
        opix, ow, ofl = orc.particle_footprint(q, cfg)
        assert np.array_equal(np.sort(pix), opix)
        if kind == 3:
            o = np.argsort(pix)
            wn = w[o].astype(np.float64)
            wn /= wn.sum()
            assert np.max(np.abs(wn - ow) / ow) <= 5e-6
            opix0, _, _ = orc.particle_footprint(q, cfg0)
            assert set(opix0.tolist()) <= set(opix.tolist())
            grew += len(opix) > len(opix0)
    assert grew > 0
