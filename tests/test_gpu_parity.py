"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Tolerances (per-pixel relative error over pixels above 1e-6 of the map
maximum; total-flux relative error):
  fp64          1e-9 / 1e-12   (reference arithmetic on the device)
  compensated   1e-5 / 1e-6    (FP32 weights, two-float accumulation)
  fp32          2e-5 / 1e-5    (FP32 weights, plain FP32 atomics: it loses
                                precision with the number of contributions
                                per pixel, which is what the compensated
                                form wins back; its error is reported)
Integer work (which particles are kept / sub-pixel / fallback / extended and
the exact footprint pixel counts) must match the oracle exactly.
"""
import numpy as np
import pytest

from conftest import assert_map, rel_err_report

pytestmark = pytest.mark.gpu

TOL = {"fp64": (1e-9, 1e-12), "compensated": (1e-5, 1e-6), "fp32": (2e-5, 1e-5)}


def _run_gpu(pw, recs, nside, precision, obs=(8.0, 0.0, 0.0), chunks=1, **ini):
    p = pw.params_for_observer(nside, precision, obs_pos=obs, **ini)
    n = len(recs)
    cap = max(1, (n + chunks - 1) // chunks)
    with pw.Projector(p, cap) as pr:
        for i in range(0, n, cap):
            pr.render(recs[i:i + cap])
        m = pr.get_map()
        st = pr.stats()
    return m, st


def _run_oracle(orc, pw, recs, nside, obs=(8.0, 0.0, 0.0), nthreads=8, **ini):
    import os
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = pw.write_ini(os.path.join(d, "o.ini"), nside=nside, obs_pos=obs, **ini)
        cfg = orc.config_from_params(orc.params_from_ini(path))
    return orc.genmap(recs, cfg, nthreads=nthreads)


def _check_counts(st, ost):
    assert st["kept"] == ost["kept"]
    assert st["subpixel"] == ost["subpixel"]
    assert st["fallback"] == ost["fallback"]
    assert st["extended"] == ost["extended"]
    assert st["contributions"] == ost["contributions"]


@pytest.fixture(scope="module")
def nfw_1e5(pw):
    return pw.synth.to_records(pw.synth.make("nfw_1e5"))


@pytest.fixture(scope="module")
def oracle_nfw_1e5(orc, pw, nfw_1e5):
    return _run_oracle(orc, pw, nfw_1e5, 64)


@pytest.mark.parametrize("precision", ["fp64", "compensated", "fp32"])
def test_nfw_1e5_nside64(pw, orc, gpu, nfw_1e5, oracle_nfw_1e5, precision):
    print(f"case nfw_1e5 nside=64 precision={precision}", flush=True)
    ref, ost = oracle_nfw_1e5
    m, st = _run_gpu(pw, nfw_1e5, 64, precision)
    _check_counts(st, ost)
    rep = rel_err_report(m, ref)
    print(rep)
    tol, ttol = TOL[precision]
    assert_map(rep, tol, ttol)


@pytest.mark.parametrize("nside", [256, 512])
@pytest.mark.parametrize("precision", ["compensated", "fp32"])
def test_subhalo_sample_high_nside(pw, orc, gpu, precision, nside):
    """A 2e4-particle sample of the 1e7 NFW + 100 subhalo set (seed 1)."""
    print(f"case sub100 sample nside={nside} precision={precision}", flush=True)
    parts = pw.synth.make("nfw_sub100_1e7", n=20000)
    recs = pw.synth.to_records(parts)
    ref, ost = _run_oracle(orc, pw, recs, nside)
    m, st = _run_gpu(pw, recs, nside, precision)
    _check_counts(st, ost)
    rep = rel_err_report(m, ref)
    print(rep)
    tol, ttol = TOL[precision]
    assert_map(rep, tol, ttol)


def test_chunked_accumulation_matches_single_call(pw, gpu, nfw_1e5):
    """Renders accumulate across calls like calculateMapByGPU chunks."""
    print("case chunked accumulation", flush=True)
    m1, st1 = _run_gpu(pw, nfw_1e5, 64, "fp64", chunks=1)
    m4, st4 = _run_gpu(pw, nfw_1e5, 64, "fp64", chunks=7)
    assert st1 == st4
    rep = rel_err_report(m4, m1)
    assert_map(rep, 1e-12)


@pytest.mark.parametrize("k", [0, 1, 16, 64])
@pytest.mark.parametrize("precision", ["compensated", "fp32"])
def test_grouping_sizes(pw, orc, gpu, precision, k):
    """Grouped rendering (K neighbours per wavefront sharing an LDS window;
    K = 0 is the per-particle kernel) against the oracle at nside 512."""
    print(f"case grouping k={k} precision={precision}", flush=True)
    recs = pw.synth.to_records(pw.synth.make("nfw_sub100_1e7", n=20000))
    ref, ost = _run_oracle(orc, pw, recs, 512)
    p = pw.params_for_observer(512, precision)
    with pw.Projector(p, len(recs)) as pr:
        pr.set_grouping(k)
        pr.render(recs)
        m = pr.get_map()
        st = pr.stats()
    _check_counts(st, ost)
    rep = rel_err_report(m, ref)
    print(rep)
    tol, ttol = TOL[precision]
    assert_map(rep, tol, ttol)


def _sky_spread(pw, n, seed, rad=0.06):
    """n particles at 5 kpc from the observer (8, 0, 0) in seeded random
    directions, footprint radius 2r = rad."""
    import numpy as np
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    d = 5.0
    a = np.zeros(n, dtype=pw.PARTICLE_DTYPE)
    a["mass"] = 4100.0
    a["dens"] = rng.uniform(1e5, 1e7, n)
    a["hsmooth"] = 0.5 * rad * d
    pos = np.array([8.0, 0.0, 0.0]) + d * u
    a["posx"], a["posy"], a["posz"] = pos[:, 0], pos[:, 1], pos[:, 2]
    return a


_HEADLINE = {}  # nside -> (records, oracle map, oracle statistics)


@pytest.mark.parametrize("nside", [2048, 4096])
@pytest.mark.parametrize("precision", ["fp64", "compensated"])
def test_headline_model_sample_high_nside(pw, orc, gpu, precision, nside):
    """3000 particles of the 1.1e9-particle model (one of them near the
    observer, covering much of the sky) at the benchmark's nside 2048 and
    the precision-at-scale nside 4096, against the oracle."""
    print(f"case headline sample nside={nside} precision={precision}", flush=True)
    if nside not in _HEADLINE:
        hm = pw.synth.model("nfw_sub2000_near_1p1e9", "cpu")
        recs = pw.synth.to_records(hm.sample(3000))
        _HEADLINE[nside] = (recs,) + tuple(_run_oracle(orc, pw, recs, nside))
    recs, ref, ost = _HEADLINE[nside]
    m, st = _run_gpu(pw, recs, nside, precision)
    _check_counts(st, ost)
    rep = rel_err_report(m, ref)
    print(rep, st["large"])
    tol, ttol = TOL[precision]
    assert_map(rep, tol, ttol)


@pytest.mark.parametrize("precision", ["fp64", "compensated", "fp32"])
def test_sky_covering_size_classes(pw, orc, gpu, precision):
    """900 particles at 5 kpc with footprint radii 2r log-uniform in
    0.05 .. 3.5 rad (from ~100 pixels to the whole sky) at nside 128: the
    sky-covering class (about half of them, more than one 256-particle
    chunk) spans all of its kernel's size classes, which are rendered
    largest first; every particle must be rendered exactly once."""
    import numpy as np
    print(f"case sky-covering size classes precision={precision}", flush=True)
    rng = np.random.default_rng(21)
    n = 900
    rad = np.exp(rng.uniform(np.log(0.05), np.log(3.5), n))
    recs = _sky_spread(pw, n, seed=22)
    recs["hsmooth"] = (0.5 * rad * 5.0).astype(np.float32)
    ref, ost = _run_oracle(orc, pw, recs, 128)
    m, st = _run_gpu(pw, recs, 128, precision)
    _check_counts(st, ost)
    assert st["large"] > 256
    rep = rel_err_report(m, ref)
    print(rep, st["large"])
    tol, ttol = TOL[precision]
    assert_map(rep, tol, ttol)


@pytest.mark.parametrize("precision", ["compensated", "fp32"])
def test_grouping_wide_windows(pw, orc, gpu, precision):
    """One group of 60 particles spread over the whole sky (sort-cell nside
    1, 64 per group): its window spans whole rings (the ring wrap goes to
    the spare slots) and, at nside 1024, rings wider than the LDS buffer
    (deposited with global atomics)."""
    print(f"case wide windows precision={precision}", flush=True)
    recs = _sky_spread(pw, 60, seed=11)
    ref, ost = _run_oracle(orc, pw, recs, 1024)
    p = pw.params_for_observer(1024, precision)
    with pw.Projector(p, len(recs)) as pr:
        pr.set_grouping(64, nside_c=1)
        pr.render(recs)
        m = pr.get_map()
        st = pr.stats()
    _check_counts(st, ost)
    assert st["large"] < 10
    rep = rel_err_report(m, ref)
    print(rep)
    tol, ttol = TOL[precision]
    assert_map(rep, tol, ttol)


@pytest.mark.parametrize("precision", ["fp64", "compensated"])
def test_flux_plugins(pw, orc, gpu, precision):
    """Maps of every flux plug-in (the reference's src/flux/flux_<kind>.cpp)
    against the oracle with the same plug-in, nside 64."""
    import os
    import tempfile
    recs = pw.synth.to_records(pw.synth.make("nfw_1e5", n=20000)).copy()
    with tempfile.TemporaryDirectory() as d:
        ini = pw.write_ini(os.path.join(d, "o.ini"), nside=64, centerpos=(0.5, 0.25, -0.75))
        cfg = orc.config_from_params(orc.params_from_ini(ini))
        p, _, _ = pw.params_from_ini(ini, precision)
    recs["sigmav"] = np.geomspace(0.05, 300.0, len(recs)) * 1e5 / cfg.vel_to_cgs
    tol, ttol = TOL[precision]
    for k, kind in enumerate(pw.capi.FLUX_KINDS):
        if kind == "ann_core_nosom":
            continue  # needs a core table: test_halo_core_plugin
        print(f"case flux plug-in {kind} precision={precision}", flush=True)
        cfg.flux_kind = k
        ref, ost = orc.genmap(recs, cfg, nthreads=8)
        with pw.Projector(p.with_flux(kind), len(recs)) as pr:
            pr.render(recs)
            m = pr.get_map()
            st = pr.stats()
        _check_counts(st, ost)
        rep = rel_err_report(m, ref)
        print(kind, rep)
        assert_map(rep, tol, ttol)


@pytest.mark.parametrize("precision", ["fp64", "compensated"])
def test_cuda_cutoff(pw, orc, gpu, precision):
    """Footprints of radius 2r + theta_pix/2 (the reference CUDA render's
    cutoff, footprint_pad = pi / (4 nside)) against the oracle with the same
    pad, nside 128."""
    import os
    import tempfile
    print(f"case cuda cutoff precision={precision}", flush=True)
    recs = pw.synth.to_records(pw.synth.make("nfw_1e5", n=20000))
    with tempfile.TemporaryDirectory() as d:
        ini = pw.write_ini(os.path.join(d, "o.ini"), nside=128)
        cfg = orc.config_from_params(orc.params_from_ini(ini))
        p, _, _ = pw.params_from_ini(ini, precision)
    p = p.with_cuda_cutoff()
    cfg.footprint_pad = p.footprint_pad
    ref, ost = orc.genmap(recs, cfg, nthreads=8)
    with pw.Projector(p, len(recs)) as pr:
        pr.render(recs)
        m = pr.get_map()
        st = pr.stats()
    _check_counts(st, ost)
    rep = rel_err_report(m, ref)
    print(rep)
    tol, ttol = TOL[precision]
    assert_map(rep, tol, ttol)


@pytest.mark.parametrize("precision", ["fp64", "compensated"])
def test_halo_core_plugin(pw, orc, gpu, precision):
    """The halo-core flux plug-in (device grid over the cores) against the
    oracle's loop over every core: (a) 3000 synthetic cores spread through
    the sampled halo, radii of 0.2-2 kpc; (b) the reference's VL2 halo file
    with particles placed in and around its cores (VL2 ini)."""
    import os
    import tempfile
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    rng = np.random.default_rng(21)
    recs = pw.synth.to_records(pw.synth.make("nfw_1e5", n=20000)).copy()
    pos = np.stack([recs["posx"], recs["posy"], recs["posz"]], 1).astype(np.float64)
    pick = rng.choice(len(recs), 3000, replace=False)
    halos = np.zeros((3000, 4), np.float32)
    halos[:, :3] = pos[pick] + rng.normal(size=(3000, 3)) * 0.5
    halos[:, 3] = rng.uniform(0.2, 2.0, 3000) * 0.73 * 40000.0
    with tempfile.TemporaryDirectory() as d:
        ini = pw.write_ini(os.path.join(d, "o.ini"), nside=64)
        cases = [(ini, recs, halos)]
        vl2 = os.path.join(gold, "VL2_info.ini")
        vh = pw.read_halo_file(os.path.join(gold, "VL2_Halos.data"))
        # cores at least 0.003 box units (~90 kpc) from the VL2 observer
        obs = np.array([0.36116062, 0.21047031, 0.0069339937])
        far = np.nonzero(np.linalg.norm(vh[:, :3] - obs, axis=1) > 0.003)[0]
        sel = rng.choice(far, 4000, replace=False)
        rc = vh[sel, 3].astype(np.float64) / (0.73 * 40000.0)
        off = rng.normal(size=(4000, 3))
        off *= (rc * rng.uniform(0.0, 1.5, 4000) / np.linalg.norm(off, axis=1))[:, None]
        vrec = np.zeros(4000, dtype=pw.PARTICLE_DTYPE)
        vrec["posx"], vrec["posy"], vrec["posz"] = (vh[sel, :3] + off).T
        vrec["mass"] = 4e-13  # below the VL2 high-resolution cut
        vrec["dens"] = rng.uniform(1.0, 1e4, 4000)
        vrec["hsmooth"] = rng.uniform(2e-6, 4e-5, 4000)
        cases.append((vl2, vrec, vh))
        for path, rr, hh in cases:
            print(f"case halo cores {os.path.basename(path)} precision={precision}", flush=True)
            cfg = orc.config_from_params(orc.params_from_ini(path))
            cfg.flux_kind = orc.FLUX_KINDS.index("ann_core_nosom")
            cfg.set_halos(hh)
            p, _, _ = pw.params_from_ini(path, precision)
            ref, ost = orc.genmap(rr, cfg, nthreads=8)
            with pw.Projector(p.with_flux("ann_core_nosom"), len(rr)) as pr:
                pr.set_halo_cores(hh)
                pr.render(rr)
                m = pr.get_map()
                st = pr.stats()
            _check_counts(st, ost)
            rep = rel_err_report(m, ref)
            print(rep)
            tol, ttol = TOL[precision]
            if precision == "fp64" and path == vl2:
                # footprints of a few pixels with r well below the pixel
                # spacing: the reference's acos(dot) weight amplifies a
                # one-ulp difference between device and host sin/cos in the
                # pixel vectors to ~1e-9 of a weight
                tol = 1e-8
            assert_map(rep, tol, ttol)
