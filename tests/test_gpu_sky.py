"""The sky-covering class rendered as a per-ring gather (one wavefront per
HEALPix ring, LDS accumulation of up to 64 particles, no atomics) against
the oracle and against the per-particle kernel with per-pixel atomics.

* 3e4 particles of the 1.1e9 model (NFW host + 2000 subhalos + near-observer
  ball), with 3000 of its near-observer particles (footprints up to the
  whole sky), nside 256: gather, compensated, against the oracle -- integer
  statistics exact, 1e-5 per pixel, 1e-6 on the total (a few hundred
  particles take the sky-covering class here, with footprints of 6 % of
  the sky and more).
* 2e6 particles of the same model with 2000 of its near-observer particles,
  nside 2048 (rings of 8192 pixels: four LDS segments): gather against the
  per-particle kernel, both compensated, and against the FP64 form, 1e-5
  per pixel; FP32 gather against FP32 per-particle; both timed (the
  sky-covering stage from the render's events).
"""
import os
import tempfile

import pytest

from conftest import assert_map, progress, record, rel_err_report

pytestmark = pytest.mark.gpu


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _render(pw, torch, parts, p, gather=True, timing=False):
    with pw.Projector(p, parts.shape[0]) as pr:
        if not gather:
            pr.set_sky_gather(0)
        pr.render(parts)  # warm-up (module load)
        pr.reset()
        if timing:
            pr.timing(True)
        pr.render(parts)
        out = torch.empty(pr.npix, dtype=torch.float64, device=parts.device)
        pr.get_map_device(out)
        torch.cuda.synchronize()
        t = pr.timing(False)[0] if timing else None
        st = pr.stats()
    return out.cpu().numpy(), st, t


def test_sky_gather_vs_oracle(pw, orc, gpu):
    import torch
    print("case sky gather vs oracle, nside 256", flush=True)
    hm = pw.synth.model("nfw_sub2000_near_1p1e9", "cpu")
    sub = hm.sample(30000, near_min=3000)
    recs = pw.synth.to_records(sub)
    nside = 256
    p = pw.params_for_observer(nside, "compensated")
    m, st, _ = _render(pw, torch, sub.to(gpu), p)
    assert st["large"] >= 100
    with tempfile.TemporaryDirectory() as d:
        path = pw.write_ini(os.path.join(d, "o.ini"), nside=nside)
        cfg = orc.config_from_params(orc.params_from_ini(path))
    ref, ost = orc.genmap(recs, cfg, nthreads=_threads())
    for k in ("kept", "subpixel", "fallback", "extended", "contributions"):
        assert st[k] == ost[k], (k, st[k], ost[k])
    rep = rel_err_report(m, ref)
    record("sky gather 3e4 (3000 near) nside256 vs oracle", max_rel=rep["max_rel"], total_rel=rep["total_rel"],
           large=st["large"], large_contributions=st["large_contributions"], contributions=st["contributions"])
    assert_map(rep, 1e-5, 1e-6)


@pytest.mark.timeout(600)
def test_sky_gather_vs_scatter_nside2048(pw, gpu):
    import torch
    print("case sky gather vs per-particle kernel, nside 2048", flush=True)
    hm = pw.synth.model("nfw_sub2000_near_1p1e9", "cuda")
    parts = hm.sample(2_000_000, near_min=2_000)
    nside = 2048
    pc = pw.params_for_observer(nside, "compensated")
    g, st_g, t_g = _render(pw, torch, parts, pc, gather=True, timing=True)
    progress("sky gather: compensated gather rendered")
    s, st_s, t_s = _render(pw, torch, parts, pc, gather=False, timing=True)
    progress("sky gather: compensated per-particle rendered")
    ref, st_r, _ = _render(pw, torch, parts, pw.params_for_observer(nside, "fp64"))
    progress("sky gather: fp64 rendered")
    for k in ("kept", "subpixel", "fallback", "extended", "contributions", "large", "large_contributions"):
        assert st_g[k] == st_s[k] == st_r[k], k
    e_gr = rel_err_report(g, ref)
    e_sr = rel_err_report(s, ref)
    e_gs = rel_err_report(g, s)
    pf = pw.params_for_observer(nside, "fp32")
    gf, _, tf_g = _render(pw, torch, parts, pf, gather=True, timing=True)
    sf, _, tf_s = _render(pw, torch, parts, pf, gather=False, timing=True)
    e_ff = rel_err_report(gf, sf)
    record("sky gather 2e6 (2000 near) nside2048", large=st_g["large"], large_contributions=st_g["large_contributions"],
           gather_vs_fp64=e_gr["max_rel"], scatter_vs_fp64=e_sr["max_rel"], gather_vs_scatter=e_gs["max_rel"],
           fp32_gather_vs_scatter=e_ff["max_rel"], comp_gather_ms=t_g["large"], comp_scatter_ms=t_s["large"],
           fp32_gather_ms=tf_g["large"], fp32_scatter_ms=tf_s["large"])
    print(e_gr, e_sr, e_gs, e_ff, t_g, t_s, tf_g, tf_s, flush=True)
    assert_map(e_gr, 1e-5, 1e-6)
    assert_map(e_gs, 1e-5, 1e-6)
    assert_map(e_ff, 1e-4, 1e-5)
    torch.cuda.synchronize()
