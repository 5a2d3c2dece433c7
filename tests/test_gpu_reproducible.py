"""The reproducible form (DMANN_PREC_FIXED): the same particles rendered in
the same calls give the same map bit for bit -- across contexts, across
launch shapes (workgroups per CU of the footprint kernels, the gather's
occupancy), for device and staged host input, and when a later render
shifts the fixed-point scale down -- while staying within the compensated
form's bar of the oracle.

Workload: 3e5 particles of the 1.1e9 model (NFW host + 2000 subhalos +
near-observer ball) with 300 of its near-observer particles, nside 512:
sub-pixel, medium (grouped) and sky-covering (gathered) particles together.
"""
import os
import tempfile

import numpy as np
import pytest

from conftest import assert_map, progress, record, rel_err_report

pytestmark = pytest.mark.gpu

NSIDE = 512


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _render(pw, torch, parts, precision, chunks=(), tune=None):
    p = pw.params_for_observer(NSIDE, precision)
    with pw.Projector(p, parts.shape[0]) as pr:
        if tune:
            pr.set_tuning(*tune["tuning"])
            pr.set_sky_gather(tune["sky"])
        bounds = [0, *chunks, parts.shape[0]]
        for i0, i1 in zip(bounds[:-1], bounds[1:]):
            pr.render(parts[i0:i1])
        m = pr.get_map()
        st = pr.stats()
    return m, st


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


@pytest.fixture(scope="module")
def parts(pw, gpu):
    hm = pw.synth.model("nfw_sub2000_near_1p1e9", "cpu")
    return hm.sample(300_000, near_min=300)


def test_fixed_is_bit_reproducible(pw, orc, gpu, parts):
    import torch
    print("case reproducible form, 3e5 particles, nside 512", flush=True)
    dev = parts.to(gpu)
    a, sa = _render(pw, torch, dev, "fixed")
    b, sb = _render(pw, torch, dev, "fixed")
    c, sc = _render(pw, torch, dev, "fixed", tune={"tuning": (4, 2, 4194304), "sky": 5})
    h, sh = _render(pw, torch, pw.synth.to_records(parts), "fixed")
    progress("reproducible form: four renders")
    assert sa["large"] > 0 and sa["subpixel"] > 0 and sa["extended"] > sa["large"]
    for s in (sb, sc, sh):
        assert s == sa
    assert _same(a, b) and _same(a, c) and _same(a, h)
    # the compensated form on the same particles, and the oracle
    comp, _ = _render(pw, torch, dev, "compensated")
    with tempfile.TemporaryDirectory() as d:
        path = pw.write_ini(os.path.join(d, "o.ini"), nside=NSIDE)
        cfg = orc.config_from_params(orc.params_from_ini(path))
    ref, ost = orc.genmap(pw.synth.to_records(parts), cfg, nthreads=_threads())
    for k in ("kept", "subpixel", "fallback", "extended", "contributions"):
        assert sa[k] == ost[k], k
    rep = rel_err_report(a, ref)
    rep_c = rel_err_report(a, comp)
    record("reproducible form 3e5 nside512", max_rel_vs_oracle=rep["max_rel"], total_rel=rep["total_rel"],
           max_rel_vs_compensated=rep_c["max_rel"], large=sa["large"], subpixel=sa["subpixel"])
    assert_map(rep, 1e-5, 1e-6)
    assert_map(rep_c, 2e-6, 1e-6)


def test_fixed_chunks_and_scale_shift(pw, gpu, parts):
    """Renders in three calls, the first of 2000 particles: the later calls
    carry far more flux than the scale was set for, so the map is shifted
    down; the same calls again give the same bits, and the map stays within
    2e-6 of the one-call render (the calls group their particles
    differently)."""
    import torch
    print("case reproducible form, chunks and scale shift", flush=True)
    dev = parts.to(gpu)
    whole, _ = _render(pw, torch, dev, "fixed")
    x, sx = _render(pw, torch, dev, "fixed", chunks=(2000, 150_000))
    y, sy = _render(pw, torch, dev, "fixed", chunks=(2000, 150_000))
    assert sx == sy and _same(x, y)
    assert_map(rel_err_report(x, whole), 2e-6, 1e-6)
    torch.cuda.synchronize()
