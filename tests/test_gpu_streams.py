"""Stream ordering of the C ABI and domain errors of batches.

A projector bound to a side stream (torch.cuda.stream) must order its map
clear, render and read after the work queued on the stream it used before,
and give the same map as on the default stream; a batch renders on its own
stream but forks from and joins back into torch's current stream.  A kept
particle at an observer's position has no direction: the single-map path
reports DMANN_ERR_DOMAIN (the reference exits, src/anglefuncs.cpp:60-63),
and so must a batch, under both schedules.
"""
import math

import numpy as np
import pytest

from conftest import assert_map, rel_err_report

pytestmark = pytest.mark.gpu

DOMAIN = 4


def test_project_map_on_side_stream(pw, gpu):
    import torch
    print("case project_map on a side stream", flush=True)
    parts = pw.synth.make("nfw_sub100_1e7", n=20000).to(gpu)
    p = pw.params_for_observer(256, "compensated")
    with pw.Projector(p, parts.shape[0]) as pr:
        ref = pw.project_map(parts, pr).clone()
        torch.cuda.synchronize()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            # the particles are produced on the side stream right before
            # the render reads them
            q = parts * 1.0
            m1 = pw.project_map(q, pr)
            m2 = pw.project_map(q, pr)
        torch.cuda.current_stream().wait_stream(side)
        got1 = m1.cpu().numpy()
        got2 = m2.cpu().numpy()
        # back on the default stream
        m3 = pw.project_map(parts, pr).cpu().numpy()
    ref = ref.cpu().numpy()
    for got in (got1, got2, m3):
        assert_map(rel_err_report(got, ref), 2e-6, 1e-6)


def _observer_params(pw, k, nmaps, nside, precision):
    a = 2 * math.pi * k / nmaps
    return pw.params_for_observer(nside, precision, obs_pos=(8 * math.cos(a), 8 * math.sin(a), 0.0))


@pytest.mark.parametrize("chunk", [0, 64])
def test_batch_reports_undefined_direction(pw, gpu, chunk):
    import torch
    print(f"case batch domain error chunk={chunk}", flush=True)
    plist = [_observer_params(pw, k, 4, 64, "compensated") for k in range(4)]
    parts = pw.synth.make("nfw_1e5", n=2000)
    # one particle exactly at observer 0's position (8, 0, 0), which FP32
    # represents exactly (the other observers' positions are not)
    o = plist[0].opos
    assert list(o) == [8.0, 0.0, 0.0]
    parts[7, 3], parts[7, 4], parts[7, 5] = o[0], o[1], o[2]
    parts = parts.to(gpu)
    with pw.Batch(plist) as b:
        if chunk:
            b.set_tuning(chunk, 8)
        b.render(parts, use_graph=False)
        with pytest.raises(pw.DmannError) as ei:
            b.synchronize()
        assert ei.value.code == DOMAIN and "map 0" in str(ei.value)
    # the single-map path says the same for that observer
    with pw.Projector(plist[0], parts.shape[0]) as pr:
        pr.render(parts)
        with pytest.raises(pw.DmannError) as ei:
            pr.synchronize()
        assert ei.value.code == DOMAIN
    torch.cuda.synchronize()


def test_batch_follows_caller_stream(pw, gpu):
    """Particles written on a side stream just before the render; the maps
    read on that stream just after it."""
    import torch
    print("case batch caller stream", flush=True)
    plist = [_observer_params(pw, k, 4, 128, "compensated") for k in range(4)]
    base = pw.synth.make("nfw_sub100_1e7", n=20000).to(gpu)
    with pw.Batch(plist) as b:
        b.render(base, use_graph=True)
        b.synchronize()
        ref = [b.get_map(i) for i in range(4)]
        buf = torch.zeros_like(base)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            buf.copy_(base)
            b.render(buf, use_graph=True)
            ptrs = []
            for i in range(4):
                ptrs.append(b.map_device(i).clone())
        torch.cuda.current_stream().wait_stream(side)
        got = [t.cpu().numpy() for t in ptrs]
    for i in range(4):
        assert_map(rel_err_report(got[i], ref[i]), 2e-6, 1e-6)
