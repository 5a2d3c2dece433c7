"""The allocation-free entry point (dmann_render_map_noalloc) against the
context path (dmann_create + dmann_render_device + dmann_get_map_device) on
the same particles, through the C ABI: two calls in a row, the second with
fewer particles and a smaller map, so a stale scratch (records, counters)
or an uncleared output would show.  Tolerances as the parity tests: FP64
1e-12 per pixel (same FP64 weights, only the atomic order differs);
compensated 1e-5 per pixel and 1e-6 on the total (the context path's
grouped kernel and the per-particle kernel use different FP32 weight
forms, and FP64 versus two-float accumulation).  Integer statistics must be
equal.  The first call is larger than one 2^20-particle chunk, so the
chunked pass through the scratch records is covered too.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import assert_map, rel_err_report

pytestmark = pytest.mark.gpu

TOL = {"fp64": (1e-12, 1e-12), "compensated": (1e-5, 1e-6)}
KEYS = ("kept", "subpixel", "fallback", "extended", "contributions", "large", "large_contributions")


def _context_map(pw, p, parts):
    import torch
    with pw.Projector(p, parts.shape[0]) as pr:
        pr.render(parts)
        out = torch.empty(pr.npix, dtype=torch.float64, device=parts.device)
        pr.get_map_device(out)
        st = pr.stats()
    return out.cpu().numpy(), st


@pytest.mark.parametrize("precision", ["compensated", "fp64"])
def test_noalloc_matches_context_path(pw, gpu, precision):
    import torch
    tol, ttol = TOL[precision]
    big = pw.synth.make("nfw_sub100_1e7", n=1_200_000 if precision == "compensated" else 300_000).to(gpu)
    small = pw.synth.make("nfw_1e5", n=20_000).to(gpu)
    cases = [(big, 256), (small, 64)]
    outs = []
    for parts, nside in cases:
        print(f"case noalloc precision={precision} n={parts.shape[0]} nside={nside}", flush=True)
        p = pw.params_for_observer(nside, precision)
        out = torch.full((12 * nside * nside,), 123.0, dtype=torch.float64, device=gpu)
        st = pw.capi.render_map_noalloc(p, parts, out, stats=True)
        outs.append((out.cpu().numpy(), st, p, parts))
    for got, st, p, parts in outs:
        ref, st_ref = _context_map(pw, p, parts)
        for k in KEYS:
            assert st[k] == st_ref[k], (k, st, st_ref)
        rep = rel_err_report(got, ref)
        print(rep, flush=True)
        assert_map(rep, tol, ttol)
    torch.cuda.synchronize()


def test_noalloc_async_and_domain_error(pw, gpu):
    """Without statistics the call only enqueues (the result is read on
    torch's stream); a particle at the observer gives DMANN_ERR_DOMAIN when
    statistics are asked for; bad arguments are rejected."""
    import torch
    print("case noalloc async + domain", flush=True)
    parts = pw.synth.make("nfw_1e5", n=5000).to(gpu)
    p = pw.params_for_observer(64, "compensated")
    out = torch.empty(12 * 64 * 64, dtype=torch.float64, device=gpu)
    assert pw.capi.render_map_noalloc(p, parts, out) is None
    ref, _ = _context_map(pw, p, parts)
    assert_map(rel_err_report(out.cpu().numpy(), ref), 1e-5, 1e-6)
    bad = parts.clone()
    bad[3, 3:6] = torch.tensor(list(p.opos), dtype=torch.float32, device=gpu)
    with pytest.raises(pw.DmannError) as ei:
        pw.capi.render_map_noalloc(p, bad, out, stats=True)
    assert ei.value.code == 4
    L = pw.load_library()
    q = p.copy()
    q.nside = 0
    assert L.dmann_render_map_noalloc(C.byref(q), C.c_void_p(parts.data_ptr()), parts.shape[0],
                                      C.c_void_p(out.data_ptr()), None, None) == 1
    assert b"nside" in L.dmann_last_error(None)
    assert L.dmann_render_map_noalloc(C.byref(p), None, 5, C.c_void_p(out.data_ptr()), None, None) == 1
    torch.cuda.synchronize()
