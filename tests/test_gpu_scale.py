"""The production path at realistic sizes against the FP64 GPU form.

The FP64 form (FP64 acos/exp weights, FP64 atomics, one particle per
wavefront) is pinned against the CPU oracle in test_gpu_parity.py; here it
is the reference for samples too large for the oracle to finish quickly:
5e6 particles of the 1.1e9-particle model at nside 2048 (its footprints, so
the grouped kernel runs multi-band windows, tall members and dense sort
cells), rendered in chunks that accumulate like the reference's
calculateMapByGPU calls. Tolerances as in test_gpu_parity.py: compensated
1e-5 per pixel (above 1e-6 of the map maximum) and 1e-6 on the total; the
integer statistics must match exactly.
"""
import math

import numpy as np
import pytest

from conftest import assert_map, rel_err_report

pytestmark = pytest.mark.gpu


def _render(pw, parts, nside, precision, k=None, chunks=1):
    p = pw.params_for_observer(nside, precision)
    n = parts.shape[0]
    cap = (n + chunks - 1) // chunks
    with pw.Projector(p, cap) as pr:
        if k is not None:
            pr.set_grouping(k)
        for i in range(0, n, cap):
            pr.render(parts[i:i + cap])
        m = pr.get_map()
        st = pr.stats()
    return m, st


@pytest.mark.parametrize("k", [32, 64])
def test_sample_5e6_nside2048_compensated_vs_fp64(pw, gpu, k):
    import torch
    print(f"case scale 5e6 nside 2048 k={k}", flush=True)
    hm = pw.synth.model("nfw_sub2000_near_1p1e9", "cuda")
    parts = hm.sample(5_000_000)
    torch.cuda.synchronize()
    ref, st_ref = _render(pw, parts, 2048, "fp64")
    got, st = _render(pw, parts, 2048, "compensated", k=k, chunks=3)
    for key in ("kept", "subpixel", "fallback", "extended", "contributions", "large"):
        assert st[key] == st_ref[key], key
    rep = rel_err_report(got, ref)
    print(rep, st, flush=True)
    assert_map(rep, 1e-5, 1e-6)
    # invariant of src/genmap.cpp:161-172: sum(map) dOmega = sum of the
    # kept particles' fluxes (FP64 from the same inputs)
    p = pw.params_for_observer(2048, "fp64")
    opos = torch.tensor(list(p.opos), dtype=torch.float64, device=parts.device)
    keep = (parts[:, 0].double() < p.mass_cut) & (parts[:, 1] >= 0)
    d2 = ((parts[:, 3:6].double() - opos) ** 2).sum(1)
    flux = p.flux_factor * parts[:, 1].double() * parts[:, 0].double() / (4.0 * math.pi * d2)
    total = float(torch.where(keep, flux, torch.zeros_like(flux)).sum())
    d_omega = 4.0 * math.pi / (12 * 2048 * 2048)
    assert abs(float(np.sum(got)) * d_omega - total) / total <= 1e-6
    torch.cuda.synchronize()
