"""Adjoint kernel against FP64 central finite differences.

Workload: 1e6 particles, NFW host + 10 Plummer subhalos (seed 3), nside 256,
observer (8,0,0) kpc.  Target: the same particles with masses x lognormal
(sigma 0.1 in ln) and positions + N(0, 0.05 kpc) (seed 5), rendered in FP64.
Loss = sum over pixels with target > 0 of (map - target)^2 / target.

Finite differences: for 1000 random particles (seed 6), central differences
with relative step 1e-4 (of m, rho, h; of d for positions) and 5e-5,
combined by Richardson extrapolation (truncation error O(step^4): plain
central differences at 1e-4 are off by ~1e-3 relative for particles whose
position gradient nearly cancels), evaluated in FP64 with the particle's
footprint held fixed (the map is discontinuous where pixel centres cross
the footprint edge; the adjoint differentiates the smooth part).
Required: relative error <= 1e-4 per quantity (positions as a 3-vector).

Exact reference: for the same 1000 particles, torch.autograd through an FP64
torch restatement of the particle's contribution on the oracle's footprint
(pixel centres from the oracle's pix2vec), contracted with the same dL/dmap:
the adjoint kernel must agree to 1e-9 relative per quantity (positions as a
3-vector).  The FP64 map path itself is pinned against the CPU oracle by
test_gpu_parity.py.
"""
import math

import numpy as np
import pytest

from conftest import record

pytestmark = pytest.mark.gpu

NSIDE = 256
C_W = 0.5 / 0.333


def _contrib(q, pix_u, cfg_R, opos, K, dOmega):
    """FP64 map contribution of particle q on a fixed pixel set (unit
    vectors pix_u): flux * w / sum(w) / dOmega."""
    m, rho, h = q[0], q[1], q[2]
    dv = q[3:6] - opos
    d = math.sqrt(float(dv @ dv))
    flux = K * rho * m / (4.0 * math.pi * d * d)
    t = cfg_R.T @ dv
    v = np.array([-t[0], -t[1], t[2]]) / d
    r = h / d
    psi = np.arctan2(np.linalg.norm(np.cross(pix_u, v), axis=1), pix_u @ v)
    w = np.exp(-C_W * psi ** 2 / r ** 2)
    return flux * w / w.sum() / dOmega


def _autograd_reference(q, pix_u, R, opos, K, dOmega, gpix):
    """d/d(m, rho, h, x, y, z) of sum_p gpix_p c_p(q), c as in _contrib, by
    torch.autograd in FP64."""
    import torch
    qt = torch.tensor(q[:6], dtype=torch.float64, requires_grad=True)
    u = torch.from_numpy(np.ascontiguousarray(pix_u, dtype=np.float64))
    m, rho, h = qt[0], qt[1], qt[2]
    dv = qt[3:6] - torch.from_numpy(np.asarray(opos, dtype=np.float64))
    d = torch.sqrt((dv * dv).sum())
    flux = K * rho * m / (4.0 * math.pi * d * d)
    t = torch.from_numpy(np.ascontiguousarray(R.T)) @ dv
    v = torch.stack([-t[0], -t[1], t[2]]) / d
    r = h / d
    cr = torch.linalg.cross(u, v.expand_as(u), dim=1)
    psi = torch.atan2(cr.norm(dim=1), u @ v)
    w = torch.exp(-C_W * psi ** 2 / r ** 2)
    c = flux * w / w.sum() / dOmega
    (c * torch.from_numpy(np.ascontiguousarray(gpix, dtype=np.float64))).sum().backward()
    return qt.grad.numpy()


def test_adjoint_vs_finite_differences(pw, orc, gpu, tmp_path):
    import torch
    print("case adjoint nfw_sub10_1e6 nside=256", flush=True)
    hm = pw.synth.model("nfw_sub10_1e6", "cpu")
    parts = hm.generate()
    rng = np.random.default_rng(5)
    pert = parts.clone()
    pert[:, 0] *= torch.from_numpy(np.exp(0.1 * rng.standard_normal(len(parts))).astype(np.float32))
    pert[:, 3:6] += torch.from_numpy((0.05 * rng.standard_normal((len(parts), 3))).astype(np.float32))
    p = pw.params_for_observer(NSIDE, "fp64")
    dev_parts = parts.to(gpu).requires_grad_(True)
    with pw.Projector(p, len(parts)) as pr:
        pr.render(pert.to(gpu))
        target = torch.empty(pr.npix, dtype=torch.float64, device=gpu)
        pr.get_map_device(target)
        torch.cuda.synchronize()
        m = pw.project_map(dev_parts, pr)
        mask = target > 0
        loss = torch.where(mask, (m - target) ** 2 / torch.where(mask, target, torch.ones_like(target)),
                           torch.zeros_like(m)).sum()
        loss.backward()
        torch.cuda.synchronize()
        grads = dev_parts.grad.detach().cpu().numpy().astype(np.float64)
        g6 = pr.adjoint(dev_parts.detach(), torch.where(mask, 2 * (m - target) / torch.where(
            mask, target, torch.ones_like(target)), torch.zeros_like(m)).detach()).cpu().numpy()
    M = m.detach().cpu().numpy()
    T = target.cpu().numpy()
    R = np.array(list(p.rotmatrix)).reshape(3, 3)
    opos = np.array(list(p.opos))
    dOmega = 4 * math.pi / (12 * NSIDE * NSIDE)
    ini = pw.write_ini(str(tmp_path / "a.ini"), nside=NSIDE)
    cfg = orc.config_from_params(orc.params_from_ini(ini))
    recs = pw.synth.to_records(parts)
    pick = np.random.default_rng(6).choice(len(parts), 1000, replace=False)
    worst = {"m": 0.0, "rho": 0.0, "h": 0.0, "pos": 0.0}
    exact = {"m": 0.0, "rho": 0.0, "h": 0.0, "pos": 0.0}
    Gfull = np.where(T > 0, 2 * (M - T) / np.where(T > 0, T, 1.0), 0.0)
    crossed = 0
    for i in pick:
        q = parts[i].numpy().astype(np.float64)
        pix, wn, flux = orc.particle_footprint(recs[i], cfg)
        if len(pix) < 2:
            continue
        u = orc.pix2vec(NSIDE, pix)
        tF = T[pix]
        keep = tF > 0
        base = M[pix] - _contrib(q, u, R, opos, cfg.flux_factor, dOmega)
        d = float(np.linalg.norm(q[3:6] - opos))

        def fd1(k, step):
            qp, qm = q.copy(), q.copy()
            qp[k] += step
            qm[k] -= step
            cp = _contrib(qp, u, R, opos, cfg.flux_factor, dOmega)
            cm = _contrib(qm, u, R, opos, cfg.flux_factor, dOmega)
            dl = ((cp - cm) * (2 * (base - tF) + cp + cm))[keep] / tF[keep]
            return dl.sum() / (2 * step)

        def fd(k, step):
            return (4.0 * fd1(k, 0.5 * step) - fd1(k, step)) / 3.0

        fm = fd(0, 1e-4 * q[0])
        fr = fd(1, 1e-4 * q[1])
        fh = fd(2, 1e-4 * q[2])
        fp = np.array([fd(3 + a, 1e-4 * d) for a in range(3)])
        ga = _autograd_reference(q, u, R, opos, cfg.flux_factor, dOmega, Gfull[pix])
        exact["m"] = max(exact["m"], abs(g6[i][0] - ga[0]) / abs(ga[0]))
        exact["rho"] = max(exact["rho"], abs(g6[i][1] - ga[1]) / abs(ga[1]))
        exact["h"] = max(exact["h"], abs(g6[i][2] - ga[2]) / max(abs(ga[2]), 1e-300))
        exact["pos"] = max(exact["pos"], float(np.linalg.norm(g6[i][3:6] - ga[3:6]) / np.linalg.norm(ga[3:6])))
        g = grads[i]
        assert np.allclose(g6[i].astype(np.float32), grads[i][:6], rtol=1e-6, atol=0)
        assert not grads[i][6:].any()
        worst["m"] = max(worst["m"], abs(g6[i][0] - fm) / abs(fm))
        worst["rho"] = max(worst["rho"], abs(g6[i][1] - fr) / abs(fr))
        worst["h"] = max(worst["h"], abs(g6[i][2] - fh) / max(abs(fh), 1e-300))
        worst["pos"] = max(worst["pos"], float(np.linalg.norm(g6[i][3:6] - fp) / np.linalg.norm(fp)))
        # membership change under the h step (informational)
        r2 = recs[i].copy()
        r2["hsmooth"] = np.float32(q[2] * (1 + 1e-4))
        if len(orc.particle_footprint(r2, cfg)[0]) != len(pix):
            crossed += 1
        del g
    print("worst relative errors vs FD:", worst, "footprints changed by the h step:", crossed)
    print("worst relative errors vs the FP64 autograd reference:", exact)
    record("adjoint 1e6 nside256 vs FP64 autograd", **{f"max_rel_{k}": v for k, v in exact.items()})
    record("adjoint 1e6 nside256 vs finite differences", **{f"max_rel_{k}": v for k, v in worst.items()})
    for k, v in exact.items():
        assert v <= 1e-9, (k, v)
    for k, v in worst.items():
        assert v <= 1e-4, (k, v)
