"""Seeded synthetic particle sets for the projector's workloads.

Units match pywrapper.capi.write_ini: positions in kpc, masses in Msun,
density in units of the critical density (h100 = 0.73).  Output is an
(n, 12) float32 tensor in DMParticle field order (mass, dens, hsmooth,
pos xyz, vel xyz, sigmav, eps, phi); velocities are zero (the nosom flux
does not use them).

Model (fixed here, recorded in the design notes):
  host    NFW, r_s = 20 kpc, truncated at c = 10 (200 kpc), isotropic,
          rho_s = N_host m / (4 pi r_s^3 mu(c)), mu(x) = ln(1+x) - x/(1+x)
  subhalo Plummer, M log-uniform in [1e6, 1e9] Msun, a uniform in
          [0.5, 3] kpc, centres at radius uniform in [10, 200] kpc,
          isotropic; round(M/m) particles each, truncated at 10 a
  near    particles uniform in a ball of radius near_radius around the
          observer (host density there)
  density host NFW (+ own Plummer for subhalo particles) x 10^(0.3 N(0,1))
  h       (32 m / ((4 pi / 3) rho))^(1/3)
Generation is chunked (2^22 particles per chunk, each chunk seeded from the
seed and its index) so a set is identical whatever device memory allows,
for a given device type.
"""
from __future__ import annotations

import math

import torch

RHO_CRIT_MSUN_KPC3 = None  # filled lazily (depends on h100 only)
CHUNK = 1 << 22


def rho_crit_msun_kpc3(h100: float = 0.73) -> float:
    """Critical density in Msun/kpc^3, from the same constants as the
    parameter chain (3 H^2 / 8 pi G with H = 100 h km/s/Mpc)."""
    Mpc_cm = 3.08568025e24
    G = 6.67259e-8
    Msun = 1.9889212e33
    rho_cgs = 3.0 * (100.0 * h100 / (Mpc_cm * 1e-5)) ** 2 / (8 * math.pi * G)
    kpc_cm = Mpc_cm * 1e-3
    return rho_cgs * kpc_cm ** 3 / Msun


def _mu(x):
    return torch.log1p(x) - x / (1 + x)


def _nfw_inverse_table(c: float, device, n: int = 8192):
    x = torch.logspace(-6, math.log10(c), n, dtype=torch.float64, device=device)
    x = torch.cat([torch.zeros(1, dtype=torch.float64, device=device), x])
    u = _mu(x) / _mu(torch.tensor(c, dtype=torch.float64, device=device))
    return u, x


def _interp(u_tab, x_tab, u):
    i = torch.searchsorted(u_tab, u).clamp(1, len(u_tab) - 1)
    u0, u1 = u_tab[i - 1], u_tab[i]
    x0, x1 = x_tab[i - 1], x_tab[i]
    t = ((u - u0) / (u1 - u0).clamp_min(1e-300)).clamp(0, 1)
    return x0 + t * (x1 - x0)


def _iso(n, g, device):
    v = torch.randn(n, 3, generator=g, device=device, dtype=torch.float64)
    return v / v.norm(dim=1, keepdim=True).clamp_min(1e-300)


class HaloModel:
    def __init__(self, n: int, seed: int, n_sub: int = 0, n_near: int = 0, observer=(8.0, 0.0, 0.0),
                 near_radius: float = 2.0, rs: float = 20.0, c: float = 10.0, mass: float = 4.1e3,
                 sigma_dex: float = 0.3, h100: float = 0.73, device="cpu"):
        self.n, self.seed, self.device = int(n), int(seed), torch.device(device)
        self.rs, self.c, self.m, self.sigma_dex = rs, c, mass, sigma_dex
        self.observer = torch.tensor(observer, dtype=torch.float64, device=self.device)
        self.near_radius = near_radius
        self.rho_crit = rho_crit_msun_kpc3(h100)
        g = torch.Generator(device="cpu").manual_seed(self.seed * 1000003 + 17)
        # subhalo catalogue (small; drawn on the CPU generator for every device)
        if n_sub:
            lm = torch.rand(n_sub, generator=g, dtype=torch.float64) * 3.0 + 6.0
            M = 10.0 ** lm
            a = torch.rand(n_sub, generator=g, dtype=torch.float64) * 2.5 + 0.5
            rad = torch.rand(n_sub, generator=g, dtype=torch.float64) * 190.0 + 10.0
            v = torch.randn(n_sub, 3, generator=g, dtype=torch.float64)
            ctr = v / v.norm(dim=1, keepdim=True) * rad[:, None]
            cnt = torch.round(M / mass).to(torch.int64)
        else:
            M = a = torch.zeros(0, dtype=torch.float64)
            ctr = torch.zeros(0, 3, dtype=torch.float64)
            cnt = torch.zeros(0, dtype=torch.int64)
        self.n_near = int(n_near)
        n_host = self.n - int(cnt.sum()) - self.n_near
        if n_host < 0:
            raise ValueError("subhalos and near-observer particles exceed n")
        self.n_host = n_host
        self.sub_M, self.sub_a, self.sub_ctr = M.to(self.device), a.to(self.device), ctr.to(self.device)
        self.sub_cnt = cnt
        # component boundaries: [host | sub_0 | sub_1 | ... | near]
        b = torch.cat([torch.tensor([0, n_host]), n_host + torch.cumsum(cnt, 0)])
        self.bounds = b.to(self.device)
        self.rho_s = n_host * mass / (4 * math.pi * rs ** 3 * float(_mu(torch.tensor(c, dtype=torch.float64))))
        self.u_tab, self.x_tab = _nfw_inverse_table(c, self.device)

    def host_density(self, pos):
        r = pos.norm(dim=1)
        x = (r / self.rs).clamp_min(1e-6)
        rho = self.rho_s / (x * (1 + x) ** 2)
        return torch.where(r <= self.rs * self.c, rho, torch.zeros_like(rho))

    def chunk(self, i0: int, i1: int) -> torch.Tensor:
        dev = self.device
        n = i1 - i0
        g = torch.Generator(device=dev).manual_seed(self.seed * 7919 + i0 // CHUNK + 1)
        idx = torch.arange(i0, i1, device=dev, dtype=torch.int64)
        comp = torch.searchsorted(self.bounds, idx, right=True) - 1  # 0 host, 1..S subhalo, S+1 near
        S = self.sub_M.numel()
        u = torch.rand(n, generator=g, device=dev, dtype=torch.float64)
        dirs = _iso(n, g, dev)
        pos = torch.empty(n, 3, dtype=torch.float64, device=dev)
        rho = torch.empty(n, dtype=torch.float64, device=dev)
        # host
        hm = comp == 0
        if hm.any():
            x = _interp(self.u_tab, self.x_tab, u[hm])
            pos[hm] = dirs[hm] * (x * self.rs)[:, None]
        # subhalos
        sm = (comp >= 1) & (comp <= S)
        if sm.any():
            k = comp[sm] - 1
            a = self.sub_a[k]
            umax = 1000.0 / 101.0 ** 1.5  # mass fraction inside 10 a
            us = (u[sm] * umax).clamp_min(1e-12)
            r = a / torch.sqrt(us ** (-2.0 / 3.0) - 1.0)
            pos[sm] = self.sub_ctr[k] + dirs[sm] * r[:, None]
        # near-observer ball
        nm = comp == S + 1
        if nm.any():
            r = self.near_radius * u[nm] ** (1.0 / 3.0)
            pos[nm] = self.observer + dirs[nm] * r[:, None]
        rho = self.host_density(pos)
        if sm.any():
            k = comp[sm] - 1
            a, M = self.sub_a[k], self.sub_M[k]
            rr = (pos[sm] - self.sub_ctr[k]).norm(dim=1)
            rho[sm] = rho[sm] + 3 * M / (4 * math.pi * a ** 3) * (1 + (rr / a) ** 2) ** (-2.5)
        rho = rho.clamp_min(self.m / (4.0 / 3.0 * math.pi * 200.0 ** 3))
        rho = rho * 10.0 ** (self.sigma_dex * torch.randn(n, generator=g, device=dev, dtype=torch.float64))
        h = (32.0 * self.m / ((4.0 * math.pi / 3.0) * rho)) ** (1.0 / 3.0)
        out = torch.zeros(n, 12, dtype=torch.float32, device=dev)
        out[:, 0] = self.m
        out[:, 1] = (rho / self.rho_crit).to(torch.float32)
        out[:, 2] = h.to(torch.float32)
        out[:, 3:6] = pos.to(torch.float32)
        return out

    def sample(self, m: int, windows: int = 16, near_min: int = 1) -> torch.Tensor:
        """m particles of the full model (same densities and smoothing
        lengths as the full set): `windows` evenly spaced contiguous index
        windows, plus the near-observer particles pro rata (at least
        near_min of them)."""
        m = int(min(m, self.n))
        near = 0
        if self.n_near:
            near = max(near_min, int(round(m * self.n_near / self.n)))
        body = m - near
        per = max(1, body // windows)
        n_body = self.n - self.n_near
        parts = []
        got = 0
        for w in range(windows):
            if got >= body:
                break
            take = min(per if w < windows - 1 else body - got, body - got)
            i0 = int((n_body - take) * w / max(1, windows - 1))
            parts.append(self.chunk(i0, i0 + take))
            got += take
        if near:
            parts.append(self.chunk(self.n - near, self.n))
        return torch.cat(parts, 0)

    def generate(self, out: torch.Tensor | None = None) -> torch.Tensor:
        if out is None:
            out = torch.empty(self.n, 12, dtype=torch.float32, device=self.device)
        for i0 in range(0, self.n, CHUNK):
            i1 = min(self.n, i0 + CHUNK)
            out[i0:i1] = self.chunk(i0, i1)
        return out


# The workloads, by what they are.
WORKLOADS = {
    # NFW host only, 1e5 particles (CPU-runnable case), nside 64, seed 0
    "nfw_1e5": dict(n=100_000, seed=0, n_sub=0, n_near=0),
    # NFW host + 100 Plummer subhalos, 1e7 particles, seed 1
    "nfw_sub100_1e7": dict(n=10_000_000, seed=1, n_sub=100, n_near=0),
    # NFW host + 10 subhalos, 1e6 particles (adjoint case), seed 3
    "nfw_sub10_1e6": dict(n=1_000_000, seed=3, n_sub=10, n_near=0),
    # NFW host + 2000 subhalos + 1e4 particles within 2 kpc of the observer,
    # 1.1e9 particles, seed 4
    "nfw_sub2000_near_1p1e9": dict(n=1_100_000_000, seed=4, n_sub=2000, n_near=10_000),
}


def model(workload: str, device="cpu", **over) -> HaloModel:
    kw = dict(WORKLOADS[workload])
    kw.update(over)
    return HaloModel(device=device, **kw)


def make(workload: str, device="cpu", n: int | None = None, **over) -> torch.Tensor:
    """Particles of a named workload; with n, a subsample of n particles of
    the full-size model (HaloModel.sample), so footprints keep their size."""
    hm = model(workload, device, **over)
    if n is not None and n < hm.n:
        return hm.sample(n)
    return hm.generate()


def to_records(t: torch.Tensor):
    """(n, 12) float32 tensor -> numpy DMParticle record array (host)."""
    from .capi import PARTICLE_DTYPE
    a = t.detach().to("cpu").contiguous().numpy()
    return a.view(PARTICLE_DTYPE).reshape(-1)
