#!/usr/bin/env python3
"""Throughput of the MI355X annihilation-map projector.

Metric: Mparticles/s into an nside-2048 HEALPix map (compensated FP32, with
plain FP32 timed beside it).  A step is one full render of the workload's
particles, already resident in HBM, into a zeroed map, including the
pre-pass, both footprint kernels and the finalised FP64 RING map on the
device.

Default workload: 1.1e9 synthetic particles (NFW host r_s = 20 kpc + 2000
Plummer subhalos + 1e4 particles within 2 kpc of the observer at
(8, 0, 0) kpc), nside 2048.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dm-annihilation-cxx_amd"))

METRIC = "Mparticles/s into nside-2048 map, compensated vs plain FP32, ≤1e-5 rel err"

# MI355X peaks used for the roofline
HBM_PEAK_GBS = 8000.0
# FP32 VALU lane-op issue rate without packed FP32: 256 CU x 4 SIMD x 16 lanes x 2.4 GHz
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
# lane-op slots per weight evaluation and evaluations per contribution
# (norm pass + deposit pass), the algorithmic work per footprint pixel
SLOTS_PER_EVAL = 16
EVALS_PER_CONTRIB = 2


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="nfw_sub2000_near_1p1e9")
    ap.add_argument("--n", type=int, default=None, help="override particle count (model sample)")
    ap.add_argument("--nside", type=int, default=2048)
    ap.add_argument("--precision", default="compensated", choices=["compensated", "fp32", "fp64"])
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP32 timing")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    return ap.parse_args()


def time_form(pw, torch, parts, nside, precision, steps, warmup):
    p = pw.params_for_observer(nside, precision)
    n = parts.shape[0]
    out = torch.empty(12 * nside * nside, dtype=torch.float64, device="cuda")
    pr = pw.Projector(p, n)
    try:
        for _ in range(warmup):
            pr.reset()
            pr.render(parts)
            pr.get_map_device(out)
        torch.cuda.synchronize()
        stats = pr.stats()
        pr.timing(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            pr.reset()
            pr.render(parts)
            pr.get_map_device(out)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        kms, calls = pr.timing(False)
        pr.synchronize()
        if warmup == 0:
            stats = {k: v // max(1, steps) for k, v in pr.stats().items()}
    finally:
        pr.close()
    sec = (t1 - t0) / steps
    return {"sec": sec, "kernel_ms": {k: v / max(1, calls) for k, v in kms.items()}, "stats": stats,
            "map": out}


def total_flux_check(torch, parts, params, m, nside):
    """Relative error of sum(map)*dOmega against sum(flux) of the kept
    particles, both in FP64 (the invariant of src/genmap.cpp:161-172)."""
    K = params.flux_factor
    opos = torch.tensor(list(params.opos), dtype=torch.float64, device=parts.device)
    tot = torch.zeros((), dtype=torch.float64, device=parts.device)
    for i in range(0, parts.shape[0], 1 << 24):
        c = parts[i:i + (1 << 24)]
        keep = (c[:, 0].double() < params.mass_cut) & (c[:, 1] >= 0)
        d2 = ((c[:, 3:6].double() - opos) ** 2).sum(1)
        f = K * c[:, 1].double() * c[:, 0].double() / (4.0 * math.pi * d2)
        tot += torch.where(keep, f, torch.zeros_like(f)).sum()
    dOmega = 4 * math.pi / (12 * nside * nside)
    got = float(m.sum()) * dOmega
    return abs(got - float(tot)) / float(tot)


def cpu_baseline(pw, args, nside):
    """The oracle (OpenMP, all host cores given to the job) on a bounded
    sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    oracle.build()
    cores = min(16, os.cpu_count() or 1)
    hm = pw.synth.model(args.workload, "cpu")
    ini = os.path.join("/tmp", f"dmann_bench_{os.getpid()}.ini")
    pw.write_ini(ini, nside=nside)
    cfg = oracle.config_from_params(oracle.params_from_ini(ini))
    os.unlink(ini)
    # calibrate on a small sample, then size the timed sample to ~cpu_seconds
    m = 2000
    recs = pw.synth.to_records(hm.sample(m))
    t0 = time.perf_counter()
    oracle.genmap(recs, cfg, nthreads=cores, finalize=False)
    dt = time.perf_counter() - t0
    m2 = int(max(2000, min(hm.n, m * args.cpu_seconds / max(dt, 1e-3))))
    recs = pw.synth.to_records(hm.sample(m2))
    t0 = time.perf_counter()
    _, st = oracle.genmap(recs, cfg, nthreads=cores, finalize=False)
    dt = time.perf_counter() - t0
    return {"value": m2 / dt / 1e6, "unit": "Mparticles/s", "cores": cores, "kind": "port",
            "sample": f"{m2} particles of the same model (16 index windows + near-observer share), "
                      f"nside {nside}, {st['contributions']} footprint pixels, {dt:.1f} s with "
                      f"{cores} OpenMP threads"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only "
                          "(see the design notes); run with --gpus 1"}))
        sys.exit(2)
    import torch
    import pywrapper as pw
    pw.load_library()
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU"}))
        sys.exit(1)
    nside = args.nside
    hm = pw.synth.model(args.workload, "cuda")
    n = hm.n if args.n is None else args.n
    log(f"generating {n} particles of {args.workload} on the GPU")
    t0 = time.perf_counter()
    parts = hm.generate() if n == hm.n else hm.sample(n)
    torch.cuda.synchronize()
    log(f"generated in {time.perf_counter() - t0:.1f}s")

    res = time_form(pw, torch, parts, nside, args.precision, args.steps, args.warmup)
    log(f"{args.precision}: {res['sec'] * 1e3:.1f} ms/step, kernels {res['kernel_ms']}, stats {res['stats']}")
    params = pw.params_for_observer(nside, args.precision)
    flux_err = total_flux_check(torch, parts, params, res["map"], nside)
    plain = None
    if not args.no_plain and args.precision != "fp32":
        pl = time_form(pw, torch, parts, nside, "fp32", args.steps, args.warmup)
        pl_err = total_flux_check(torch, parts, params, pl["map"], nside)
        log(f"fp32: {pl['sec'] * 1e3:.1f} ms/step, kernels {pl['kernel_ms']}")
        plain = {"value": n / pl["sec"] / 1e6, "ms_per_step": pl["sec"] * 1e3,
                 "medium_kernel_ms": pl["kernel_ms"]["medium"], "total_flux_rel_err": pl_err,
                 "compensated_over_plain_time": res["sec"] / pl["sec"]}
        del pl
    C = res["stats"]["contributions"]
    kms = res["kernel_ms"]
    t_med = kms["medium"] / 1e3
    # dominant kernel: the medium footprint kernel
    frac_med = 1.0 - res["stats"]["large"] / max(1, res["stats"]["extended"])
    valu_ops = EVALS_PER_CONTRIB * SLOTS_PER_EVAL * C
    s_bytes = 8 if args.precision != "fp32" else 4
    hbm_bytes = 48 * n + 2 * s_bytes * 12 * nside * nside
    achieved_valu = valu_ops / (t_med + kms["large"] / 1e3) / 1e12
    roofline = {
        "bound": "valu",
        "achieved": achieved_valu,
        "peak": VALU_PEAK_TOPS,
        "unit": "Tlane-op/s",
        "frac": achieved_valu / VALU_PEAK_TOPS,
        "traffic": None,
        "kernel": "k_disc_medium + k_disc_large (footprint kernels)",
        "kernel_ms": kms,
        "contributions": C,
        "hbm": {"achieved": hbm_bytes / res["sec"] / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": hbm_bytes / res["sec"] / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes": hbm_bytes},
    }
    del frac_med
    line = {
        "metric": METRIC,
        "value": n / res["sec"] / 1e6,
        "unit": "Mparticles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["sec"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision != "fp64" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n} particles, NFW host + subhalos"
                               + (" + 1e4 near-observer" if "near" in args.workload else "")
                               + f", observer (8,0,0) kpc, nside {nside}",
                   "nside": nside, "n_particles": n, "precision": args.precision,
                   "contributions": C},
        "total_flux_rel_err": flux_err,
        "plain_fp32": plain,
        "roofline": roofline,
    }
    if not args.no_cpu:
        del parts
        torch.cuda.empty_cache()
        line["cpu_baseline"] = cpu_baseline(pw, args, nside)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
