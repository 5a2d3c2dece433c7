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

# Memory traffic of one launch of the grouped kernel, from two rocprofv3
# counter passes (FETCH_SIZE, WRITE_SIZE in KiB) scaled by what the same
# counters read for k_finalize, a streaming kernel of known size (fetch x2,
# write x1): profiles/k_disc_group_nfw_sub2000_near_1p1e9_nside2048_traffic.md
MEASURED_TRAFFIC = {
    ("nfw_sub2000_near_1p1e9", 2048, "compensated"): (1.2344e9 * 2.0 + 2.7896e9 * 1.0) * 1024,
}


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
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-observer batch leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint (backward) leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the FP64 GPU map that the per-pixel errors are measured against")
    ap.add_argument("--fp64-chunks", type=int, default=1,
                    help="render the FP64 reference map in this many particle chunks (accumulating into one "
                         "map), with a progress line after each; for maps whose FP64 render takes minutes")
    ap.add_argument("--host-n", type=int, default=0,
                    help="also time N particles handed over in pinned host memory (PCIe-inclusive rate)")
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
        for k in range(steps):
            pr.reset()
            pr.render(parts)
            pr.get_map_device(out)
            if k % 5 == 4 and k + 1 < steps:
                # a progress line: wait for the queued steps first (the
                # timed region stays one stream of renders; the wait adds
                # no GPU work)
                torch.cuda.synchronize()
                log(f"{precision}: {k + 1} of {steps} timed steps, {time.perf_counter() - t0:.1f} s")
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


def pixel_errors(torch, m, ref):
    """Per-pixel relative error on pixels above 1e-6 of the map maximum and
    the relative error of the total, against the FP64 map."""
    big = ref > 1e-6 * ref.max()
    rel = ((m - ref).abs() / ref.abs().clamp_min(1e-300))[big]
    return {"max_rel": float(rel.max()), "pixels_checked": int(big.sum()),
            "total_rel": float(abs(m.sum() - ref.sum()) / ref.sum())}


def host_input_rate(pw, torch, parts, nside, precision, n):
    """Mparticles/s when the particles arrive in pinned host memory, the
    host-buffer call of the C ABI: H2D copy of the 48-byte records + render."""
    import numpy as np
    sample = parts[:n].cpu().pin_memory()
    host = sample.numpy()
    p = pw.params_for_observer(nside, precision)
    out = torch.empty(12 * nside * nside, dtype=torch.float64, device="cuda")
    pr = pw.Projector(p, n)
    try:
        pr.render(host[:min(n, 1 << 20)])
        pr.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pr.render(host)
        pr.get_map_device(out)
        torch.cuda.synchronize()
        t_host = time.perf_counter() - t0
        dev = sample.cuda()
        pr.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pr.render(dev)
        pr.get_map_device(out)
        torch.cuda.synchronize()
        t_dev = time.perf_counter() - t0
    finally:
        pr.close()
    del np
    return {"n": n, "value": n / t_host / 1e6, "device_resident_value": n / t_dev / 1e6,
            "h2d_GBps": 48 * n / max(t_host - t_dev, 1e-9) / 1e9, "unit": "Mparticles/s"}


def fp64_reference(pw, torch, parts, nside, chunks):
    """The FP64 form (the GPU oracle) over all particles, rendered in
    `chunks` pieces that accumulate into one map (renders add into the map
    until it is reset), with a progress line after each piece."""
    if chunks <= 1:
        return time_form(pw, torch, parts, nside, "fp64", 1, 0)
    n = parts.shape[0]
    step = (n + chunks - 1) // chunks
    out = torch.empty(12 * nside * nside, dtype=torch.float64, device="cuda")
    pr = pw.Projector(pw.params_for_observer(nside, "fp64"), step)
    try:
        pr.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(0, n, step):
            pr.render(parts[i:i + step])
            pr.synchronize()
            log(f"fp64 reference: {min(n, i + step)} of {n} particles rendered, {time.perf_counter() - t0:.0f} s")
        pr.get_map_device(out)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
    finally:
        pr.close()
    return {"sec": sec, "map": out}


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


def cpu_baseline(pw, args, nside, n_full, c_full):
    """The oracle (OpenMP, all host cores given to the job) on a bounded
    sample of the same workload, sized to about --cpu-seconds; its
    footprint-pixel rate gives the time of the whole workload (work is
    proportional to the footprint pixels), reported as Mparticles/s.  The
    sample takes the near-observer particles only pro rata (none in a small
    sample): one sky-covering particle would otherwise hold a thread for the
    whole run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.build()
    # the job's share of the host cores (OMP_NUM_THREADS where the machine
    # sets it, else every core); nproc is reported beside it
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    hm = pw.synth.model(args.workload, "cpu")
    ini = os.path.join("/tmp", f"dmann_bench_{os.getpid()}.ini")
    pw.write_ini(ini, nside=nside)
    cfg = oracle.config_from_params(oracle.params_from_ini(ini))
    os.unlink(ini)

    def timed(m, threads):
        recs = pw.synth.to_records(hm.sample(m, near_min=0))
        t0 = time.perf_counter()
        _, st = oracle.genmap(recs, cfg, nthreads=threads, finalize=False)
        return time.perf_counter() - t0, st["contributions"]

    def sized(threads, seconds):
        # marginal rate between a small and a large sample, so that the
        # per-thread map set-up and reduction (fixed costs) cancel; the
        # large sample grows until its extra time reaches half the target
        dt_a, c_a = timed(1000, threads)
        m = 4000
        for _ in range(8):
            dt, c = timed(m, threads)
            if dt - dt_a >= 0.5 * seconds or m >= hm.n:
                break
            m = int(min(hm.n, m * min(8.0, seconds / max(dt - dt_a, 1e-2))))
        return m, dt, c, (c - c_a) / max(dt - dt_a, 1e-9)

    m, dt, c, rate = sized(cores, args.cpu_seconds)
    m1, dt1, c1, rate1 = sized(1, args.cpu_seconds / 3)
    # the CPU-runnable case in full: 1e5 particles of the NFW host, nside 64
    small = pw.synth.to_records(pw.synth.make("nfw_1e5"))
    ini = os.path.join("/tmp", f"dmann_bench_{os.getpid()}_64.ini")
    pw.write_ini(ini, nside=64)
    cfg64 = oracle.config_from_params(oracle.params_from_ini(ini))
    os.unlink(ini)
    full = {}
    for threads in (1, cores):
        t0 = time.perf_counter()
        _, st64 = oracle.genmap(small, cfg64, nthreads=threads)
        sec = time.perf_counter() - t0
        full[f"threads_{threads}"] = {"seconds": sec, "value": len(small) / sec / 1e6}
    full["workload"] = f"nfw_1e5: {len(small)} particles (NFW host), nside 64, {st64['contributions']} footprint pixels"
    return {"value": n_full / (c_full / rate) / 1e6, "unit": "Mparticles/s", "cores": cores, "kind": "port",
            "nproc": os.cpu_count(),
            "sample": f"{m} particles of the same model (16 index windows, near-observer share pro rata), "
                      f"nside {nside}: "
                      f"{c} footprint pixels in {dt:.1f} s on {cores} OpenMP threads, marginal rate over a "
                      f"1000-particle sample {rate:.3e} pixels/s, "
                      f"extrapolated to the workload's {c_full} footprint pixels ({c_full / rate:.0f} s)",
            "sample_value": m / dt / 1e6,
            "single_core": {"value": n_full / (c_full / rate1) / 1e6,
                            "sample": f"{m1} particles, {c1} footprint pixels in {dt1:.1f} s on 1 thread, "
                                      f"marginal rate {rate1:.3e} pixels/s"},
            "full_small_case": full}


def batch_leg(pw, torch):
    """64 observers on the 8 kpc circle in the z = 0 plane, 1e7 particles of
    the NFW + 100 subhalo set, nside of each map drawn (seed 2) from
    {256, 512, 1024}, compensated: one batch replayed from a hipGraph against
    64 single-observer renders of the same maps."""
    import numpy as np
    nmaps = 64
    rng = np.random.default_rng(2)
    nsides = [int(x) for x in rng.choice([256, 512, 1024], nmaps)]
    parts = pw.synth.make("nfw_sub100_1e7", "cuda")
    plist = []
    for k in range(nmaps):
        a = 2 * math.pi * k / nmaps
        plist.append(pw.params_for_observer(nsides[k], "compensated", obs_pos=(8 * math.cos(a), 8 * math.sin(a), 0.0)))
    with pw.Batch(plist) as b:
        b.render(parts, use_graph=True)
        b.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.render(parts, use_graph=True)
        b.synchronize()
        t_graph = time.perf_counter() - t0
        contrib = sum(b.stats(i)["contributions"] for i in range(nmaps))
    t_single = 0.0
    for k in range(nmaps):
        with pw.Projector(plist[k], parts.shape[0]) as pr:
            out = torch.empty(pr.npix, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pr.render(parts)
            pr.get_map_device(out)
            torch.cuda.synchronize()
            t_single += time.perf_counter() - t0
    n = parts.shape[0]
    return {"workload": f"nfw_sub100_1e7: {n} particles, {nmaps} observers on the 8 kpc circle, nside from "
                        "{256, 512, 1024} (seed 2), compensated",
            "graph_ms": 1e3 * t_graph, "single_renders_ms": 1e3 * t_single,
            "value": nmaps * n / t_graph / 1e6, "unit": "Mparticle-maps/s", "contributions": contrib}


# Algorithmic FP64 work the adjoint does per footprint pixel, counted from
# its formulas (k_adjoint; per-particle constants such as 1/r^2 hoisted):
# pixel longitude and vector, cos of the angle (dot product) and sin (norm of
# the cross product), the weight and its r- and direction-derivatives, and
# the ten running sums -- 54 FP64 operations, plus sin, cos, sqrt, atan2,
# exp and one division counted as 10 operations each: 114 per pixel.
ADJOINT_FLOP_PER_PIXEL = 114
FP64_PEAK_TFLOPS = 78.6


def adjoint_leg(pw, torch):
    """Backward pass of the adjoint workload: 1e6 particles of the NFW + 10
    subhalo set (seed 3), nside 256, loss = sum (map - target)^2 / target."""
    hm = pw.synth.model("nfw_sub10_1e6", "cuda")
    parts = hm.generate()
    target_parts = parts.clone()
    target_parts[:, 0] *= 1.1
    p = pw.params_for_observer(256, "fp64")
    with pw.Projector(p, parts.shape[0]) as pr:
        target = pw.project_map(target_parts, pr).detach().clone()
        x = parts.clone().requires_grad_(True)
        m = pw.project_map(x, pr)
        mask = target > 0
        loss = torch.where(mask, (m - target) ** 2 / torch.where(mask, target, torch.ones_like(target)),
                           torch.zeros_like(m)).sum()
        loss.backward()  # warm
        g = torch.where(mask, 2 * (m - target) / torch.where(mask, target, torch.ones_like(target)),
                        torch.zeros_like(m)).detach()
        out = torch.empty(parts.shape[0], 6, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pr.adjoint(parts, g, out)
        torch.cuda.synchronize()
        t_bwd = time.perf_counter() - t0
        t0 = time.perf_counter()
        pr.reset()
        pr.render(parts)
        pr.get_map_device(torch.empty(pr.npix, dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        C = pr.stats()["contributions"]
    achieved = ADJOINT_FLOP_PER_PIXEL * C / t_bwd / 1e12
    return {"workload": f"nfw_sub10_1e6: {parts.shape[0]} particles, nside 256, FP64",
            "backward_ms": 1e3 * t_bwd, "forward_fp64_ms": 1e3 * t_fwd, "contributions": C,
            "roofline": {"bound": "fp64 valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP64_PEAK_TFLOPS, "flop_per_pixel": ADJOINT_FLOP_PER_PIXEL}}


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
    pl_map = None
    if not args.no_plain and args.precision != "fp32":
        # a side figure: a few steps suffice
        pl_steps, pl_warm = min(args.steps, 3), min(args.warmup, 1)
        pl = time_form(pw, torch, parts, nside, "fp32", pl_steps, pl_warm)
        pl_err = total_flux_check(torch, parts, params, pl["map"], nside)
        log(f"fp32: {pl['sec'] * 1e3:.1f} ms/step, kernels {pl['kernel_ms']}")
        plain = {"value": n / pl["sec"] / 1e6, "ms_per_step": pl["sec"] * 1e3, "steps": pl_steps,
                 "kernel_ms": pl["kernel_ms"], "total_flux_rel_err": pl_err,
                 "compensated_over_plain_time": res["sec"] / pl["sec"]}
        pl_map = pl["map"]
        del pl
    vs_fp64 = None
    if not args.no_fp64 and args.precision != "fp64":
        # the FP64 form (FP64 acos/exp weights, FP64 atomics) is the GPU oracle;
        # one untimed-warmup-free render is enough for the error figures
        ref = fp64_reference(pw, torch, parts, nside, args.fp64_chunks)
        log(f"fp64: {ref['sec'] * 1e3:.1f} ms/step")
        vs_fp64 = {args.precision: pixel_errors(torch, res["map"], ref["map"]),
                   "fp64_ms": ref["sec"] * 1e3,
                   "fp64_total_flux_rel_err": total_flux_check(torch, parts, params, ref["map"], nside)}
        if pl_map is not None:
            vs_fp64["fp32"] = pixel_errors(torch, pl_map, ref["map"])
        log(f"vs fp64: {vs_fp64}")
        del ref
    del pl_map
    host_in = None
    if args.host_n > 0:
        host_in = host_input_rate(pw, torch, parts, nside, args.precision, min(args.host_n, n))
        log(f"host input: {host_in}")
    C = res["stats"]["contributions"]
    C_large = res["stats"]["large_contributions"]
    C_med = C - C_large
    kms = res["kernel_ms"]
    # dominant kernel: the medium footprint kernel (grouped), timed alone
    t_med = kms["medium"] / 1e3
    achieved_valu = EVALS_PER_CONTRIB * SLOTS_PER_EVAL * C_med / t_med / 1e12
    both = EVALS_PER_CONTRIB * SLOTS_PER_EVAL * C / ((kms["medium"] + kms["large"]) / 1e3) / 1e12
    s_bytes = 8 if args.precision != "fp32" else 4
    hbm_bytes = 48 * n + 2 * s_bytes * 12 * nside * nside
    roofline = {
        "bound": "valu",
        "achieved": achieved_valu,
        "peak": VALU_PEAK_TOPS,
        "unit": "Tlane-op/s",
        "frac": achieved_valu / VALU_PEAK_TOPS,
        "traffic": MEASURED_TRAFFIC.get((args.workload, nside, args.precision)) if args.n is None else None,
        "kernel": "k_disc_group (medium footprint kernel)",
        "kernel_ms": kms,
        "contributions": {"medium": C_med, "large": C_large},
        "footprint_kernels_frac": both / VALU_PEAK_TOPS,
        "hbm": {"achieved": hbm_bytes / res["sec"] / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": hbm_bytes / res["sec"] / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes": hbm_bytes},
    }
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
        "vs_fp64": vs_fp64,
        "roofline": roofline,
    }
    if host_in is not None:
        line["host_input"] = host_in
    del parts
    torch.cuda.empty_cache()
    if not args.no_batch:
        line["batch"] = batch_leg(pw, torch)
        log(f"batch: {line['batch']}")
        torch.cuda.empty_cache()
    if not args.no_adjoint:
        line["adjoint"] = adjoint_leg(pw, torch)
        log(f"adjoint: {line['adjoint']}")
        torch.cuda.empty_cache()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(pw, args, nside, n, C)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
