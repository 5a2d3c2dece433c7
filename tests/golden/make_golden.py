#!/usr/bin/env python3
"""Writes the golden known-answer fixtures of tests/golden/ from the
reference's own compiled code (oracle/_ref/libref_kat.so, built by
`make -C oracle ref` from src/parameters.cpp, src/datatypes.cpp,
src/flux/flux_ann_nosom.cpp, src/datareader.cpp, src/tipsy_io.c).

Fixtures:
  params_kat.json   Parameters outputs (rotation matrix, observer, unit
                    chain, masses, map) and getflux values, full precision,
                    for the ini files below.
  flux_kat.json     getflux of every flux plug-in (oracle/_ref/libref_flux_<kind>.so,
                    one per src/flux/flux_<kind>.cpp) on records spanning the
                    velocity saturation points, for two ini files.
  core_kat.json, VL2_Halos.data
                    getflux of the halo-core plug-in on records in and around
                    the cores of the reference's halo file (a copy of
                    src/halocore/VL2_Halos.data, data of the reference).
  tipsy_kat.npz     the bytes of a small tipsy XDR file (gas, dark, star
                    records), its smooth side files and a mask, and the
                    records DataReader's tipsy path returns for several
                    (test, mem_buffer, mask) settings.
  reader_kat.npz    a small native particle file and mask, and the records
                    DataReader returns for several (test, mem_buffer) settings.
  VL2_info.ini, test_VL2_info.ini
                    the reference's own ini fixtures (src/VL2_info.ini,
                    src/test/VL2_info.ini), inputs of the KAT.
  synth_obs8.ini, synth_offaxis.ini
                    ini files of the synthetic workloads (written here).
Run in a container that has the reference tree; the fixtures are committed.
"""
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "dm-annihilation-cxx_amd"))

import oracle  # noqa: E402

REF = "/root/reference/src"
NAMES = (["rot%d" % i for i in range(9)] + ["opos%d" % i for i in range(3)] + ["cpos%d" % i for i in range(3)]
         + ["otheta", "ophi", "mass_to_cgs", "mass_to_Msun", "length_to_Mpc", "length_to_cgs", "time_to_cgs",
            "velocity_to_cgs", "density_to_cgs", "annihilation_flux_to_cgs", "rho_crit_in_cgs", "Delta_vir"]
         + ["particle_masses%d" % i for i in range(10)]
         + ["Nside", "Npix", "dOmega", "memParts", "testNum", "isNative", "isMask"])


def write_synth_inis():
    from pywrapper.capi import write_ini
    write_ini(os.path.join(HERE, "synth_obs8.ini"), nside=64)
    write_ini(os.path.join(HERE, "synth_offaxis.ini"), nside=256, obs_pos=(3.0, -4.5, 6.25),
              centerpos=(0.5, 0.25, -0.75), align_vec=(0.3, -0.2, 0.93), particle_masses=(4100.0, 8200.0, 0.0),
              mem_buffer=4096)


def write_flux_kat(rng):
    kinds = [k for k in oracle.FLUX_KINDS if k != "ann_core_nosom"]  # core_kat.json
    out = {"kinds": list(kinds), "cases": {}}
    for ini in ("VL2_info.ini", "synth_offaxis.ini"):
        path = os.path.join(HERE, ini)
        p = oracle.params_from_ini(path)
        n = 24
        recs = np.zeros(n, dtype=oracle.PARTICLE_DTYPE)
        recs["mass"] = rng.uniform(1e-9, 5e3, n)
        recs["dens"] = rng.uniform(1e-3, 1e6, n)
        recs["hsmooth"] = rng.uniform(0.01, 10, n)
        # velocity dispersions from 0.1 to 100 km/s (both saturation points)
        recs["sigmav"] = np.geomspace(0.1, 100.0, n) * 1e5 / p.velocity_to_cgs
        off = rng.normal(size=(n, 3)) * np.geomspace(1e-4, 2.0, n)[:, None]
        recs["posx"] = p.cpos[0] + off[:, 0]
        recs["posy"] = p.cpos[1] + off[:, 1]
        recs["posz"] = p.cpos[2] + off[:, 2]
        dists = rng.uniform(0.05, 300.0, n)
        case = {"records": recs.view(np.float32).reshape(n, 12).tolist(),
                "dists": [float(d).hex() for d in dists], "flux": {}}
        for k in kinds:
            R = oracle.ref_flux(k)
            if R is None:
                sys.exit(f"oracle/_ref/libref_flux_{k}.so is not built")
            vals = []
            for r, d in zip(recs, dists):
                f32 = np.ascontiguousarray(np.array([r], dtype=oracle.PARTICLE_DTYPE)).view(np.float32).copy()
                vals.append(float(R.ref_flux(path.encode(), f32, float(d))).hex())
            case["flux"][k] = vals
        out["cases"][ini] = case
    with open(os.path.join(HERE, "flux_kat.json"), "w") as fh:
        json.dump(out, fh, indent=1)


def tipsy_files(rng, nsph=7, ndark=1234, nstar=5):
    """Bytes of a tipsy XDR file (big-endian: header, gas, dark, star
    records) and of its three smooth side files (int32 nbodies + one float
    per body, host byte order)."""
    import struct
    nbodies = nsph + ndark + nstar
    head = struct.pack(">diiiiii", 0.375, nbodies, 3, nsph, ndark, nstar, 0)
    gas = rng.standard_normal((nsph, 12)).astype(">f4")
    dark = (rng.standard_normal((ndark, 9)) * np.arange(1, 10)).astype(">f4")
    dark[:, 0] = np.abs(dark[:, 0]) + 1.0
    star = rng.standard_normal((nstar, 11)).astype(">f4")
    part = head + gas.tobytes() + dark.tobytes() + star.tobytes()
    side = []
    for scale in (1e3, 0.1, 30.0):
        vals = (np.abs(rng.standard_normal(nbodies)) * scale).astype("<f4")
        side.append(np.array([nbodies], dtype="<i4").tobytes() + vals.tobytes())
    mask = rng.random(ndark) > 0.25
    mbytes = np.array([ndark], dtype="<i4").tobytes() + mask.astype(np.uint8).tobytes()
    return part, side, mbytes


def write_core_kat(rng):
    """getflux of the halo-core plug-in (flux_ann_core_nosom.cpp with
    halocore.cpp, which reads ./halocore/VL2_Halos.data from the working
    directory) on records placed inside and around halo cores."""
    shutil.copyfile(os.path.join(REF, "halocore", "VL2_Halos.data"), os.path.join(HERE, "VL2_Halos.data"))
    halos = np.loadtxt(os.path.join(HERE, "VL2_Halos.data"), dtype=np.float64)
    R = oracle.ref_flux("ann_core_nosom")
    if R is None:
        sys.exit("oracle/_ref/libref_flux_ann_core_nosom.so is not built")
    ini = os.path.join(HERE, "VL2_info.ini")
    n = 64
    pick = rng.choice(len(halos), n, replace=False)
    rc = halos[pick, 3] / (0.73 * 40000.0)
    off = rng.normal(size=(n, 3))
    off /= np.linalg.norm(off, axis=1, keepdims=True)
    off *= (rc * rng.uniform(0.0, 1.6, n))[:, None]
    recs = np.zeros(n, dtype=oracle.PARTICLE_DTYPE)
    recs["mass"] = rng.uniform(1e-9, 5e3, n)
    recs["dens"] = rng.uniform(1e-3, 1e6, n)
    recs["posx"] = halos[pick, 0] + off[:, 0]
    recs["posy"] = halos[pick, 1] + off[:, 1]
    recs["posz"] = halos[pick, 2] + off[:, 2]
    dists = rng.uniform(0.05, 300.0, n)
    cwd = os.getcwd()
    os.chdir(REF)
    try:
        vals = []
        for r, d in zip(recs, dists):
            f32 = np.ascontiguousarray(np.array([r], dtype=oracle.PARTICLE_DTYPE)).view(np.float32).copy()
            vals.append(float(R.ref_flux(ini.encode(), f32, float(d))).hex())
    finally:
        os.chdir(cwd)
    with open(os.path.join(HERE, "core_kat.json"), "w") as fh:
        json.dump({"ini": "VL2_info.ini", "halos": "VL2_Halos.data",
                   "records": recs.view(np.float32).reshape(n, 12).tolist(),
                   "dists": [float(d).hex() for d in dists], "flux": vals}, fh, indent=1)


def write_tipsy_kat(R, rng):
    part, side, mbytes = tipsy_files(rng)
    names = ["_t.tipsy", "_t.den32", "_t.hsm32", "_t.sigma32", "_t.mask"]
    paths = [os.path.join(HERE, nm) for nm in names]
    for pth, b in zip(paths, [part] + side + [mbytes]):
        with open(pth, "wb") as fh:
            fh.write(b)
    out = {"part": np.frombuffer(part, np.uint8), "dens": np.frombuffer(side[0], np.uint8),
           "hsm": np.frombuffer(side[1], np.uint8), "sigma": np.frombuffer(side[2], np.uint8),
           "mask": np.frombuffer(mbytes, np.uint8)}
    for name, test, membuf, use_mask in (("all_chunk100", -1, 100, False), ("test500_chunk64", 500, 64, False),
                                         ("mask_chunk333", -1, 333, True), ("mask_test900_chunk2000", 900, 2000, True)):
        buf = np.zeros(2000 * 12, dtype=np.float32)
        got = R.ref_read_tipsy(paths[0].encode(), paths[1].encode(), paths[2].encode(), paths[3].encode(), test,
                               membuf, (paths[4] if use_mask else "").encode(), buf, 2000)
        out[f"{name}__out"] = buf[:got * 12].reshape(got, 12).copy()
        out[f"{name}__cfg"] = np.array([test, membuf, int(use_mask)])
    for pth in paths:
        os.unlink(pth)
    np.savez_compressed(os.path.join(HERE, "tipsy_kat.npz"), **out)


def main():
    oracle.build()
    R = oracle.ref()
    if R is None:
        sys.exit("oracle/_ref/libref_kat.so is not built (needs the reference tree)")
    shutil.copyfile(os.path.join(REF, "VL2_info.ini"), os.path.join(HERE, "VL2_info.ini"))
    shutil.copyfile(os.path.join(REF, "test", "VL2_info.ini"), os.path.join(HERE, "test_VL2_info.ini"))
    write_synth_inis()
    inis = ["VL2_info.ini", "test_VL2_info.ini", "synth_obs8.ini", "synth_offaxis.ini"]
    rng = np.random.default_rng(12345)
    recs = np.zeros(8, dtype=oracle.PARTICLE_DTYPE)
    recs["mass"] = rng.uniform(1e-9, 5e3, 8)
    recs["dens"] = rng.uniform(1e-3, 1e6, 8)
    recs["hsmooth"] = rng.uniform(0.01, 10, 8)
    dists = rng.uniform(0.05, 300.0, 8)
    out = {"fields": NAMES, "cases": {}}
    for ini in inis:
        path = os.path.join(HERE, ini)
        v = np.zeros(44)
        R.ref_params(path.encode(), v)
        flux = []
        for r, d in zip(recs, dists):
            f32 = np.ascontiguousarray(np.array([r], dtype=oracle.PARTICLE_DTYPE)).view(np.float32).copy()
            flux.append(R.ref_getflux(path.encode(), f32, float(d)))
        out["cases"][ini] = {"values": [float(x).hex() for x in v],
                             "flux_records": recs.view(np.float32).reshape(8, 12).tolist(),
                             "flux_dists": [float(d).hex() for d in dists],
                             "flux": [float(f).hex() for f in flux]}
    with open(os.path.join(HERE, "params_kat.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    write_flux_kat(np.random.default_rng(777))
    write_tipsy_kat(R, np.random.default_rng(4242))
    write_core_kat(np.random.default_rng(99))

    # native reader KAT
    n = 1000
    parts = np.zeros(n, dtype=oracle.PARTICLE_DTYPE)
    for i, f in enumerate(oracle.PARTICLE_DTYPE.names):
        parts[f] = rng.standard_normal(n).astype(np.float32) * (i + 1)
    mask = (rng.random(n) > 0.3).astype(np.uint8)
    tmp = os.path.join(HERE, "_tmp_native.bin")
    tmpm = os.path.join(HERE, "_tmp_mask.bin")
    with open(tmp, "wb") as fh:
        np.array([n], dtype="<i4").tofile(fh)
        parts.tofile(fh)
    with open(tmpm, "wb") as fh:
        np.array([n], dtype="<i4").tofile(fh)
        mask.tofile(fh)
    cases = {}
    for name, test, membuf, use_mask in (("all_chunk97", -1, 97, False), ("test333_chunk50", 333, 50, False),
                                         ("mask_chunk128", -1, 128, True), ("mask_test700_chunk1000", 700, 1000, True)):
        buf = np.zeros(n * 12, dtype=np.float32)
        got = R.ref_read_native(tmp.encode(), test, membuf, (tmpm if use_mask else "").encode(), buf, n)
        cases[name] = (test, membuf, use_mask, buf[:got * 12].reshape(got, 12).copy())
    np.savez_compressed(os.path.join(HERE, "reader_kat.npz"), particles=parts.view(np.float32).reshape(n, 12),
                        mask=mask, **{f"{k}__out": v[3] for k, v in cases.items()},
                        **{f"{k}__cfg": np.array([v[0], v[1], int(v[2])]) for k, v in cases.items()})
    os.unlink(tmp)
    os.unlink(tmpm)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
