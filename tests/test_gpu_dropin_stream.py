"""The drop-in executable end to end on 1e7 particles (the NFW + 100 subhalo
set, seed 1) at nside 512: a native particle file read in mem_buffer chunks
of 2048 (the reference's own src/VL2_info.ini:57 setting) and of 1e7, the
chunks gathered into staged launches of 2^22 particles whose host-to-device
copies overlap the renders, the FITS map written.  Both maps against the
projector on the same particles resident in device memory (compensated,
2e-6 per pixel: launch composition and atomic order differ), wall time of
each run (process start, file read, render, FITS write) recorded; the
reference-linked executable too where it was built.
"""
import os
import subprocess
import time

import numpy as np
import pytest

from conftest import PKG, ROOT, assert_map, progress, record, rel_err_report

pytestmark = pytest.mark.gpu
EXE = os.path.join(PKG, "annmap_hip")
REF_EXE = os.path.join(ROOT, "oracle", "_ref", "annmap_ref")
NSIDE = 512


@pytest.mark.timeout(600)
def test_dropin_1e7_streamed(pw, gpu, tmp_path):
    import torch
    print("case drop-in end to end, 1e7 particles, mem_buffer 2048 and 1e7", flush=True)
    assert os.path.exists(EXE), "build annmap_hip first (make -C dm-annihilation-cxx_amd)"
    parts = pw.synth.make("nfw_sub100_1e7", "cuda")
    recs = pw.synth.to_records(parts)
    data = str(tmp_path / "parts.bin")
    with open(data, "wb") as fh:
        np.array([len(recs)], "<i4").tofile(fh)
        recs.tofile(fh)
    p = pw.params_for_observer(NSIDE, "compensated")
    with pw.Projector(p, parts.shape[0]) as pr:
        pr.render(parts)
        ref = pr.get_map()
        st = pr.stats()
    progress("drop-in: device-resident reference map rendered")
    runs = [(EXE, 2048), (EXE, 10_000_000)]
    if os.path.exists(REF_EXE):
        runs.append((REF_EXE, 2048))
    figures = {}
    for exe, mb in runs:
        name = f"{os.path.basename(exe)}_mb{mb}"
        out = str(tmp_path / f"{name}.fits")
        ini = pw.write_ini(str(tmp_path / f"{name}.ini"), nside=NSIDE, mem_buffer=mb, datafile=data, outputmap=out)
        t0 = time.perf_counter()
        r = subprocess.run([exe, "-info", ini], capture_output=True, text=True, timeout=300)
        sec = time.perf_counter() - t0
        print(name, r.stdout[-300:], r.stderr[-300:], flush=True)
        assert r.returncode == 0, r.stdout[-500:]
        assert f"{st['kept']} particles rendered" in r.stdout
        got, ns = pw.read_fits(out)
        assert ns == NSIDE
        rep = rel_err_report(got, ref)
        assert_map(rep, 2e-6, 1e-6)
        figures[f"{name}_s"] = sec
        figures[f"{name}_Mparticles_per_s"] = len(recs) / sec / 1e6
        progress(f"drop-in: {name} {sec:.2f} s")
    record("drop-in 1e7 nside512 end to end", **figures)
    torch.cuda.synchronize()
