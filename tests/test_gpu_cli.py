"""End-to-end drop-in: <exe> -info <ini> reads a native particle file (with a
mask) in mem_buffer chunks, renders on the GPU and writes the HEALPix FITS
map; the map must match the oracle run on the same file (compensated form:
1e-5 per pixel, 1e-6 total), the FITS must round-trip, and a missing ini
must exit non-zero with the reference's message.

Two executables: annmap_hip (the package's own driver and Parameters /
DataReader mirrors) and oracle/_ref/annmap_ref -- the reference's own
src/main.cpp, parameters.cpp, datareader.cpp linked with the package's
MapGenerator::start (host/genmap_hip.cpp) compiled against the reference's
headers, i.e. the link-time drop-in itself (built where the reference tree
exists; skipped otherwise)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG, ROOT, assert_map, rel_err_report

pytestmark = pytest.mark.gpu
EXE = os.path.join(PKG, "annmap_hip")
REF_EXE = os.path.join(ROOT, "oracle", "_ref", "annmap_ref")


@pytest.mark.parametrize("exe", [EXE, REF_EXE], ids=["annmap_hip", "reference_main"])
def test_annmap_hip_end_to_end(pw, orc, gpu, tmp_path, exe):
    print(f"case cli {os.path.basename(exe)}", flush=True)
    if exe == REF_EXE and not os.path.exists(exe):
        pytest.skip("the reference-linked drop-in is built only where the reference tree exists")
    assert os.path.exists(exe), "build annmap_hip first (make -C dm-annihilation-cxx_amd)"
    EXE_ = exe
    recs = pw.synth.to_records(pw.synth.make("nfw_1e5", n=6000))
    data = str(tmp_path / "parts.bin")
    mask = str(tmp_path / "mask.bin")
    with open(data, "wb") as fh:
        np.array([len(recs)], "<i4").tofile(fh)
        recs.tofile(fh)
    m = (np.arange(len(recs)) % 7 != 3).astype(np.uint8)
    with open(mask, "wb") as fh:
        np.array([len(recs)], "<i4").tofile(fh)
        m.tofile(fh)
    out = str(tmp_path / "sky.fits")
    ini = pw.write_ini(str(tmp_path / "run.ini"), nside=64, mem_buffer=1000, datafile=data, outputmap=out,
                       extra={"mask": mask})
    r = subprocess.run([EXE_, "-info", ini], capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:], r.stderr[-400:])
    assert r.returncode == 0
    assert "Total Conts" in r.stdout
    got, nside = pw.read_fits(out)
    assert nside == 64
    p = orc.params_from_ini(ini)
    cfg = orc.config_from_params(p)
    parts = orc.read_native(data, -1, mask)
    ref, st = orc.genmap(parts, cfg, nthreads=4)
    rep = rel_err_report(got, ref)
    print(rep)
    assert_map(rep, 1e-5, 1e-6)
    # a second run overwrites the existing file
    r = subprocess.run([EXE_, "-info", ini], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "File exists" in r.stdout


def test_annmap_hip_missing_ini(tmp_path):
    r = subprocess.run([EXE, "-info", str(tmp_path / "nope.ini")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Info file incorrect" in r.stdout
