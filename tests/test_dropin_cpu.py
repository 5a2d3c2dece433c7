"""The drop-in boundary and the FITS output, checked without a GPU.

* Link-time drop-in: host/genmap_hip.cpp (MapGenerator::start) compiled
  against the reference's own headers (src/parameters.h, datareader.h,
  mapgenerator.h: the preprocessor must resolve them there, not to the
  package's mirrors) and linked with the reference's src/main.cpp,
  parameters.cpp, datatypes.cpp, datareader.cpp, tipsy_io.c and
  libdmann.so, by oracle/Makefile into oracle/_ref/annmap_ref; its usage
  path runs without a GPU.  Skipped where the reference tree is absent.
* FITS: the library's HEALPix writer read back by an independent reader,
  cfitsio (the library HEALPix C++ writes through in
  src/mapgenerator.cpp:30-62), where a cfitsio shared library is present:
  the binary-table extension's keywords, its column format and every value.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

REF = "/root/reference/src"
ORACLE = os.path.join(ROOT, "oracle")
CFITSIO = ["/opt/conda/lib/libcfitsio.so", "libcfitsio.so", "libcfitsio.so.8", "libcfitsio.so.10"]


def test_dropin_builds_against_reference_headers():
    if not os.path.isdir(REF):
        pytest.skip("reference tree absent")
    subprocess.run(["make", "-C", ORACLE, "-s", "_ref/annmap_ref"], check=True, capture_output=True)
    exe = os.path.join(ORACLE, "_ref", "annmap_ref")
    assert os.path.exists(exe)
    src = os.path.join(ORACLE, "_ref", "dropin", "genmap_hip.cpp")
    pp = subprocess.run(["g++", "-std=c++17", "-E", f"-I{REF}", f"-I{ROOT}/include", "-I/usr/include/tirpc", src],
                        check=True, capture_output=True, text=True).stdout
    used = {line.split('"')[1] for line in pp.splitlines() if line.startswith("# ") and '"' in line}
    for h in ("parameters.h", "datareader.h", "mapgenerator.h", "datatypes.h"):
        assert os.path.join(REF, h) in used, h
        assert not any(u.endswith("/host/" + h) for u in used), h
    r = subprocess.run([exe, "-help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.startswith("Usage:")


def _cfitsio():
    for name in CFITSIO:
        try:
            return C.CDLL(name)
        except OSError:
            continue
    return None


def test_fits_read_by_cfitsio(pw, tmp_path):
    L = _cfitsio()
    if L is None:
        pytest.skip("no cfitsio shared library")
    for nside in (16, 4):  # npix divisible by 1024 (vector column) and not (scalar column)
        npix = 12 * nside * nside
        m = np.random.default_rng(nside).lognormal(size=npix) * 1e-3
        m[5] = 0.0
        path = str(tmp_path / f"map{nside}.fits")
        pw.write_fits(path, m, nside)
        fp = C.c_void_p()
        st = C.c_int(0)
        assert L.ffopen(C.byref(fp), path.encode(), 0, C.byref(st)) == 0, st.value
        ext = C.c_int()
        assert L.ffmahd(fp, 2, C.byref(ext), C.byref(st)) == 0, st.value
        assert ext.value == 2  # BINARY_TBL

        def key_s(k):
            v = C.create_string_buffer(80)
            assert L.ffgkys(fp, k.encode(), v, None, C.byref(st)) == 0, (k, st.value)
            return v.value.decode().strip()

        def key_j(k):
            v = C.c_long()
            assert L.ffgkyj(fp, k.encode(), C.byref(v), None, C.byref(st)) == 0, (k, st.value)
            return v.value

        assert key_s("PIXTYPE") == "HEALPIX"
        assert key_s("ORDERING") == "RING"
        assert key_s("INDXSCHM") == "IMPLICIT"
        assert key_j("NSIDE") == nside
        assert key_j("FIRSTPIX") == 0 and key_j("LASTPIX") == npix - 1
        rep = 1024 if npix % 1024 == 0 else 1
        assert key_s("TFORM1") == f"{rep}D"
        nrow = C.c_long()
        assert L.ffgnrw(fp, C.byref(nrow), C.byref(st)) == 0
        assert nrow.value * rep == npix
        out = np.zeros(npix, np.float64)
        anynul = C.c_int()
        assert L.ffgcvd(fp, 1, C.c_longlong(1), C.c_longlong(1), C.c_longlong(npix), C.c_double(0.0),
                        out.ctypes.data_as(C.c_void_p), C.byref(anynul), C.byref(st)) == 0, st.value
        assert L.ffclos(fp, C.byref(st)) == 0
        assert np.array_equal(out, m)
