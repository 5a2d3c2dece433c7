"""Particle-file reader and FITS output.

Reader: the oracle's native-format reader and the library's reader (native
and tipsy + smooth side files) against records that the reference's own
DataReader (src/datareader.cpp) returned for the same bytes
(tests/golden/reader_kat.npz, tipsy_kat.npz): chunked buffers, `test`
truncation, mask.
FITS: the library's HEALPix FITS writer against the FITS standard (2880-byte
blocks, header keywords, big-endian float64 column) and a round trip.
"""
import os
import struct

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def kat():
    return np.load(os.path.join(GOLD, "reader_kat.npz"))


def _write_native(path, parts, maskpath=None, mask=None):
    with open(path, "wb") as fh:
        np.array([len(parts)], dtype="<i4").tofile(fh)
        np.ascontiguousarray(parts, dtype=np.float32).tofile(fh)
    if maskpath is not None:
        with open(maskpath, "wb") as fh:
            np.array([len(mask)], dtype="<i4").tofile(fh)
            np.asarray(mask, dtype=np.uint8).tofile(fh)


@pytest.mark.parametrize("case", ["all_chunk97", "test333_chunk50", "mask_chunk128", "mask_test700_chunk1000"])
def test_oracle_reader_matches_reference_datareader(orc, kat, tmp_path, case):
    test, membuf, use_mask = (int(x) for x in kat[f"{case}__cfg"])
    f = str(tmp_path / "p.bin")
    m = str(tmp_path / "m.bin")
    _write_native(f, kat["particles"], m, kat["mask"])
    got = orc.read_native(f, test, m if use_mask else None)
    want = kat[f"{case}__out"]
    assert np.array_equal(got.view(np.float32).reshape(-1, 12), want)


def test_reader_errors(orc, tmp_path):
    with pytest.raises(IOError):
        orc.read_native(str(tmp_path / "missing.bin"))
    f = str(tmp_path / "p.bin")
    m = str(tmp_path / "m.bin")
    parts = np.zeros((10, 12), np.float32)
    _write_native(f, parts, m, np.ones(9, np.uint8))  # count mismatch
    with pytest.raises(IOError):
        orc.read_native(f, -1, m)


def test_fits_layout_and_round_trip(pw, tmp_path):
    nside = 16
    npix = 12 * nside * nside
    rng = np.random.default_rng(0)
    m = rng.standard_normal(npix)
    path = str(tmp_path / "map.fits")
    pw.write_fits(path, m, nside)
    raw = open(path, "rb").read()
    assert len(raw) % 2880 == 0
    cards = [raw[i:i + 80].decode() for i in range(0, 2880, 80)]
    assert cards[0].startswith("SIMPLE  =                    T")
    ext = [raw[2880 + i:2880 + i + 80].decode() for i in range(0, 2880, 80)]
    kv = {c[:8].strip(): c[10:].split("/")[0].strip() for c in ext if "=" in c[:10]}
    assert kv["XTENSION"] == "'BINTABLE'"
    assert kv["NAXIS1"] == str(8 * 1024) and kv["NAXIS2"] == str(npix // 1024)
    assert kv["TFORM1"] == "'1024D   '"
    assert kv["ORDERING"] == "'RING    '" and kv["PIXTYPE"] == "'HEALPIX '"
    assert kv["NSIDE"] == str(nside) and kv["LASTPIX"] == str(npix - 1)
    data = np.frombuffer(raw[5760:5760 + 8 * npix], dtype=">f8")
    assert np.array_equal(data, m)
    back, ns = pw.read_fits(path)
    assert ns == nside and np.array_equal(back, m)
    # small maps use one value per row
    pw.write_fits(path, m[:48], 2)
    back, ns = pw.read_fits(path)
    assert ns == 2 and np.array_equal(back, m[:48])
    assert struct.unpack(">d", open(path, "rb").read()[5760:5768])[0] == m[0]


@pytest.mark.parametrize("case", ["all_chunk97", "test333_chunk50", "mask_chunk128", "mask_test700_chunk1000"])
def test_product_reader_native_matches_reference_datareader(pw, kat, tmp_path, case):
    """The library's DataReader (C ABI dmann_read_particles) on the native
    format against the records the reference's DataReader returned."""
    test, membuf, use_mask = (int(x) for x in kat[f"{case}__cfg"])
    f = str(tmp_path / "p.bin")
    m = str(tmp_path / "m.bin")
    _write_native(f, kat["particles"], m, kat["mask"])
    extra = {"datafile": f, "test": test}
    if use_mask:
        extra["mask"] = m
    ini = pw.write_ini(str(tmp_path / "r.ini"), nside=16, mem_buffer=membuf, extra=extra)
    got = pw.read_particles(ini)
    assert np.array_equal(got.view(np.float32).reshape(-1, 12), kat[f"{case}__out"])


@pytest.fixture(scope="module")
def tipsy_kat():
    return np.load(os.path.join(GOLD, "tipsy_kat.npz"))


@pytest.mark.parametrize("case", ["all_chunk100", "test500_chunk64", "mask_chunk333", "mask_test900_chunk2000"])
def test_product_reader_tipsy_matches_reference_datareader(pw, tipsy_kat, tmp_path, case):
    """The tipsy path (XDR particle file with gas and star records around the
    dark ones, smooth side files, mask) against the reference's DataReader
    on the same bytes."""
    test, membuf, use_mask = (int(x) for x in tipsy_kat[f"{case}__cfg"])
    paths = {}
    for key in ("part", "dens", "hsm", "sigma", "mask"):
        paths[key] = str(tmp_path / f"t.{key}")
        tipsy_kat[key].tofile(paths[key])
    extra = {"partfile": paths["part"], "densfile": paths["dens"], "hsmfile": paths["hsm"],
             "vdspfile": paths["sigma"], "test": test}
    if use_mask:
        extra["mask"] = paths["mask"]
    ini = pw.write_ini(str(tmp_path / "r.ini"), nside=16, mem_buffer=membuf, extra=extra)
    got = pw.read_particles(ini)
    assert np.array_equal(got.view(np.float32).reshape(-1, 12), tipsy_kat[f"{case}__out"])


def test_product_reader_errors(pw, tipsy_kat, tmp_path):
    paths = {}
    for key in ("part", "dens", "hsm", "sigma"):
        paths[key] = str(tmp_path / f"t.{key}")
        tipsy_kat[key].tofile(paths[key])
    # a side file with the wrong body count
    bad = tipsy_kat["dens"].copy()
    bad[:4] = np.frombuffer(np.array([7], dtype="<i4").tobytes(), np.uint8)
    bad.tofile(paths["dens"])
    extra = {"partfile": paths["part"], "densfile": paths["dens"], "hsmfile": paths["hsm"],
             "vdspfile": paths["sigma"]}
    ini = pw.write_ini(str(tmp_path / "r.ini"), nside=16, extra=extra)
    with pytest.raises(pw.DmannError):
        pw.read_particles(ini)
    ini = pw.write_ini(str(tmp_path / "n.ini"), nside=16, extra={"datafile": str(tmp_path / "missing.bin")})
    with pytest.raises(pw.DmannError):
        pw.read_particles(ini)


CFITSIO = "/opt/conda/lib/libcfitsio.so"


@pytest.mark.skipif(not os.path.exists(CFITSIO), reason="cfitsio is not installed here")
@pytest.mark.parametrize("nside", [2, 64])
def test_fits_read_by_cfitsio(pw, tmp_path, nside):
    """The FITS map read back by an independent reader, cfitsio (the
    library the reference's HEALPix fitshandle writes with,
    src/mapgenerator.cpp:30-62): second HDU a binary table with the
    HEALPix keywords, one float64 column holding the map in RING order."""
    import ctypes as C
    L = C.CDLL(CFITSIO)
    npix = 12 * nside * nside
    m = np.random.default_rng(1).standard_normal(npix) * 1e-3
    path = str(tmp_path / "map.fits")
    pw.write_fits(path, m, nside)
    f = C.c_void_p()
    st = C.c_int(0)
    assert L.ffopen(C.byref(f), path.encode(), 0, C.byref(st)) == 0, st.value
    nhdu = C.c_int(0)
    L.ffthdu(f, C.byref(nhdu), C.byref(st))
    assert nhdu.value == 2
    ext = C.c_int(0)
    assert L.ffmahd(f, 2, C.byref(ext), C.byref(st)) == 0 and ext.value == 2  # BINARY_TBL

    def key_long(name):
        v = C.c_long(0)
        assert L.ffgkyj(f, name.encode(), C.byref(v), None, C.byref(st)) == 0, (name, st.value)
        return v.value

    def key_str(name):
        v = C.create_string_buffer(72)
        assert L.ffgkys(f, name.encode(), v, None, C.byref(st)) == 0, (name, st.value)
        return v.value.decode().strip()

    assert key_long("NSIDE") == nside and key_long("LASTPIX") == npix - 1
    assert key_str("ORDERING") == "RING" and key_str("PIXTYPE") == "HEALPIX"
    ncol = C.c_int(0)
    L.ffgncl(f, C.byref(ncol), C.byref(st))
    assert ncol.value == 1
    typecode, repeat, width = C.c_int(0), C.c_long(0), C.c_long(0)
    assert L.ffgtcl(f, 1, C.byref(typecode), C.byref(repeat), C.byref(width), C.byref(st)) == 0
    assert typecode.value == 82  # TDOUBLE
    nrows = C.c_long(0)
    L.ffgnrw(f, C.byref(nrows), C.byref(st))
    assert nrows.value * repeat.value == npix
    out = np.zeros(npix, dtype=np.float64)
    anynul = C.c_int(0)
    L.ffgcvd.argtypes = [C.c_void_p, C.c_int, C.c_longlong, C.c_longlong, C.c_longlong, C.c_double,
                         C.c_void_p, C.c_void_p, C.c_void_p]
    assert L.ffgcvd(f, 1, 1, 1, npix, 0.0, out.ctypes.data, C.byref(anynul), C.byref(st)) == 0, st.value
    L.ffclos(f, C.byref(st))
    assert st.value == 0
    assert np.array_equal(out, m)
