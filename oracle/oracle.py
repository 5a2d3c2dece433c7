"""ctypes binding of the CPU oracle (liboracle.so) and of the reference
known-answer library (_ref/libref_kat.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product path never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_kat.so")
REF_SRC = "/root/reference/src"

PARTICLE_DTYPE = np.dtype([(n, "<f4") for n in (
    "mass", "dens", "hsmooth", "posx", "posy", "posz",
    "velx", "vely", "velz", "sigmav", "eps", "phi")])
assert PARTICLE_DTYPE.itemsize == 48


class OrcParams(C.Structure):
    _fields_ = [
        ("Omega_M", C.c_double), ("h100", C.c_double), ("n_s", C.c_double),
        ("sigma_8", C.c_double), ("Lbox_in_Mpc", C.c_double), ("mass_to_Msun_in", C.c_double),
        ("particle_masses_in", C.c_double * 10), ("particle_numbers_in", C.c_long * 10),
        ("centerpos", C.c_double * 3), ("centervel", C.c_double * 3),
        ("align_vec", C.c_double * 3), ("obs_pos", C.c_double * 3),
        ("nside", C.c_int), ("test", C.c_int), ("mem_buffer", C.c_int),
        ("is_native", C.c_int), ("is_mask", C.c_int),
        ("datafile", C.c_char * 512), ("outputmap", C.c_char * 512), ("maskfile", C.c_char * 512),
        ("rho_crit_in_cgs", C.c_double), ("Delta_vir", C.c_double),
        ("mass_to_cgs", C.c_double), ("mass_to_Msun", C.c_double), ("length_to_Mpc", C.c_double),
        ("length_to_cgs", C.c_double), ("time_to_cgs", C.c_double),
        ("velocity_to_cgs", C.c_double), ("density_to_cgs", C.c_double),
        ("annihilation_flux_to_cgs", C.c_double),
        ("opos", C.c_double * 3), ("cpos", C.c_double * 3), ("otheta", C.c_double), ("ophi", C.c_double),
        ("particle_masses", C.c_double * 10),
        ("rotmatrix", C.c_double * 9),
        ("npix", C.c_int), ("dOmega", C.c_double),
        ("flux_factor", C.c_double),
    ]


# flux plug-ins (the reference's src/flux/ files), OrcConfig.flux_kind
FLUX_KINDS = ("ann_nosom", "ann_pwave", "ann_somv", "ann_somv2", "decay", "ann_adiabatic", "ann_core_nosom")


class OrcConfig(C.Structure):
    _fields_ = [("opos", C.c_double * 3), ("rotmatrix", C.c_double * 9),
                ("flux_factor", C.c_double), ("mass_cut", C.c_double), ("nside", C.c_int),
                ("flux_kind", C.c_int), ("decay_factor", C.c_double), ("vel_to_cgs", C.c_double),
                ("c_in_cgs", C.c_double), ("cpos", C.c_double * 3), ("footprint_pad", C.c_double),
                ("halos", C.c_void_p), ("nhalos", C.c_int64)]

    def set_halos(self, halos):
        """Attach halo-core rows ((n, 4) float32); the array is kept alive
        with the config."""
        self._halos = np.ascontiguousarray(halos, dtype=np.float32).reshape(-1, 4)
        self.halos = self._halos.ctypes.data
        self.nhalos = self._halos.shape[0]


class OrcStats(C.Structure):
    _fields_ = [("kept", C.c_int64), ("subpixel", C.c_int64), ("fallback", C.c_int64),
                ("extended", C.c_int64), ("contributions", C.c_int64), ("total_flux", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_ref = None


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-B" if force else "-s", "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    if os.path.isdir(REF_SRC) and (force or not os.path.exists(REF_PATH)):
        subprocess.run(["make", "-C", HERE, "-B" if force else "-s", "ref"], check=True,
                       stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
        L.orc_ring_above.argtypes = [C.c_int, C.c_double]
        L.orc_ring_above.restype = C.c_int
        L.orc_ring2z.argtypes = [C.c_int, C.c_int]
        L.orc_ring2z.restype = C.c_double
        L.orc_ang2pix_ring.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_ang2pix_ring.restype = C.c_int64
        L.orc_pix2vec_ring.argtypes = [C.c_int, C.c_int64, f64p]
        L.orc_query_disc.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, i64p, C.c_int64]
        L.orc_query_disc.restype = C.c_int64
        L.orc_params_from_ini.argtypes = [C.c_char_p, C.POINTER(OrcParams)]
        L.orc_params_from_ini.restype = C.c_int
        L.orc_config_from_params.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcConfig)]
        L.orc_getflux.argtypes = [C.c_double, C.c_void_p, C.c_double]
        L.orc_getflux.restype = C.c_double
        L.orc_core_correction.argtypes = [C.c_void_p, C.c_int64, C.c_float, C.c_float, C.c_float]
        L.orc_core_correction.restype = C.c_double
        L.orc_getflux_cfg.argtypes = [C.POINTER(OrcConfig), C.c_void_p, C.c_double]
        L.orc_getflux_cfg.restype = C.c_double
        L.orc_calc_angles.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_read_native.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.POINTER(C.c_void_p)]
        L.orc_read_native.restype = C.c_int64
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_genmap.argtypes = [C.c_void_p, C.c_int64, C.POINTER(OrcConfig), f64p, C.POINTER(OrcStats)]
        L.orc_genmap_omp.argtypes = [C.c_void_p, C.c_int64, C.POINTER(OrcConfig), f64p,
                                     C.POINTER(OrcStats), C.c_int]
        L.orc_finalize.argtypes = [f64p, C.c_int64, C.c_double]
        L.orc_finalize.restype = C.c_double
        L.orc_particle_footprint.argtypes = [C.c_void_p, C.POINTER(OrcConfig), i64p, f64p, C.c_int64,
                                             C.POINTER(C.c_double)]
        L.orc_particle_footprint.restype = C.c_int64
        _lib = L
    return _lib


def ref():
    """The reference KAT library, or None when it was not built here."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        R.ref_params.argtypes = [C.c_char_p, np.ctypeslib.ndpointer(np.float64, flags="C")]
        R.ref_getflux.argtypes = [C.c_char_p, np.ctypeslib.ndpointer(np.float32, flags="C"), C.c_double]
        R.ref_getflux.restype = C.c_double
        R.ref_read_native.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_char_p,
                                      np.ctypeslib.ndpointer(np.float32, flags="C"), C.c_longlong]
        R.ref_read_native.restype = C.c_longlong
        R.ref_read_tipsy.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_char_p,
                                     np.ctypeslib.ndpointer(np.float32, flags="C"), C.c_longlong]
        R.ref_read_tipsy.restype = C.c_longlong
        R.ref_sizeof_particle.restype = C.c_int
        _ref = R
    return _ref


_ref_flux = {}


def ref_flux(kind: str):
    """ref_flux(ini, rec12, d) of the reference's flux plug-in `kind`
    (oracle/_ref/libref_flux_<kind>.so), or None when it was not built."""
    if kind not in _ref_flux:
        path = os.path.join(os.path.dirname(REF_PATH), f"libref_flux_{kind}.so")
        R = None
        if os.path.exists(path):
            R = C.CDLL(path)
            R.ref_flux.argtypes = [C.c_char_p, np.ctypeslib.ndpointer(np.float32, flags="C"), C.c_double]
            R.ref_flux.restype = C.c_double
        _ref_flux[kind] = R
    return _ref_flux[kind]


# ---------------------------------------------------------------- wrappers

def params_from_ini(path: str) -> OrcParams:
    p = OrcParams()
    rc = lib().orc_params_from_ini(path.encode(), C.byref(p))
    if rc != 0:
        raise FileNotFoundError(path)
    return p


def config_from_params(p: OrcParams) -> OrcConfig:
    c = OrcConfig()
    lib().orc_config_from_params(C.byref(p), C.byref(c))
    return c


def make_config(opos, rotmatrix, flux_factor, nside, mass_cut=None) -> OrcConfig:
    c = OrcConfig()
    for i in range(3):
        c.opos[i] = float(opos[i])
    rm = np.asarray(rotmatrix, dtype=np.float64).reshape(9)
    for i in range(9):
        c.rotmatrix[i] = float(rm[i])
    c.flux_factor = float(flux_factor)
    c.nside = int(nside)
    c.mass_cut = float(np.float64(np.float32(3.4028234663852886e38)) * 1.1) if mass_cut is None else float(mass_cut)
    return c


def genmap(parts: np.ndarray, cfg: OrcConfig, nthreads: int = 1, finalize: bool = True):
    """Returns (map, stats).  map is divided by dOmega when finalize."""
    parts = np.ascontiguousarray(parts, dtype=PARTICLE_DTYPE)
    npix = 12 * cfg.nside * cfg.nside
    m = np.zeros(npix, dtype=np.float64)
    st = OrcStats()
    if nthreads == 1:
        lib().orc_genmap(parts.ctypes.data, len(parts), C.byref(cfg), m, C.byref(st))
    else:
        lib().orc_genmap_omp(parts.ctypes.data, len(parts), C.byref(cfg), m, C.byref(st), nthreads)
    if finalize:
        lib().orc_finalize(m, npix, 4 * np.pi / npix)
    return m, st.as_dict()


def query_disc(nside: int, theta: float, phi: float, radius: float) -> np.ndarray:
    cap = 1 << 12
    while True:
        out = np.zeros(cap, dtype=np.int64)
        n = lib().orc_query_disc(nside, theta, phi, radius, out, cap)
        if n <= cap:
            return out[:n].copy()
        cap = int(n)


def pix2vec(nside: int, pix) -> np.ndarray:
    pix = np.atleast_1d(np.asarray(pix, dtype=np.int64))
    out = np.zeros((len(pix), 3))
    v = np.zeros(3)
    for i, p in enumerate(pix):
        lib().orc_pix2vec_ring(nside, int(p), v)
        out[i] = v
    return out


def ang2pix(nside: int, theta: float, phi: float) -> int:
    return int(lib().orc_ang2pix_ring(nside, theta, phi))


def getflux(flux_factor: float, rec: np.ndarray, d: float) -> float:
    rec = np.ascontiguousarray(rec, dtype=PARTICLE_DTYPE)
    return float(lib().orc_getflux(flux_factor, rec.ctypes.data, d))


def getflux_cfg(cfg: OrcConfig, rec: np.ndarray, d: float) -> float:
    """getflux of the plug-in cfg.flux_kind (index into FLUX_KINDS)."""
    rec = np.ascontiguousarray(rec, dtype=PARTICLE_DTYPE)
    return float(lib().orc_getflux_cfg(C.byref(cfg), rec.ctypes.data, d))


def calc_angles(x, y, z, d, rotmatrix):
    rm = (C.c_double * 9)(*np.asarray(rotmatrix, dtype=np.float64).reshape(9))
    ct, ph = C.c_double(), C.c_double()
    lib().orc_calc_angles(x, y, z, d, rm, C.byref(ct), C.byref(ph))
    return ct.value, ph.value


def read_native(path: str, test: int = -1, maskfile: str | None = None) -> np.ndarray:
    ptr = C.c_void_p()
    n = lib().orc_read_native(path.encode(), test, (maskfile or "").encode(), C.byref(ptr))
    if n < 0:
        raise IOError(path)
    buf = (C.c_char * (48 * n)).from_address(ptr.value) if n else b""
    arr = np.frombuffer(bytes(buf), dtype=PARTICLE_DTYPE).copy()
    lib().orc_free(ptr)
    return arr


def particle_footprint(rec: np.ndarray, cfg: OrcConfig):
    """(pix, w/norm, flux) of one particle, as genmap.cpp deposits it."""
    rec = np.ascontiguousarray(rec, dtype=PARTICLE_DTYPE)
    cap = 1 << 14
    while True:
        pix = np.zeros(cap, dtype=np.int64)
        w = np.zeros(cap, dtype=np.float64)
        f = C.c_double()
        n = lib().orc_particle_footprint(rec.ctypes.data, C.byref(cfg), pix, w, cap, C.byref(f))
        if n <= cap:
            return pix[:n].copy(), w[:n].copy(), f.value
        cap = int(n)
