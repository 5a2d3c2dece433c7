"""ctypes binding of libdmann.so (include/dmann.h).

The library is built in-tree by `make -C dm-annihilation-cxx_amd`; importing
this module fails loudly when it is missing: there is no CPU fallback.
Device-resident inputs are torch tensors on a ROCm device; the context then
runs on torch's current stream.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import tempfile

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libdmann.so")

PREC_FP32, PREC_COMPENSATED, PREC_FP64, PREC_FIXED = 0, 1, 2, 3
PRECISIONS = {"fp32": PREC_FP32, "compensated": PREC_COMPENSATED, "fp64": PREC_FP64, "fixed": PREC_FIXED}

# 48-byte DMParticle record (src/datatypes.h:46-105)
PARTICLE_FIELDS = ("mass", "dens", "hsmooth", "posx", "posy", "posz",
                   "velx", "vely", "velz", "sigmav", "eps", "phi")
PARTICLE_DTYPE = np.dtype([(n, "<f4") for n in PARTICLE_FIELDS])

FLT_MAX = float(np.finfo(np.float32).max)


# flux plug-ins (the reference's src/flux/flux_<kind>.cpp), DmannParams.flux_kind
FLUX_KINDS = ("ann_nosom", "ann_pwave", "ann_somv", "ann_somv2", "decay", "ann_adiabatic", "ann_core_nosom")


class DmannParams(C.Structure):
    _fields_ = [("opos", C.c_double * 3), ("rotmatrix", C.c_double * 9),
                ("flux_factor", C.c_double), ("mass_cut", C.c_double),
                ("nside", C.c_int32), ("precision", C.c_int32),
                ("flux_kind", C.c_int32), ("reserved", C.c_int32),
                ("decay_factor", C.c_double), ("vel_to_cgs", C.c_double), ("c_in_cgs", C.c_double),
                ("cpos", C.c_double * 3), ("footprint_pad", C.c_double)]

    def with_cuda_cutoff(self) -> "DmannParams":
        """A copy whose footprints take every pixel within 2r + theta_pix/2,
        the reference CUDA render's cutoff (theta_pix = 2 pi / (4 nside))."""
        p = self.copy()
        p.footprint_pad = math.pi / (4 * self.nside)
        return p

    def with_flux(self, kind: str) -> "DmannParams":
        """A copy that uses the flux plug-in `kind` (one of FLUX_KINDS)."""
        p = self.copy()
        p.flux_kind = FLUX_KINDS.index(kind)
        return p

    def copy(self) -> "DmannParams":
        p = DmannParams()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(DmannParams))
        return p


class DmannStats(C.Structure):
    _fields_ = [("kept", C.c_int64), ("subpixel", C.c_int64), ("fallback", C.c_int64),
                ("extended", C.c_int64), ("contributions", C.c_int64), ("large", C.c_int64),
                ("large_contributions", C.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


EXPORTS = ("dmann_params_from_ini", "dmann_create", "dmann_render", "dmann_render_device",
           "dmann_set_stream", "dmann_synchronize", "dmann_get_map", "dmann_get_map_device",
           "dmann_reset", "dmann_get_stats", "dmann_npix", "dmann_destroy",
           "dmann_write_healpix_fits", "dmann_read_healpix_fits", "dmann_host_footprint", "dmann_host_ring_check",
           "dmann_host_sky_check", "dmann_host_nest_cells",
           "dmann_timing", "dmann_adjoint_device", "dmann_last_error", "dmann_set_tuning", "dmann_set_grouping",
           "dmann_batch_create", "dmann_batch_set_tuning", "dmann_batch_render_device", "dmann_batch_set_stream",
           "dmann_batch_synchronize", "dmann_batch_npix", "dmann_batch_map_device", "dmann_batch_get_map",
           "dmann_batch_get_stats", "dmann_batch_graph_replays", "dmann_batch_destroy",
           "dmann_batch_last_error", "dmann_read_particles", "dmann_set_halo_cores", "dmann_read_halo_file",
           "dmann_render_map_noalloc", "dmann_set_sky_gather",
           "dmann_set_subpixel_sort")

_lib = None


class DmannError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dmann status {code}: {msg}")
        self.code = code


def load_library():
    """Loads libdmann.so with its prototypes (no GPU call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `make -C {PKG_DIR}`")
    L = C.CDLL(LIB_PATH)
    vp, i64 = C.c_void_p, C.c_int64
    L.dmann_params_from_ini.argtypes = [C.c_char_p, C.c_int, C.POINTER(DmannParams),
                                        C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.dmann_create.argtypes = [C.POINTER(vp), C.POINTER(DmannParams), i64, C.c_int]
    L.dmann_render.argtypes = [vp, vp, i64]
    L.dmann_render_device.argtypes = [vp, vp, i64]
    L.dmann_set_stream.argtypes = [vp, vp]
    L.dmann_synchronize.argtypes = [vp]
    L.dmann_get_map.argtypes = [vp, vp]
    L.dmann_get_map_device.argtypes = [vp, vp]
    L.dmann_reset.argtypes = [vp]
    L.dmann_get_stats.argtypes = [vp, C.POINTER(DmannStats)]
    L.dmann_npix.argtypes = [vp]
    L.dmann_npix.restype = i64
    L.dmann_destroy.argtypes = [vp]
    L.dmann_destroy.restype = None
    L.dmann_set_halo_cores.argtypes = [vp, vp, i64]
    L.dmann_read_halo_file.argtypes = [C.c_char_p, vp, i64]
    L.dmann_read_halo_file.restype = i64
    L.dmann_read_particles.argtypes = [C.c_char_p, vp, i64]
    L.dmann_read_particles.restype = i64
    L.dmann_write_healpix_fits.argtypes = [C.c_char_p, vp, i64, C.c_int]
    L.dmann_read_healpix_fits.argtypes = [C.c_char_p, vp, i64, C.POINTER(C.c_int), C.POINTER(i64)]
    L.dmann_host_footprint.argtypes = [C.POINTER(DmannParams), vp, vp, vp, i64, C.POINTER(C.c_double),
                                       C.POINTER(C.c_int)]
    L.dmann_host_footprint.restype = i64
    L.dmann_host_ring_check.argtypes = [C.POINTER(DmannParams), vp]
    L.dmann_host_ring_check.restype = i64
    L.dmann_host_sky_check.argtypes = [C.POINTER(DmannParams), vp]
    L.dmann_host_sky_check.restype = i64
    L.dmann_host_nest_cells.argtypes = [C.c_int, vp, vp, i64, vp]
    L.dmann_host_nest_cells.restype = i64
    L.dmann_adjoint_device.argtypes = [vp, vp, i64, vp, vp]
    L.dmann_set_tuning.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.dmann_set_grouping.argtypes = [vp, C.c_int, C.c_int]
    L.dmann_set_sky_gather.argtypes = [vp, C.c_int]
    L.dmann_set_subpixel_sort.argtypes = [vp, C.c_int]
    L.dmann_timing.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.dmann_last_error.argtypes = [vp]
    L.dmann_last_error.restype = C.c_char_p
    L.dmann_batch_create.argtypes = [C.POINTER(vp), vp, C.c_int, C.c_int]
    L.dmann_batch_set_tuning.argtypes = [vp, C.c_int, C.c_int]
    L.dmann_batch_render_device.argtypes = [vp, vp, i64, C.c_int]
    L.dmann_batch_set_stream.argtypes = [vp, vp]
    L.dmann_batch_synchronize.argtypes = [vp]
    L.dmann_batch_npix.argtypes = [vp, C.c_int]
    L.dmann_batch_npix.restype = i64
    L.dmann_batch_map_device.argtypes = [vp, C.c_int, C.POINTER(vp)]
    L.dmann_batch_get_map.argtypes = [vp, C.c_int, vp]
    L.dmann_batch_get_stats.argtypes = [vp, C.c_int, C.POINTER(DmannStats)]
    L.dmann_batch_graph_replays.argtypes = [vp]
    L.dmann_batch_graph_replays.restype = C.c_longlong
    L.dmann_batch_destroy.argtypes = [vp]
    L.dmann_batch_destroy.restype = None
    L.dmann_batch_last_error.argtypes = [vp]
    L.dmann_batch_last_error.restype = C.c_char_p
    L.dmann_render_map_noalloc.argtypes = [C.POINTER(DmannParams), vp, i64, vp, vp, C.POINTER(DmannStats)]
    _lib = L
    return L


def _check(code, ctx=None):
    if code != 0:
        msg = load_library().dmann_last_error(ctx)
        raise DmannError(code, (msg or b"").decode(errors="replace"))


# ------------------------------------------------------------------ params

def params_from_ini(path: str, precision="compensated"):
    """(DmannParams, mem_buffer, test) from a reference ini file."""
    prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    p = DmannParams()
    mb, tn = C.c_int(), C.c_int()
    _check(load_library().dmann_params_from_ini(path.encode(), prec, C.byref(p), C.byref(mb), C.byref(tn)))
    return p, mb.value, tn.value


def write_ini(path: str, *, nside: int, obs_pos=(8.0, 0.0, 0.0), centerpos=(0.0, 0.0, 0.0),
              align_vec=(0.0, 0.0, 1.0), particle_masses=(4100.0,), Omega_M=0.238, h100=0.73,
              Lbox_in_Mpc=0.001, mass_to_Msun=1.0, mem_buffer=10000000, datafile="particles.bin",
              outputmap="skymap.fits", extra: dict | None = None) -> str:
    """Writes an ini for the synthetic workloads: code length 1 kpc, code
    mass 1 Msun, density in units of the critical density."""
    f3 = lambda v: " ".join(repr(float(x)) for x in v)
    lines = [f"Omega_M = {Omega_M}", f"h100 = {h100}", f"Lbox_in_Mpc = {Lbox_in_Mpc}",
             f"mass_to_Msun = {mass_to_Msun}",
             "particle_masses = " + " ".join(repr(float(m)) for m in particle_masses),
             f"centerpos = {f3(centerpos)}", f"align_vec = {f3(align_vec)}", f"obs_pos = {f3(obs_pos)}",
             f"nside = {int(nside)}", f"mem_buffer = {int(mem_buffer)}", f"datafile = {datafile}",
             f"outputmap = {outputmap}"]
    for k, v in (extra or {}).items():
        lines.append(f"{k} = {v}")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path


def params_for_observer(nside: int, precision="compensated", **ini_kw) -> DmannParams:
    """Map constants for an observer, through the library's own Parameters."""
    with tempfile.TemporaryDirectory() as d:
        path = write_ini(os.path.join(d, "obs.ini"), nside=nside, **ini_kw)
        p, _, _ = params_from_ini(path, precision)
    return p


# ------------------------------------------------------------------ context

def _as_host_records(parts) -> np.ndarray:
    a = np.ascontiguousarray(parts)
    if a.dtype == PARTICLE_DTYPE:
        return a
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != 12:
        raise ValueError("particles must be a DMParticle record array or an (n, 12) float32 array")
    return a


class Projector:
    """One map on one device (a dmann_ctx)."""

    def __init__(self, params: DmannParams, max_parts: int, device: int = 0):
        L = load_library()
        self._L = L
        self.params = params.copy()
        h = C.c_void_p()
        _check(L.dmann_create(C.byref(h), C.byref(self.params), int(max_parts), int(device)))
        self._h = h
        self.device = device
        self.npix = int(L.dmann_npix(h))
        self.max_parts = int(max_parts)

    # -- rendering
    def render(self, parts) -> None:
        """Adds particles (host numpy records, or a torch (n, 12) float32
        tensor on the context's device) to the map."""
        try:
            import torch
            is_tensor = isinstance(parts, torch.Tensor)
        except ImportError:  # pragma: no cover
            is_tensor = False
        if is_tensor:
            t = parts
            if t.dtype != torch.float32 or t.dim() != 2 or t.shape[1] != 12 or not t.is_contiguous():
                raise ValueError("device particles must be a contiguous (n, 12) float32 tensor")
            if not t.is_cuda:
                self.render(t.numpy())
                return
            self._bind_stream()
            _check(self._L.dmann_render_device(self._h, C.c_void_p(t.data_ptr()), int(t.shape[0])), self._h)
            return
        a = _as_host_records(parts)
        n = a.shape[0]
        self._bind_stream()
        _check(self._L.dmann_render(self._h, C.c_void_p(a.ctypes.data), int(n)), self._h)

    def _bind_stream(self):
        """Every call runs on torch's current stream of the device; the
        library orders a newly bound stream after the previous one."""
        s = _torch_stream(self.device)
        _check(self._L.dmann_set_stream(self._h, C.c_void_p(s)), self._h)

    def synchronize(self) -> None:
        self._bind_stream()
        _check(self._L.dmann_synchronize(self._h), self._h)

    def get_map(self) -> np.ndarray:
        out = np.empty(self.npix, dtype=np.float64)
        self._bind_stream()
        _check(self._L.dmann_get_map(self._h, C.c_void_p(out.ctypes.data)), self._h)
        return out

    def get_map_device(self, out):
        """Writes the finished map into a float64 CUDA tensor of npix."""
        import torch
        if out.dtype != torch.float64 or out.numel() != self.npix or not out.is_cuda:
            raise ValueError("out must be a float64 device tensor of npix elements")
        self._bind_stream()
        _check(self._L.dmann_get_map_device(self._h, C.c_void_p(out.data_ptr())), self._h)
        return out

    def adjoint(self, parts, grad_map, out=None):
        """dL/d(mass, dens, hsmooth, posx, posy, posz) per particle (FP64,
        (n, 6) device tensor) for dL/dmap `grad_map` (float64 device tensor of
        npix).  `parts` is a (n, 12) float32 device tensor."""
        import torch
        n = parts.shape[0]
        if out is None:
            out = torch.empty(n, 6, dtype=torch.float64, device=parts.device)
        g = grad_map.contiguous()
        if g.dtype != torch.float64 or g.numel() != self.npix or not g.is_cuda:
            raise ValueError("grad_map must be a float64 device tensor of npix elements")
        if not (parts.is_cuda and parts.dtype == torch.float32 and parts.is_contiguous()):
            raise ValueError("parts must be a contiguous float32 device tensor")
        self._bind_stream()
        _check(self._L.dmann_adjoint_device(self._h, C.c_void_p(parts.data_ptr()), int(n),
                                            C.c_void_p(g.data_ptr()), C.c_void_p(out.data_ptr())), self._h)
        return out

    def set_tuning(self, medium_blocks_per_cu: int = 8, large_blocks_per_cu: int = 4, large_pix: int = 65536):
        _check(self._L.dmann_set_tuning(self._h, int(medium_blocks_per_cu), int(large_blocks_per_cu),
                                        int(large_pix)), self._h)

    def set_sky_gather(self, blocks_per_cu: int = 12):
        """Sky-covering class as a per-ring gather with that many
        single-wavefront workgroups per CU (0: one workgroup per particle
        with per-pixel atomics)."""
        _check(self._L.dmann_set_sky_gather(self._h, int(blocks_per_cu)), self._h)

    def set_subpixel_sort(self, on: bool = True):
        """Sub-pixel deposits reduced per wavefront by pixel (True) or one
        atomic per particle (False)."""
        _check(self._L.dmann_set_subpixel_sort(self._h, 1 if on else 0), self._h)

    def set_grouping(self, k: int = 64, nside_c: int | None = None):
        """K neighbours per wavefront (0: one particle per wavefront), sorted
        by coarse NESTED cell at nside_c (default min(nside, 32)), then size
        class, then NESTED cell at the map's nside."""
        ns = self.params.nside
        _check(self._L.dmann_set_grouping(self._h, int(k), int(nside_c or min(ns, 32))), self._h)

    def timing(self, enable: bool):
        """Returns ({'prepass','large','sort','medium'}: summed ms, calls)
        of the render calls since the last timing() call, then
        enables/disables event recording for the next ones.  'medium' is
        the medium footprint kernel alone, 'sort' the coarse-cell sort that
        precedes it."""
        ms = (C.c_double * 4)()
        calls = C.c_int()
        self._bind_stream()
        _check(self._L.dmann_timing(self._h, int(bool(enable)), ms, C.byref(calls)), self._h)
        return {"prepass": ms[0], "large": ms[1], "sort": ms[2], "medium": ms[3]}, calls.value

    def set_halo_cores(self, halos) -> None:
        """Halo-core rows (n, 4) float32 (xc, yc, zc, radius) for the
        ann_core_nosom flux plug-in."""
        h = np.ascontiguousarray(halos, dtype=np.float32).reshape(-1, 4)
        _check(self._L.dmann_set_halo_cores(self._h, C.c_void_p(h.ctypes.data), int(h.shape[0])), self._h)

    def reset(self) -> None:
        self._bind_stream()
        _check(self._L.dmann_reset(self._h), self._h)

    def stats(self) -> dict:
        st = DmannStats()
        self._bind_stream()
        _check(self._L.dmann_get_stats(self._h, C.byref(st)), self._h)
        return st.as_dict()

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.dmann_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def _torch_stream(device: int) -> int:
    """Handle of torch's current stream on `device` (0, the null stream,
    when torch has no GPU)."""
    try:
        import torch
    except ImportError:  # pragma: no cover
        return 0
    if not torch.cuda.is_available():
        return 0
    return int(torch.cuda.current_stream(device).cuda_stream)


def read_halo_file(path: str) -> np.ndarray:
    """(n, 4) float32 rows (xc, yc, zc, radius) of a halo-core file."""
    L = load_library()
    n = int(L.dmann_read_halo_file(path.encode(), None, 0))
    if n < 0:
        raise DmannError(1, f"cannot read {path}")
    out = np.zeros((n, 4), dtype=np.float32)
    L.dmann_read_halo_file(path.encode(), C.c_void_p(out.ctypes.data), n)
    return out


def read_particles(ini: str) -> np.ndarray:
    """The particles an ini file names (native or tipsy + smooth files, with
    its test truncation and mask), as DMParticle records."""
    L = load_library()
    n = int(L.dmann_read_particles(ini.encode(), None, 0))
    if n < 0:
        raise DmannError(-n, f"cannot read the particles of {ini}")
    out = np.zeros(n, dtype=PARTICLE_DTYPE)
    m = int(L.dmann_read_particles(ini.encode(), C.c_void_p(out.ctypes.data), n))
    if m != n:
        raise DmannError(-m if m < 0 else 4, f"particle files of {ini} changed while read")
    return out


def host_footprint(params: DmannParams, rec):
    """(kind, pixels, FP32 weights, flux) of one particle, computed on the
    host by the kernels' own per-particle and per-ring code."""
    L = load_library()
    r = np.ascontiguousarray(np.asarray(rec).reshape(-1)[:1]).view(np.float32)
    if r.size != 12:
        r = np.ascontiguousarray(np.asarray(rec, dtype=np.float32).reshape(12))
    cap = 1 << 12
    while True:
        pix = np.zeros(cap, dtype=np.int32)
        w = np.zeros(cap, dtype=np.float32)
        fl, kind = C.c_double(), C.c_int()
        n = L.dmann_host_footprint(C.byref(params), C.c_void_p(r.ctypes.data), C.c_void_p(pix.ctypes.data),
                                   C.c_void_p(w.ctypes.data), cap, C.byref(fl), C.byref(kind))
        if n < 0:
            raise DmannError(int(n), "bad particle")
        if n <= cap:
            m = 1 if kind.value in (1, 2) else int(n)
            return kind.value, pix[:m].copy(), w[:m].copy(), fl.value
        cap = int(n)


def host_ring_check(params: DmannParams, rec) -> int:
    """For a medium-class particle: the number of its rings whose ranges in
    the grouped kernel's cheap form (ring_lite) differ from the query_disc
    construction (ring_setup); -2 for other classes."""
    L = load_library()
    r = np.ascontiguousarray(np.asarray(rec).reshape(-1)[:1]).view(np.float32)
    if r.size != 12:
        r = np.ascontiguousarray(np.asarray(rec, dtype=np.float32).reshape(12))
    return int(L.dmann_host_ring_check(C.byref(params), C.c_void_p(r.ctypes.data)))


def host_sky_check(params: DmannParams, rec) -> int:
    """For a sky-covering particle: the number of differences between the
    pixels and weights the per-ring gather visits and those of the
    per-particle ring walk (0 expected); -2 for other classes."""
    L = load_library()
    r = np.ascontiguousarray(np.asarray(rec).reshape(-1)[:1]).view(np.float32)
    if r.size != 12:
        r = np.ascontiguousarray(np.asarray(rec, dtype=np.float32).reshape(12))
    return int(L.dmann_host_sky_check(C.byref(params), C.c_void_p(r.ctypes.data)))


def host_nest_cells(nside: int, theta, phi) -> np.ndarray:
    """HEALPix NESTED index at nside (a power of two) of the directions
    (theta, phi), as the grouping's sort computes it (host build)."""
    L = load_library()
    t = np.ascontiguousarray(theta, dtype=np.float64)
    f = np.ascontiguousarray(phi, dtype=np.float64)
    out = np.empty(t.size, dtype=np.uint32)
    r = L.dmann_host_nest_cells(int(nside), C.c_void_p(t.ctypes.data), C.c_void_p(f.ctypes.data), t.size,
                                C.c_void_p(out.ctypes.data))
    if r != t.size:
        raise ValueError(f"dmann_host_nest_cells: bad arguments (nside {nside})")
    return out


# ------------------------------------------------------------------ batch

class Batch:
    """Many maps (observers, mixed nside) of one particle set, rendered
    together and replayed from a hipGraph (dmann_batch_*)."""

    def __init__(self, params_list, device: int = 0):
        L = load_library()
        self._L = L
        arr = (DmannParams * len(params_list))(*[p.copy() for p in params_list])
        self._params = arr
        h = C.c_void_p()
        code = L.dmann_batch_create(C.byref(h), C.cast(arr, C.c_void_p), len(params_list), int(device))
        if code != 0:
            raise DmannError(code, (L.dmann_batch_last_error(None) or b"").decode())
        self._h = h
        self.device = device
        self.nmaps = len(params_list)
        self.npix = [int(L.dmann_batch_npix(h, i)) for i in range(self.nmaps)]

    def _chk(self, code):
        if code != 0:
            raise DmannError(code, (self._L.dmann_batch_last_error(self._h) or b"").decode())

    def set_tuning(self, chunk: int, blocks_per_cu: int):
        self._chk(self._L.dmann_batch_set_tuning(self._h, int(chunk), int(blocks_per_cu)))

    def render(self, parts, use_graph: bool = True):
        """parts: (n, 12) float32 CUDA tensor, kept alive by the caller
        while a captured graph may replay it.  The render is ordered after
        the work on torch's current stream, and work enqueued there later
        sees the finished maps."""
        self._chk(self._L.dmann_batch_set_stream(self._h, C.c_void_p(_torch_stream(self.device))))
        self._chk(self._L.dmann_batch_render_device(self._h, C.c_void_p(parts.data_ptr()), int(parts.shape[0]),
                                                    int(bool(use_graph))))

    def synchronize(self):
        self._chk(self._L.dmann_batch_synchronize(self._h))

    def get_map(self, i: int) -> np.ndarray:
        out = np.empty(self.npix[i], dtype=np.float64)
        self._chk(self._L.dmann_batch_get_map(self._h, int(i), C.c_void_p(out.ctypes.data)))
        return out

    def stats(self, i: int) -> dict:
        st = DmannStats()
        self._chk(self._L.dmann_batch_get_stats(self._h, int(i), C.byref(st)))
        return st.as_dict()

    def map_device(self, i: int):
        """Finished map i (float64, npix, divided by dOmega) as a torch
        tensor viewing the batch's device buffer (no copy); valid until the
        batch is closed.  Read it on the stream the render was issued from
        (or after synchronize())."""
        import torch
        ptr = C.c_void_p()
        self._chk(self._L.dmann_batch_map_device(self._h, int(i), C.byref(ptr)))

        class _View:
            __cuda_array_interface__ = {"shape": (self.npix[i],), "typestr": "<f8",
                                        "data": (int(ptr.value), False), "version": 2}

        return torch.as_tensor(_View(), device=f"cuda:{self.device}")

    def replays(self) -> int:
        return int(self._L.dmann_batch_graph_replays(self._h))

    def close(self):
        if getattr(self, "_h", None):
            self._L.dmann_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ no-alloc

def render_map_noalloc(params: DmannParams, parts, out, stats: bool = False):
    """Allocation-free render (dmann_render_map_noalloc): the finished FP64
    map of `parts` ((n, 12) float32 CUDA tensor) into `out` (float64 CUDA
    tensor of npix), on torch's current stream; no device memory is
    allocated.  stats=True waits and returns the statistics dict."""
    import torch
    npix = 12 * params.nside * params.nside
    if not (out.is_cuda and out.dtype == torch.float64 and out.numel() == npix and out.is_contiguous()):
        raise ValueError("out must be a contiguous float64 CUDA tensor of npix elements")
    if not (parts.is_cuda and parts.dtype == torch.float32 and parts.dim() == 2 and parts.shape[1] == 12
            and parts.is_contiguous()):
        raise ValueError("parts must be a contiguous (n, 12) float32 CUDA tensor")
    L = load_library()
    st = DmannStats() if stats else None
    code = L.dmann_render_map_noalloc(C.byref(params), C.c_void_p(parts.data_ptr()), int(parts.shape[0]),
                                      C.c_void_p(out.data_ptr()), C.c_void_p(_torch_stream(out.device.index or 0)),
                                      C.byref(st) if st is not None else None)
    _check(code)
    return st.as_dict() if st is not None else None


# ------------------------------------------------------------------ FITS

def write_fits(path: str, m: np.ndarray, nside: int) -> None:
    m = np.ascontiguousarray(m, dtype=np.float64)
    _check(load_library().dmann_write_healpix_fits(path.encode(), C.c_void_p(m.ctypes.data), m.size, nside))


def read_fits(path: str):
    L = load_library()
    ns, n = C.c_int(), C.c_int64()
    _check(L.dmann_read_healpix_fits(path.encode(), None, 0, C.byref(ns), C.byref(n)))
    out = np.empty(n.value, dtype=np.float64)
    _check(L.dmann_read_healpix_fits(path.encode(), C.c_void_p(out.ctypes.data), n.value, C.byref(ns),
                                     C.byref(n)))
    return out, ns.value
