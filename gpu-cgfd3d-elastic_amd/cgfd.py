"""ctypes binding of libcgfd_el_iso.so, the gfx950 el_iso RK4 path.

The binding is thin on purpose: every array crosses the C ABI declared in
include/cgfd_el_iso.h as a plain pointer in the reference's host layout
(nx*ny*nz per component, nx = ni+6, i fastest).  Loading fails loudly when
the shared library has not been built; there is no CPU fallback.
"""
import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcgfd_el_iso.so")

STORE_F32, STORE_F16, STORE_BF16, STORE_F64 = 0, 1, 2, 3
# CGFD_WAVELET_* (include/cgfd_setup.h)
WAVELET_RICKER, WAVELET_GAUSSIAN, WAVELET_RICKER_DERIV, WAVELET_GAUSSIAN_DERIV, WAVELET_KLAUDER = 0, 1, 2, 3, 4
DT_F32, DT_F64 = 0, 1
LEVEL_PRE, LEVEL_TMPA, LEVEL_TMPB, LEVEL_END, LEVEL_RHS = 0, 1, 2, 3, 4
FLAG_PLAIN = 1
FLAG_PML_POINT = 2
FLAG_FULL_METRIC = 4
SPLIT_SHELL = 0
SPLIT_PLANES = 1
SPLIT_CORE = 2

# every symbol the header declares (checked by the CPU test suite)
EXPORTS = [
    "cgfd_create", "cgfd_destroy", "cgfd_footprint_bytes", "cgfd_last_error",
    "cgfd_upload_metric", "cgfd_upload_medium", "cgfd_upload_free_surface",
    "cgfd_upload_pml", "cgfd_set_storage_scale", "cgfd_upload_source",
    "cgfd_upload_receivers", "cgfd_set_level", "cgfd_reset_shot", "cgfd_run",
    "cgfd_el_iso_onestage", "cgfd_read_seismo", "cgfd_fetch_level",
    "cgfd_fetch_pml_aux", "cgfd_set_timing", "cgfd_last_run_timing",
    "cgfd_set_tile", "cgfd_get_tile", "cgfd_get_split", "cgfd_set_split", "cgfd_get_const_metric", "cgfd_set_graph", "cgfd_set_shot",
    "cgfd_run_async", "cgfd_poll", "cgfd_wait", "cgfd_reserve_receivers", "cgfd_set_seismo_ring",
    "cgfd_grid_cartesian", "cgfd_grid_topography", "cgfd_metric_calc", "cgfd_dt_estimate",
    "cgfd_media_layered", "cgfd_media_homogeneous", "cgfd_rho_to_slow", "cgfd_pml_len_dh",
    "cgfd_pml_profile", "cgfd_free_surface_matrices", "cgfd_wavelet", "cgfd_angle2moment",
    "cgfd_source_point", "cgfd_receiver_index", "cgfd_locate_points", "cgfd_media_layer_file", "cgfd_media_grid_file",
    "cgfd_media_bin_file",
    "cgfd_storage_scales",
    "cgfd_server_create", "cgfd_server_destroy", "cgfd_server_add_model", "cgfd_server_submit",
    "cgfd_server_wait", "cgfd_server_result", "cgfd_server_get_stats", "cgfd_fetch_subvolume",
    "cgfd_fetch_subvolume_noalloc",
    "cgfd_sac_write", "cgfd_station_sac", "cgfd_snap_open", "cgfd_snap_append", "cgfd_snap_close",
    "cgfd_set_peak_motion", "cgfd_read_peak_motion", "cgfd_peak_motion_write", "cgfd_hbm_stream",
    "cgfd_hbm_stream_dword", "cgfd_station_file", "cgfd_source_file",
    "cgfd_slice_open", "cgfd_slice_append", "cgfd_slice_close", "cgfd_line_sac", "cgfd_line_points",
]


class TopoDesc(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int), ("dh", C.c_double), ("stretch", C.c_double),
                ("nhill", C.c_int), ("amp_min", C.c_double), ("amp_max", C.c_double), ("sig_min", C.c_double),
                ("sig_max", C.c_double), ("rough_rms", C.c_double), ("min_wavelength", C.c_double),
                ("seed", C.c_ulonglong)]


class LayerDesc(C.Structure):
    _fields_ = [("nlayer", C.c_int), ("z_interface", C.c_void_p), ("vp", C.c_void_p), ("vpvs", C.c_void_p),
                ("rho", C.c_void_p), ("pert_std", C.c_double), ("pert_clip", C.c_double),
                ("corr_len", C.c_double), ("seed", C.c_ulonglong)]


class Desc(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int),
                ("storage", C.c_int), ("nshots", C.c_int), ("flags", C.c_int)]


class SourceDesc(C.Structure):
    _fields_ = [("nsrc", C.c_int), ("force_on", C.c_int), ("moment_on", C.c_int),
                ("max_nt", C.c_int), ("max_ext", C.c_int), ("dtype", C.c_int),
                ("it_begin", C.c_void_p), ("it_end", C.c_void_p), ("ext_num", C.c_void_p),
                ("ext_indx", C.c_void_p), ("ext_coef", C.c_void_p),
                ("F", C.c_void_p), ("M", C.c_void_p)]


class CgfdError(RuntimeError):
    pass


class SnapDesc(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int), ("out_vel", C.c_int), ("out_stress", C.c_int),
                ("out_strain", C.c_int), ("first_global", C.c_int * 3), ("first_local", C.c_int * 3),
                ("stride", C.c_int * 3), ("topo", C.c_int * 3)]


class ServerDesc(C.Structure):
    _fields_ = [("hbm_budget", C.c_longlong), ("max_batch", C.c_int), ("chunk_steps", C.c_int),
                ("group_share", C.c_double)]


class ModelDesc(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int), ("storage", C.c_int), ("dtype", C.c_int),
                ("metric", C.c_void_p), ("lam", C.c_void_p), ("mu", C.c_void_p), ("slw", C.c_void_p),
                ("free_top", C.c_int), ("matVx2Vz", C.c_void_p), ("matVy2Vz", C.c_void_p),
                ("pml_nlay", C.c_int * 6), ("pml_A", C.c_void_p * 6), ("pml_B", C.c_void_p * 6),
                ("pml_D", C.c_void_p * 6), ("scale_exp2", C.c_void_p), ("dt", C.c_double),
                ("max_nrecv", C.c_int), ("max_nt", C.c_int)]


class Request(C.Structure):
    _fields_ = [("model", C.c_int), ("nt_total", C.c_int), ("src", C.c_void_p), ("nrecv", C.c_int),
                ("recv_indx8", C.c_void_p), ("recv_dtype", C.c_int), ("recv_frac3", C.c_void_p)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("admitted", C.c_longlong), ("completed", C.c_longlong),
                ("rejected", C.c_longlong), ("groups", C.c_int), ("active_shots", C.c_int),
                ("queued", C.c_int), ("hbm_in_use", C.c_longlong), ("hbm_budget", C.c_longlong)]


def _source_desc(src):
    """cgfd_source_desc for a source dict, and the arrays it points into."""
    t = np.float64 if np.asarray(src["ext_coef"]).dtype == np.float64 else np.float32
    keep = dict(
        itb=_arr(src["it_begin"], np.int32), ite=_arr(src["it_end"], np.int32),
        en=_arr(src["ext_num"], np.int32), ei=_arr(src["ext_indx"], np.int64),
        ec=_arr(src["ext_coef"], t),
        F=_arr(src["F"], t) if src.get("F") is not None else None,
        M=_arr(src["M"], t) if src.get("M") is not None else None)
    d = SourceDesc(src["nsrc"], src["force_on"], src["moment_on"], src["max_nt"], src["max_ext"],
                   DT_F64 if t == np.float64 else DT_F32,
                   _ptr(keep["itb"]), _ptr(keep["ite"]), _ptr(keep["en"]), _ptr(keep["ei"]),
                   _ptr(keep["ec"]), _ptr(keep["F"]), _ptr(keep["M"]))
    return d, keep


_lib = None


def load(path=LIB_PATH):
    """Load the HIP library (raises if it is missing: no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise CgfdError(f"{path} not built; run __graft_entry__.build()")
    lib = C.CDLL(path)
    vp, i, ll = C.c_void_p, C.c_int, C.c_longlong
    sig = {
        "cgfd_create": ([C.POINTER(Desc), C.POINTER(vp)], i),
        "cgfd_destroy": ([vp], None),
        "cgfd_footprint_bytes": ([C.POINTER(Desc), i, i], ll),
        "cgfd_last_error": ([], C.c_char_p),
        "cgfd_upload_metric": ([vp, i, vp], i),
        "cgfd_upload_medium": ([vp, i, vp, vp, vp], i),
        "cgfd_upload_free_surface": ([vp, i, i, vp, vp], i),
        "cgfd_upload_pml": ([vp, i, i, i, vp, vp, vp], i),
        "cgfd_set_storage_scale": ([vp, vp], i),
        "cgfd_upload_source": ([vp, i, C.POINTER(SourceDesc)], i),
        "cgfd_upload_receivers": ([vp, i, i, vp, i, vp, i], i),
        "cgfd_set_level": ([vp, i, i, i, vp], i),
        "cgfd_reset_shot": ([vp, i], i),
        "cgfd_run": ([vp, i, i, C.c_double], i),
        "cgfd_el_iso_onestage": ([vp, i, i, i, i, i], i),
        "cgfd_read_seismo": ([vp, i, i, vp], i),
        "cgfd_fetch_level": ([vp, i, i, i, vp], i),
        "cgfd_fetch_pml_aux": ([vp, i, i, i, i, vp], i),
        "cgfd_set_timing": ([vp, i], i),
        "cgfd_set_graph": ([vp, i], i),
        "cgfd_set_shot": ([vp, i, i, i], i),
        "cgfd_run_async": ([vp, i, i, C.c_double], i),
        "cgfd_poll": ([vp], i),
        "cgfd_reserve_receivers": ([vp, i, i], i),
        "cgfd_set_seismo_ring": ([vp, i], i),
        "cgfd_wait": ([vp], i),
        "cgfd_set_tile": ([vp, i, i], i),
        "cgfd_get_tile": ([vp, C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(i)], i),
        "cgfd_get_split": ([vp, C.POINTER(i)], i),
        "cgfd_set_split": ([vp, i], i),
        "cgfd_get_const_metric": ([vp, C.POINTER(i)], i),
        "cgfd_last_run_timing": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)], i),
    }
    d = C.c_double
    sig.update({
        "cgfd_grid_cartesian": ([i, i, i, d, d, d, d, d, d, vp, vp, vp], i),
        "cgfd_grid_topography": ([C.POINTER(TopoDesc), vp, vp, vp], i),
        "cgfd_metric_calc": ([i, i, i, vp, vp, vp, vp], i),
        "cgfd_dt_estimate": ([i, i, i, vp, vp, vp, vp, vp, vp, d], d),
        "cgfd_media_layered": ([i, i, i, vp, vp, vp, C.POINTER(LayerDesc), vp, vp, vp], i),
        "cgfd_media_homogeneous": ([i, i, i, d, d, d, vp, vp, vp], i),
        "cgfd_rho_to_slow": ([ll, vp, vp], i),
        "cgfd_pml_len_dh": ([i, i, i, vp, vp, vp, i, i, C.POINTER(d), C.POINTER(d)], i),
        "cgfd_pml_profile": ([i, i, d, d, d, d, d, vp, vp, vp], i),
        "cgfd_free_surface_matrices": ([i, i, i, vp, vp, vp, vp, vp], i),
        "cgfd_wavelet": ([i, d, d, d], d),
        "cgfd_angle2moment": ([d, d, d, vp], i),
        "cgfd_source_point": ([i, i, i, d, d, d, i, d, d, d, i, d, d, d, vp, vp, vp, vp, vp, vp, vp], i),
        "cgfd_receiver_index": ([i, i, i, d, d, d, i, vp, vp], i),
        "cgfd_locate_points": ([i, i, i, vp, vp, vp, i, vp, vp, vp, vp, vp], i),
        "cgfd_media_layer_file": ([i, i, i, vp, vp, vp, C.c_char_p, C.c_char_p, vp, vp, vp], i),
        "cgfd_media_grid_file": ([i, i, i, vp, vp, vp, C.c_char_p, C.c_char_p, vp, vp, vp], i),
        "cgfd_media_bin_file": ([i, i, i, vp, vp, vp, vp, vp, vp, vp, C.c_char_p, C.c_char_p, C.c_char_p,
                                 vp, vp, vp], i),
        "cgfd_storage_scales": ([i, i, i, i, vp, vp, vp, vp, C.POINTER(SourceDesc), d, vp], i),
        "cgfd_fetch_subvolume": ([vp, i, i, C.c_uint, i, i, i, i, i, i, i, i, i, i, vp], i),
        "cgfd_fetch_subvolume_noalloc": ([vp, i, i, C.c_uint, i, i, i, i, i, i, i, i, i, i, vp], i),
        "cgfd_sac_write": ([C.c_char_p, vp, i, C.c_float, C.c_float, vp, vp, C.c_char_p, C.c_char_p], i),
        "cgfd_station_sac": ([C.c_char_p, C.c_char_p, vp, i, C.c_float, C.c_float, C.c_float, C.c_float, i,
                              C.c_float, C.c_float], i),
        "cgfd_snap_open": ([C.c_char_p, C.POINTER(SnapDesc), C.POINTER(vp)], i),
        "cgfd_snap_append": ([vp, C.c_float, vp, vp, vp], i),
        "cgfd_snap_close": ([vp], i),
        "cgfd_set_peak_motion": ([vp, i], i),
        "cgfd_read_peak_motion": ([vp, i, i, vp], i),
        "cgfd_peak_motion_write": ([C.c_char_p, i, i, vp, vp, vp, vp], i),
        "cgfd_server_create": ([C.POINTER(ServerDesc), C.POINTER(vp)], i),
        "cgfd_server_destroy": ([vp], None),
        "cgfd_server_add_model": ([vp, C.POINTER(ModelDesc), C.POINTER(i)], i),
        "cgfd_server_submit": ([vp, C.POINTER(Request), C.POINTER(ll)], i),
        "cgfd_server_wait": ([vp, ll, d], i),
        "cgfd_server_result": ([vp, ll, i, vp, C.POINTER(d), C.POINTER(d)], i),
        "cgfd_server_get_stats": ([vp, C.POINTER(ServerStats)], i),
        "cgfd_hbm_stream": ([ll, i, C.POINTER(d), C.POINTER(d)], i),
        "cgfd_hbm_stream_dword": ([ll, i, C.POINTER(d)], i),
        "cgfd_station_file": ([C.c_char_p, i, i, i, vp, vp, vp, i, vp, C.POINTER(i)], i),
        "cgfd_source_file": ([C.c_char_p, i, i, i, vp, vp, vp, vp, d, d, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp],
                             i),
        "cgfd_slice_open": ([C.c_char_p, i, i, i, i, vp, C.POINTER(vp)], i),
        "cgfd_slice_append": ([vp, C.c_float, vp], i),
        "cgfd_slice_close": ([vp], i),
        "cgfd_line_sac": ([C.c_char_p, C.c_char_p, i, vp, i, C.c_float, C.c_float, C.c_float, C.c_float], i),
        "cgfd_line_points": ([i, i, i, vp, vp, vp, vp, vp, i, vp, vp, vp, vp], i),
    })
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _check(rc):
    if rc != 0:
        msg = _lib.cgfd_last_error().decode(errors="replace")
        raise CgfdError(f"cgfd error {rc}: {msg}")


def _arr(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Session:
    """One device session: a shared model and nshots wavefields."""

    def __init__(self, ni, nj, nk, storage=STORE_F32, nshots=1, flags=0):
        lib = load()
        self.ni, self.nj, self.nk = ni, nj, nk
        self.nx, self.ny, self.nz = ni + 6, nj + 6, nk + 6
        self.storage, self.nshots = storage, nshots
        self._desc = Desc(ni, nj, nk, storage, nshots, flags)
        h = C.c_void_p()
        _check(lib.cgfd_create(C.byref(self._desc), C.byref(h)))
        self.h = h
        self._nrecv = [0] * nshots
        self._nt = [0] * nshots

    def close(self):
        if self.h:
            _lib.cgfd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def mdtype(self):
        return np.float64 if self.storage == STORE_F64 else np.float32

    def _dt(self, a):
        return (DT_F64, np.float64) if a.dtype == np.float64 else (DT_F32, np.float32)

    def upload_metric(self, metric):
        dt, t = self._dt(metric)
        m = _arr(metric, t)
        _check(_lib.cgfd_upload_metric(self.h, dt, _ptr(m)))

    def upload_medium(self, lam, mu, slw):
        dt, t = self._dt(lam)
        a, b, c = _arr(lam, t), _arr(mu, t), _arr(slw, t)
        _check(_lib.cgfd_upload_medium(self.h, dt, _ptr(a), _ptr(b), _ptr(c)))

    def upload_free_surface(self, free_top, mvx=None, mvy=None):
        if not free_top:
            _check(_lib.cgfd_upload_free_surface(self.h, 0, DT_F32, None, None))
            return
        dt, t = self._dt(mvx)
        a, b = _arr(mvx, t), _arr(mvy, t)
        _check(_lib.cgfd_upload_free_surface(self.h, 1, dt, _ptr(a), _ptr(b)))

    def upload_pml(self, face, nlay, A=None, B=None, D=None):
        if nlay <= 0:
            _check(_lib.cgfd_upload_pml(self.h, face, 0, DT_F32, None, None, None))
            return
        dt, t = self._dt(A)
        a, b, d = _arr(A, t), _arr(B, t), _arr(D, t)
        _check(_lib.cgfd_upload_pml(self.h, face, nlay, dt, _ptr(a), _ptr(b), _ptr(d)))

    def set_storage_scale(self, exp2):
        e = _arr(exp2, np.int32)
        _check(_lib.cgfd_set_storage_scale(self.h, _ptr(e)))

    def upload_source(self, shot, src):
        """src: dict with nsrc, force_on, moment_on, max_nt, max_ext, it_begin,
        it_end, ext_num, ext_indx, ext_coef, F (3,nsrc,max_nt,4) or None, M (6,...) or None."""
        if src is None or src["nsrc"] == 0:
            _check(_lib.cgfd_upload_source(self.h, shot, None))
            return
        d, keep = _source_desc(src)
        _check(_lib.cgfd_upload_source(self.h, shot, C.byref(d)))

    def upload_receivers(self, shot, indx8, frac3, nt_total):
        ix = _arr(indx8, np.int64).reshape(-1, 8)
        dt, t = self._dt(np.asarray(frac3))
        fr = _arr(frac3, t).reshape(-1, 3)
        _check(_lib.cgfd_upload_receivers(self.h, shot, ix.shape[0], _ptr(ix), dt, _ptr(fr), nt_total))
        self._nrecv[shot] = ix.shape[0]
        self._nt[shot] = nt_total

    def set_level(self, shot, level, w9):
        dt, t = self._dt(w9)
        w = _arr(w9, t)
        _check(_lib.cgfd_set_level(self.h, shot, level, dt, _ptr(w)))

    def reset_shot(self, shot):
        _check(_lib.cgfd_reset_shot(self.h, shot))

    def run(self, it0, nsteps, dt):
        _check(_lib.cgfd_run(self.h, it0, nsteps, C.c_double(dt)))

    def onestage(self, shot, in_level, ipair, istage, it):
        _check(_lib.cgfd_el_iso_onestage(self.h, shot, in_level, ipair, istage, it))

    def read_seismo(self, shot, dtype=np.float64):
        out = np.zeros((self._nrecv[shot], 9, self._nt[shot]), dtype=dtype)
        _check(_lib.cgfd_read_seismo(self.h, shot, DT_F64 if dtype == np.float64 else DT_F32, _ptr(out)))
        return out

    def fetch_level(self, shot, level, dtype=np.float64):
        out = np.zeros((9, self.nz, self.ny, self.nx), dtype=dtype)
        _check(_lib.cgfd_fetch_level(self.h, shot, level, DT_F64 if dtype == np.float64 else DT_F32, _ptr(out)))
        return out

    def fetch_pml_aux(self, shot, face, level, fsize, dtype=np.float64):
        out = np.zeros((9, fsize), dtype=dtype)
        _check(_lib.cgfd_fetch_pml_aux(self.h, shot, face, level, DT_F64 if dtype == np.float64 else DT_F32,
                                       _ptr(out)))
        return out

    def set_tile(self, shape, zchunk=48):
        _check(_lib.cgfd_set_tile(self.h, shape, zchunk))

    def set_split(self, split):
        """Pin the work split (SPLIT_*; -1: automatic)."""
        _check(_lib.cgfd_set_split(self.h, split))

    def get_split(self):
        """SPLIT_PLANES, SPLIT_CORE or SPLIT_SHELL: how stages divide between the kernels."""
        v = C.c_int()
        _check(_lib.cgfd_get_split(self.h, C.byref(v)))
        return v.value

    def const_metric(self):
        """True when the sweep runs with the metric as nine constants (Cartesian grid)."""
        v = C.c_int()
        _check(_lib.cgfd_get_const_metric(self.h, C.byref(v)))
        return bool(v.value)

    def get_tile(self):
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(_lib.cgfd_get_tile(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return dict(shape=a.value, tx=b.value, ty=c.value, zchunk=d.value)

    def fetch_subvolume(self, shot, level, comps, start, count, stride=(1, 1, 1), dtype=np.float32):
        """Components `comps` (indices 0..8) at start + n*stride (i, j, k order,
        physical indices) as [len(comps)][nk][nj][ni]."""
        mask = 0
        for c in comps:
            mask |= 1 << c
        (i1, j1, k1), (ni, nj, nk), (di, dj, dk) = start, count, stride
        out = np.zeros((len(comps), nk, nj, ni), dtype=dtype)
        _check(_lib.cgfd_fetch_subvolume(self.h, shot, level, mask, i1, j1, k1, ni, nj, nk, di, dj, dk,
                                         DT_F64 if dtype == np.float64 else DT_F32, _ptr(out)))
        return out

    def fetch_subvolume_noalloc(self, shot, level, comps, start, count, stride=(1, 1, 1)):
        """fetch_subvolume through the library's fixed device scratch (no device
        allocation); output in the session's model type."""
        mask = 0
        for c in comps:
            mask |= 1 << c
        (i1, j1, k1), (ni, nj, nk), (di, dj, dk) = start, count, stride
        dtype = np.float64 if self.storage == STORE_F64 else np.float32
        out = np.zeros((len(comps), nk, nj, ni), dtype=dtype)
        _check(_lib.cgfd_fetch_subvolume_noalloc(self.h, shot, level, mask, i1, j1, k1, ni, nj, nk, di, dj, dk,
                                                 DT_F64 if dtype == np.float64 else DT_F32, _ptr(out)))
        return out

    def set_peak_motion(self, on):
        _check(_lib.cgfd_set_peak_motion(self.h, 1 if on else 0))

    def read_peak_motion(self, shot, dtype=np.float64):
        """[15][ny][nx] (PGV PGVh PGVx PGVy PGVz PGA ... PGDz), ghosts zero."""
        out = np.zeros((15, self.ny, self.nx), dtype=dtype)
        _check(_lib.cgfd_read_peak_motion(self.h, shot, DT_F64 if dtype == np.float64 else DT_F32, _ptr(out)))
        return out

    def set_timing(self, on):
        _check(_lib.cgfd_set_timing(self.h, 1 if on else 0))

    def reserve_receivers(self, max_nrecv, max_nt):
        _check(_lib.cgfd_reserve_receivers(self.h, max_nrecv, max_nt))

    def set_seismo_ring(self, ring_steps):
        """Keep ring_steps samples per trace on the device, streaming chunks
        of ring_steps/2 to a pinned host record (0: whole record on device)."""
        _check(_lib.cgfd_set_seismo_ring(self.h, int(ring_steps)))

    def set_graph(self, on):
        _check(_lib.cgfd_set_graph(self.h, 1 if on else 0))

    def set_shot(self, shot, active, it_start=0):
        _check(_lib.cgfd_set_shot(self.h, shot, 1 if active else 0, it_start))

    def run_async(self, it0, nsteps, dt):
        _check(_lib.cgfd_run_async(self.h, it0, nsteps, C.c_double(dt)))

    def poll(self):
        r = _lib.cgfd_poll(self.h)
        if r < 0:
            _check(r)
        return r == 0

    def wait(self):
        _check(_lib.cgfd_wait(self.h))

    def last_run_timing(self):
        a, b, n = C.c_double(), C.c_double(), C.c_int()
        _check(_lib.cgfd_last_run_timing(self.h, C.byref(a), C.byref(b), C.byref(n)))
        return a.value, b.value, n.value


# ---------------------------------------------------------------- host setup
def _z(shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype)


def grid_cartesian(ni, nj, nk, dx, x0, dy, y0, dz, z0):
    load()
    sh = (nk + 6, nj + 6, ni + 6)
    x, y, z = _z(sh), _z(sh), _z(sh)
    _check(_lib.cgfd_grid_cartesian(ni, nj, nk, dx, x0, dy, y0, dz, z0, _ptr(x), _ptr(y), _ptr(z)))
    return x, y, z


def grid_topography(ni, nj, nk, dh, stretch, nhill, amp, sig, rough_rms, min_wavelength, seed):
    load()
    sh = (nk + 6, nj + 6, ni + 6)
    x, y, z = _z(sh), _z(sh), _z(sh)
    d = TopoDesc(ni, nj, nk, dh, stretch, nhill, amp[0], amp[1], sig[0], sig[1], rough_rms, min_wavelength, seed)
    _check(_lib.cgfd_grid_topography(C.byref(d), _ptr(x), _ptr(y), _ptr(z)))
    return x, y, z


def metric_calc(ni, nj, nk, x, y, z):
    load()
    m = _z((10,) + x.shape)
    _check(_lib.cgfd_metric_calc(ni, nj, nk, _ptr(_arr(x, np.float64)), _ptr(_arr(y, np.float64)),
                                 _ptr(_arr(z, np.float64)), _ptr(m)))
    return m


def dt_estimate(ni, nj, nk, x, y, z, lam, mu, rho, cfl=1.3):
    load()
    a = [_arr(v, np.float64) for v in (x, y, z, lam, mu, rho)]
    r = _lib.cgfd_dt_estimate(ni, nj, nk, *[_ptr(v) for v in a], cfl)
    if r <= 0:
        raise CgfdError("dt estimate failed")
    return r


def media_homogeneous(ni, nj, nk, vp, vs, rho):
    load()
    sh = (nk + 6, nj + 6, ni + 6)
    l, m, r = _z(sh), _z(sh), _z(sh)
    _check(_lib.cgfd_media_homogeneous(ni, nj, nk, vp, vs, rho, _ptr(l), _ptr(m), _ptr(r)))
    return l, m, r


def media_layered(ni, nj, nk, x, y, z, z_interface, vp, vpvs, rho, pert_std, pert_clip, corr_len, seed):
    load()
    sh = (nk + 6, nj + 6, ni + 6)
    l, m, r = _z(sh), _z(sh), _z(sh)
    zi, v, vv, rr = [_arr(a, np.float64) for a in (z_interface, vp, vpvs, rho)]
    d = LayerDesc(len(v), _ptr(zi), _ptr(v), _ptr(vv), _ptr(rr), pert_std, pert_clip, corr_len, seed)
    _check(_lib.cgfd_media_layered(ni, nj, nk, _ptr(_arr(x, np.float64)), _ptr(_arr(y, np.float64)),
                                   _ptr(_arr(z, np.float64)), C.byref(d), _ptr(l), _ptr(m), _ptr(r)))
    return l, m, r


def rho_to_slow(rho):
    load()
    r = _arr(rho, np.float64)
    s = np.zeros_like(r)
    _check(_lib.cgfd_rho_to_slow(r.size, _ptr(r), _ptr(s)))
    return s


def pml_len_dh(ni, nj, nk, x, y, z, face, nlay):
    load()
    L0, dh = C.c_double(), C.c_double()
    _check(_lib.cgfd_pml_len_dh(ni, nj, nk, _ptr(_arr(x, np.float64)), _ptr(_arr(y, np.float64)),
                                _ptr(_arr(z, np.float64)), face, nlay, C.byref(L0), C.byref(dh)))
    return L0.value, dh.value


def pml_profile(nlay, face, L0, dh, vel, amax, bmax):
    load()
    A, B, D = _z(nlay + 1), _z(nlay + 1), _z(nlay + 1)
    _check(_lib.cgfd_pml_profile(nlay, face, L0, dh, vel, amax, bmax, _ptr(A), _ptr(B), _ptr(D)))
    return A, B, D


def free_surface_matrices(ni, nj, nk, metric, lam, mu):
    load()
    mx = _z((nj + 6, ni + 6, 9))
    my = _z((nj + 6, ni + 6, 9))
    _check(_lib.cgfd_free_surface_matrices(ni, nj, nk, _ptr(_arr(metric, np.float64)), _ptr(_arr(lam, np.float64)),
                                           _ptr(_arr(mu, np.float64)), _ptr(mx), _ptr(my)))
    return mx, my


def wavelet(kind, t, c0, c1):
    load()
    return _lib.cgfd_wavelet(kind, t, c0, c1)


def angle2moment(strike, dip, rake):
    load()
    m = _z(6)
    _check(_lib.cgfd_angle2moment(strike, dip, rake, _ptr(m)))
    return m


def source_point(ni, nj, nk, sx, sy, sz, depth_flag, t0, dt, wavelet_tstart, kind, c0, c1, stf_length):
    """Returns dict(it_begin, it_end, ext_num, ext_indx[343], ext_coef[343], max_nt, stf[max_nt,4])."""
    load()
    itb, ite, en, mnt = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    ei = np.zeros(343, dtype=np.int64)
    ec = _z(343)
    args = (ni, nj, nk, sx, sy, sz, depth_flag, t0, dt, wavelet_tstart, kind, c0, c1, stf_length,
            C.byref(itb), C.byref(ite), C.byref(en), _ptr(ei), _ptr(ec), C.byref(mnt))
    _check(_lib.cgfd_source_point(*args, None))
    stf = _z((mnt.value, 4))
    _check(_lib.cgfd_source_point(*args, _ptr(stf)))
    return dict(it_begin=itb.value, it_end=ite.value, ext_num=en.value, ext_indx=ei, ext_coef=ec,
                max_nt=mnt.value, stf=stf)


def receiver_index(ni, nj, nk, rx, ry, rz, depth_flag):
    load()
    ix = np.zeros(8, dtype=np.int64)
    fr = _z(3)
    rc = _lib.cgfd_receiver_index(ni, nj, nk, rx, ry, rz, depth_flag, _ptr(ix), _ptr(fr))
    return (ix, fr) if rc == 0 else (None, None)


def locate_points(ni, nj, nk, x, y, z, pts, depth_flag=None):
    """Points given by coordinates (cgfd_locate_points): pts [n][3] = x, y and
    z or depth (where depth_flag is 1).  Returns (idx [n][3] physical 0-based
    lower corner, frac [n][3] in [0,1), found [n] bool)."""
    load()
    P = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, 3))
    n = P.shape[0]
    idx = np.zeros((n, 3), dtype=np.int32)
    fr = np.zeros((n, 3))
    fd = np.zeros(n, dtype=np.int32)
    xs = [_arr(a, np.float64) for a in (x, y, z)]
    df = None if depth_flag is None else np.ascontiguousarray(np.broadcast_to(np.asarray(depth_flag, np.int32), (n,)))
    _check(_lib.cgfd_locate_points(ni, nj, nk, *[_ptr(a) for a in xs], n, _ptr(P), None if df is None else _ptr(df),
                                   _ptr(idx), _ptr(fr), _ptr(fd)))
    return idx, fr, fd.astype(bool)


def media_layer_file(ni, nj, nk, x, y, z, path, method="loc"):
    """lam, mu, rho from a layer file in the md3lay format (cgfd_media_layer_file)."""
    load()
    sh = (nk + 6, nj + 6, ni + 6)
    lam, mu, rho = _z(sh), _z(sh), _z(sh)
    xs = [_arr(a, np.float64) for a in (x, y, z)]
    _check(_lib.cgfd_media_layer_file(ni, nj, nk, *[_ptr(a) for a in xs], str(path).encode(), method.encode(),
                                      _ptr(lam), _ptr(mu), _ptr(rho)))
    return lam, mu, rho


def media_grid_file(ni, nj, nk, x, y, z, path, method="loc"):
    """lam, mu, rho from a grid media file in the md3grd format (cgfd_media_grid_file)."""
    load()
    sh = (nk + 6, nj + 6, ni + 6)
    lam, mu, rho = _z(sh), _z(sh), _z(sh)
    xs = [_arr(a, np.float64) for a in (x, y, z)]
    _check(_lib.cgfd_media_grid_file(ni, nj, nk, *[_ptr(a) for a in xs], str(path).encode(), method.encode(),
                                     _ptr(lam), _ptr(mu), _ptr(rho)))
    return lam, mu, rho


def media_bin_file(ni, nj, nk, x, y, z, order, size, spacing, origin, rho_path, vp_path, vs_path):
    """lam, mu, rho from three raw float32 volumes (cgfd_media_bin_file).  order
    names the axis of each file dimension, fastest first: a string such as
    "xyz" or "zxy", or three of 0, 1, 2 for x, y, z."""
    load()
    if isinstance(order, str):
        order = ["xyz".index(c) for c in order]
    sh = (nk + 6, nj + 6, ni + 6)
    lam, mu, rho = _z(sh), _z(sh), _z(sh)
    xs = [_arr(a, np.float64) for a in (x, y, z)]
    o, n = np.asarray(order, np.int32), np.asarray(size, np.int32)
    h, g = np.asarray(spacing, np.float64), np.asarray(origin, np.float64)
    _check(_lib.cgfd_media_bin_file(ni, nj, nk, *[_ptr(a) for a in xs], _ptr(o), _ptr(n), _ptr(h), _ptr(g),
                                    str(rho_path).encode(), str(vp_path).encode(), str(vs_path).encode(),
                                    _ptr(lam), _ptr(mu), _ptr(rho)))
    return lam, mu, rho


def receiver_at(ni, nj, nk, idx, frac):
    """The 8 trilinear corner offsets and fractions of a receiver at a located
    point (the form cgfd_upload_receivers takes)."""
    i, j, k = (int(v) for v in idx)
    nx, ny = ni + 6, nj + 6
    p = (k + 3) * nx * ny + (j + 3) * nx + (i + 3)
    sy, sz = nx, nx * ny
    ix = np.array([p, p + 1, p + sy, p + 1 + sy, p + sz, p + 1 + sz, p + sy + sz, p + 1 + sy + sz], dtype=np.int64)
    return ix, np.asarray(frac, dtype=np.float64).copy()


def storage_scales(storage, ni, nj, nk, metric, lam, mu, slw, src, dt):
    """Power-of-two scale exponents of a reduced-precision session (cgfd_storage_scales)."""
    load()
    ex = np.zeros(9, dtype=np.int32)
    arrs = [_arr(a, np.float64) for a in (metric, lam, mu, slw)]
    if src is not None and src["nsrc"] > 0:
        d, keep = _source_desc(src)
        dp = C.byref(d)
    else:
        dp = None
    _check(_lib.cgfd_storage_scales(storage, ni, nj, nk, *[_ptr(a) for a in arrs], dp, float(dt), _ptr(ex)))
    return [int(v) for v in ex]


class Server:
    """Continuous-batching shot server (cgfd_serve.h): models stay resident on
    the device, requests on one model share a batched session."""

    def __init__(self, hbm_budget=0, max_batch=8, chunk_steps=16, group_share=0.0):
        load()
        self.h = C.c_void_p()
        d = ServerDesc(int(hbm_budget), int(max_batch), int(chunk_steps), float(group_share))
        _check(_lib.cgfd_server_create(C.byref(d), C.byref(self.h)))
        self._meta = {}

    def close(self):
        if self.h:
            _lib.cgfd_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_model(self, ni, nj, nk, metric, lam, mu, slw, dt, storage=STORE_F32, free_top=False, mvx=None,
                  mvy=None, pml=None, scale_exp2=None, max_nrecv=0, max_nt=1, dtype=np.float32):
        """pml: {face: (nlay, A, B, D)}; arrays in the reference host layout."""
        t = np.float64 if dtype == np.float64 else np.float32
        keep = [_arr(metric, t), _arr(lam, t), _arr(mu, t), _arr(slw, t)]
        md = ModelDesc()
        md.ni, md.nj, md.nk, md.storage = ni, nj, nk, storage
        md.dtype = DT_F64 if t == np.float64 else DT_F32
        md.metric, md.lam, md.mu, md.slw = [_ptr(a).value for a in keep]
        md.free_top = 1 if free_top else 0
        if free_top:
            keep += [_arr(mvx, t), _arr(mvy, t)]
            md.matVx2Vz, md.matVy2Vz = _ptr(keep[-2]).value, _ptr(keep[-1]).value
        for f, (nl, A, B, D) in (pml or {}).items():
            arrs = [_arr(A, t), _arr(B, t), _arr(D, t)]
            keep += arrs
            md.pml_nlay[f] = nl
            md.pml_A[f], md.pml_B[f], md.pml_D[f] = [_ptr(a).value for a in arrs]
        if scale_exp2 is not None:
            ex = np.ascontiguousarray(scale_exp2, dtype=np.int32)
            keep.append(ex)
            md.scale_exp2 = _ptr(ex).value
        md.dt = float(dt)
        md.max_nrecv, md.max_nt = int(max_nrecv), int(max_nt)
        mid = C.c_int()
        _check(_lib.cgfd_server_add_model(self.h, C.byref(md), C.byref(mid)))
        return mid.value

    def submit(self, model, nt_total, src, recv_indx, recv_frac):
        rq = Request()
        rq.model, rq.nt_total = model, int(nt_total)
        keep = []
        if src is not None and src["nsrc"] > 0:
            d, k = _source_desc(src)
            keep += [d, k]
            rq.src = C.cast(C.pointer(d), C.c_void_p).value
        ix = _arr(recv_indx, np.int64).reshape(-1, 8)
        fr = _arr(recv_frac, np.float64).reshape(-1, 3)
        rq.nrecv = ix.shape[0]
        rq.recv_indx8, rq.recv_dtype, rq.recv_frac3 = _ptr(ix).value, DT_F64, _ptr(fr).value
        t = C.c_longlong()
        _check(_lib.cgfd_server_submit(self.h, C.byref(rq), C.byref(t)))
        self._meta[t.value] = (ix.shape[0], int(nt_total))
        return t.value

    def wait(self, ticket, timeout=-1.0):
        """True when done, False on time-out; raises if the request was rejected."""
        r = _lib.cgfd_server_wait(self.h, ticket, float(timeout))
        if r == 1:
            return False
        _check(r)
        return True

    def result(self, ticket):
        """(seismograms [nrecv][9][nt], seconds to admission, seconds to completion)."""
        nr, nt = self._meta.pop(ticket)
        out = np.zeros((nr, 9, nt))
        ta, td = C.c_double(), C.c_double()
        _check(_lib.cgfd_server_result(self.h, ticket, DT_F64, _ptr(out), C.byref(ta), C.byref(td)))
        return out, ta.value, td.value

    def stats(self):
        st = ServerStats()
        _check(_lib.cgfd_server_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}


def hbm_stream(nbytes=2 << 30, reps=10):
    """(read GB/s, copy GB/s counting read + write bytes) of the streaming ceiling kernels."""
    load()
    r, c = C.c_double(), C.c_double()
    _check(_lib.cgfd_hbm_stream(int(nbytes), int(reps), C.byref(r), C.byref(c)))
    return r.value, c.value


def hbm_stream_dword(nbytes=2 << 30, reps=10):
    """Read GB/s of the 4-byte-per-lane streaming read (the sweep's access size)."""
    load()
    r = C.c_double()
    _check(_lib.cgfd_hbm_stream_dword(int(nbytes), int(reps), C.byref(r)))
    return r.value


def footprint_bytes(ni, nj, nk, storage=STORE_F32, nshots=1, nrecv=0, nt_total=0):
    """Device bytes a session of this shape holds (cgfd_footprint_bytes)."""
    load()
    d = Desc(ni, nj, nk, storage, nshots, 0)
    return int(_lib.cgfd_footprint_bytes(C.byref(d), nrecv, nt_total))


# ---------------------------------------------------------------- output files
def sac_write(path, data, delta, b=0.0, ev=(0, 0, 0, 0), st=(0, 0, 0, 0), kstnm="-12345", kcmpnm="-12345"):
    load()
    d = _arr(data, np.float32)
    e = _arr(ev, np.float32)
    t = _arr(st, np.float32)
    _check(_lib.cgfd_sac_write(str(path).encode(), _ptr(d), d.size, float(delta), float(b), _ptr(e), _ptr(t),
                               kstnm.encode(), kcmpnm.encode()))


def station_sac(out_dir, name, seismo, dt, x, y, z, strain=False, lam=0.0, mu=0.0):
    """The 9 (+6 strain) SAC files of one receiver; seismo [9][nt]."""
    load()
    sz = _arr(seismo, np.float32).reshape(9, -1)
    _check(_lib.cgfd_station_sac(str(out_dir).encode(), name.encode(), _ptr(sz), sz.shape[1], float(dt), float(x),
                                 float(y), float(z), 1 if strain else 0, float(lam), float(mu)))


class Snapshot:
    """NetCDF-3 snapshot file with the reference's names and attributes."""

    def __init__(self, path, ni, nj, nk, vel=True, stress=True, strain=False, first_global=(0, 0, 0),
                 first_local=(3, 3, 3), stride=(1, 1, 1), topo=(0, 0, 0)):
        load()
        d = SnapDesc(ni, nj, nk, int(vel), int(stress), int(strain), (C.c_int * 3)(*first_global),
                     (C.c_int * 3)(*first_local), (C.c_int * 3)(*stride), (C.c_int * 3)(*topo))
        self.h = C.c_void_p()
        self.n = ni * nj * nk
        _check(_lib.cgfd_snap_open(str(path).encode(), C.byref(d), C.byref(self.h)))

    def append(self, time, w9, lam=None, mu=None):
        w = _arr(w9, np.float32)
        la = _arr(lam, np.float32) if lam is not None else None
        m = _arr(mu, np.float32) if mu is not None else None
        _check(_lib.cgfd_snap_append(self.h, float(time), _ptr(w), _ptr(la), _ptr(m)))

    def close(self):
        if self.h:
            _check(_lib.cgfd_snap_close(self.h))
            self.h = C.c_void_p()


def slice_path(out_dir, axis, gindex, part="px0_py0_pz0"):
    """The reference's slice file name (io_slice_locate, io_funcs.cu:457-512):
    axis 0/1/2 -> slicex_i / slicey_j / slicez_k, 1-based plane number."""
    return os.path.join(str(out_dir), f"slice{'xyz'[axis]}_{'ijk'[axis]}{gindex + 1}_{part}.nc")


class Slice:
    """NetCDF-3 slice file with the reference's dimensions, names and attributes."""

    def __init__(self, path, axis, gindex, ni, nj, nk, topo=(0, 0, 0)):
        load()
        if axis not in (0, 1, 2):
            raise ValueError("slice axis is 0 (x), 1 (y) or 2 (z)")
        if not 0 <= gindex < (ni, nj, nk)[axis]:
            raise ValueError(f"slice plane {gindex} outside the grid's {(ni, nj, nk)[axis]} planes")
        self.axis = axis
        self.shape = [(nk, nj), (nk, ni), (nj, ni)][axis]
        self.h = C.c_void_p()
        t = (C.c_int * 3)(*topo)
        _check(_lib.cgfd_slice_open(str(path).encode(), axis, gindex + 3, self.shape[0], self.shape[1], t,
                                    C.byref(self.h)))

    def append(self, time, w9):
        """w9: [9][n_slow][n_fast] (e.g. a Session.fetch_subvolume plane, squeezed)."""
        w = _arr(np.asarray(w9).reshape(9, *self.shape), np.float32)
        _check(_lib.cgfd_slice_append(self.h, float(time), _ptr(w)))

    def close(self):
        if self.h:
            _check(_lib.cgfd_slice_close(self.h))
            self.h = C.c_void_p()


def session_slice(sess, shot, axis, gindex):
    """The plane gindex (0-based physical) of the nine W^n components, [9][slow][fast], gathered on the device."""
    ni, nj, nk = sess.ni, sess.nj, sess.nk
    start = [0, 0, 0]
    count = [ni, nj, nk]
    start[axis], count[axis] = gindex, 1
    w = sess.fetch_subvolume(shot, LEVEL_PRE, list(range(9)), tuple(start), tuple(count))
    return np.squeeze(w, axis=3 - axis)


class SourceFileInfo(C.Structure):
    """cgfd_source_file_info (include/cgfd_setup.h)."""
    _fields_ = [("event", C.c_char * 256), ("nsrc_file", C.c_int), ("nsrc", C.c_int), ("stf_given", C.c_int),
                ("stf_length", C.c_double), ("stf_dt", C.c_double), ("stf_nt", C.c_int), ("cmp_type", C.c_int),
                ("mechanism", C.c_int), ("by_coord", C.c_int), ("depth", C.c_int), ("force_on", C.c_int),
                ("moment_on", C.c_int), ("max_nt", C.c_int)]


def _host_memory_bytes():
    return os.sysconf("SC_PHYS_PAGES") * os.sysconf("SC_PAGE_SIZE")


def source_file(path, ni, nj, nk, t0, dt, x=None, y=None, z=None, mu=None, max_table_bytes=None):
    """A source file (src_read_locate_file) as (info dict, source dict for
    Session.upload_source, node [nsrc][3], shift [nsrc][3]).  Tables larger
    than max_table_bytes (default: half the host memory) raise CgfdError
    before anything is allocated."""
    load()
    xs = [_arr(a, np.float64) if a is not None else None for a in (x, y, z, mu)]
    xp = [_ptr(a) for a in xs]
    info = SourceFileInfo()
    r = _lib.cgfd_source_file(str(path).encode(), ni, nj, nk, *xp, float(t0), float(dt), C.byref(info),
                              *([None] * 9))
    if r < 0:
        _check(r)
    ns, mnt = info.nsrc, info.max_nt
    nbytes = (3 * info.force_on + 6 * info.moment_on) * max(ns, 1) * max(mnt, 1) * 4 * 8 + max(ns, 1) * 343 * 16
    limit = _host_memory_bytes() // 2 if max_table_bytes is None else max_table_bytes
    if nbytes > limit:
        raise CgfdError(f"source file {path}: tables of {nbytes} bytes ({max(ns, 1)} sources x {mnt} steps) "
                        f"exceed the {limit}-byte limit")
    itb = np.zeros(max(ns, 1), np.int32)
    ite = np.zeros(max(ns, 1), np.int32)
    en = np.zeros(max(ns, 1), np.int32)
    ei = np.zeros((max(ns, 1), 343), np.int64)
    ec = np.zeros((max(ns, 1), 343))
    F = np.zeros((3, max(ns, 1), max(mnt, 1), 4)) if info.force_on else None
    M = np.zeros((6, max(ns, 1), max(mnt, 1), 4)) if info.moment_on else None
    node = np.zeros((max(ns, 1), 3), np.int32)
    shift = np.zeros((max(ns, 1), 3))
    r = _lib.cgfd_source_file(str(path).encode(), ni, nj, nk, *xp, float(t0), float(dt), C.byref(info),
                              _ptr(itb), _ptr(ite), _ptr(en), _ptr(ei), _ptr(ec), _ptr(F), _ptr(M), _ptr(node),
                              _ptr(shift))
    if r < 0:
        _check(r)
    d = {k: getattr(info, k) for k, _ in SourceFileInfo._fields_}
    d["event"] = info.event.decode()
    src = dict(nsrc=ns, force_on=info.force_on, moment_on=info.moment_on, max_nt=mnt, max_ext=343,
               it_begin=itb[:ns], it_end=ite[:ns], ext_num=en[:ns], ext_indx=ei[:ns].ravel(),
               ext_coef=ec[:ns].ravel(), F=None if F is None else F[:, :ns].ravel(),
               M=None if M is None else M[:, :ns].ravel())
    return d, src, node[:ns], shift[:ns]


class Station(C.Structure):
    """cgfd_station (include/cgfd_setup.h)."""
    _fields_ = [("name", C.c_char * 64), ("by_coord", C.c_int), ("depth", C.c_int), ("inp", C.c_double * 3),
                ("inside", C.c_int), ("indx8", C.c_longlong * 8), ("frac", C.c_double * 3),
                ("xyz", C.c_double * 3)]


def station_file(path, ni, nj, nk, x=None, y=None, z=None, max_n=100000):
    """Stations of a station file (io_recv_read_locate): list of dicts with
    name, by_coord, depth, in, inside, indx8, frac, xyz, in file order."""
    load()
    xs = [_arr(a, np.float64) if a is not None else None for a in (x, y, z)]
    xp = [_ptr(a) for a in xs]
    nf = C.c_int(0)
    # a first pass with room for none reads and checks the file and gives its count
    r = _lib.cgfd_station_file(str(path).encode(), ni, nj, nk, *xp, 0, None, C.byref(nf))
    if r < 0:
        _check(r)
    n_out = min(nf.value, max_n)
    arr = (Station * max(n_out, 1))()
    r = _lib.cgfd_station_file(str(path).encode(), ni, nj, nk, *xp, n_out, arr, C.byref(nf))
    if r < 0:
        _check(r)
    out = []
    for q in range(r):
        s_ = arr[q]
        out.append(dict(name=s_.name.decode(), by_coord=s_.by_coord, depth=s_.depth, inp=list(s_.inp),
                        inside=bool(s_.inside), indx8=np.array(s_.indx8[:], dtype=np.int64),
                        frac=np.array(s_.frac[:]), xyz=np.array(s_.xyz[:])))
    return out


def line_points(ni, nj, nk, start, incre, count, x=None, y=None, z=None):
    """Receiver line (io_line_locate): (seq, indx8, frac3, xyz) of its points inside the grid."""
    load()
    seq = np.zeros(max(count, 1), dtype=np.int32)
    ix = np.zeros((max(count, 1), 8), dtype=np.int64)
    fr = np.zeros((max(count, 1), 3))
    xyz = np.zeros((max(count, 1), 3))
    st, inc = np.asarray(start, np.int32), np.asarray(incre, np.int32)
    xs = [_arr(a, np.float64) if a is not None else None for a in (x, y, z)]
    n = _lib.cgfd_line_points(ni, nj, nk, *[_ptr(a) for a in xs], _ptr(st), _ptr(inc), int(count), _ptr(seq),
                              _ptr(ix), _ptr(fr), _ptr(xyz) if x is not None else None)
    if n < 0:
        _check(n)
    return seq[:n], ix[:n], fr[:n], xyz[:n]


def line_sac(out_dir, line, seq, seismo, dt, x, y, z):
    """{out_dir}/{line}.no{seq}.{Vx..Txy}.sac of one line point (io_line_output_sac)."""
    load()
    s_ = _arr(seismo, np.float32)
    _check(_lib.cgfd_line_sac(str(out_dir).encode(), line.encode(), int(seq), _ptr(s_), s_.shape[-1], float(dt),
                              float(x), float(y), float(z)))


PEAK_NAMES = ["PGV", "PGVh", "PGVx", "PGVy", "PGVz", "PGA", "PGAh", "PGAx", "PGAy", "PGAz",
              "PGD", "PGDh", "PGDx", "PGDy", "PGDz"]


def peak_motion_write(path, pg, first_physical=(3, 3), count_physical=None, topo=(0, 0, 0)):
    """PG_V_A_D file from maps [15][ny][nx]."""
    load()
    m = _arr(pg, np.float32)
    ny, nx = m.shape[1], m.shape[2]
    cp = count_physical if count_physical is not None else (nx - 6, ny - 6)
    a = np.ascontiguousarray(first_physical, dtype=np.int32)
    b = np.ascontiguousarray(cp, dtype=np.int32)
    c = np.ascontiguousarray(topo, dtype=np.int32)
    _check(_lib.cgfd_peak_motion_write(str(path).encode(), nx, ny, _ptr(m), _ptr(a), _ptr(b), _ptr(c)))
