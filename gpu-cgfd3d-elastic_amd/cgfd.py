"""ctypes binding of libcgfd_el_iso.so, the gfx950 el_iso RK4 path.

The binding is thin on purpose: every array crosses the C ABI declared in
include/cgfd_el_iso.h as a plain pointer in the reference's host layout
(nx*ny*nz per component, nx = ni+6, i fastest).  Loading fails loudly when
the shared library has not been built; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcgfd_el_iso.so")

STORE_F32, STORE_F16, STORE_BF16, STORE_F64 = 0, 1, 2, 3
DT_F32, DT_F64 = 0, 1
LEVEL_PRE, LEVEL_TMPA, LEVEL_TMPB, LEVEL_END, LEVEL_RHS = 0, 1, 2, 3, 4
FLAG_PLAIN = 1

# every symbol the header declares (checked by the CPU test suite)
EXPORTS = [
    "cgfd_create", "cgfd_destroy", "cgfd_footprint_bytes", "cgfd_last_error",
    "cgfd_upload_metric", "cgfd_upload_medium", "cgfd_upload_free_surface",
    "cgfd_upload_pml", "cgfd_set_storage_scale", "cgfd_upload_source",
    "cgfd_upload_receivers", "cgfd_set_level", "cgfd_reset_shot", "cgfd_run",
    "cgfd_el_iso_onestage", "cgfd_read_seismo", "cgfd_fetch_level",
    "cgfd_fetch_pml_aux", "cgfd_set_timing", "cgfd_last_run_timing",
]


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
        "cgfd_last_run_timing": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)], i),
    }
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

    def set_timing(self, on):
        _check(_lib.cgfd_set_timing(self.h, 1 if on else 0))

    def last_run_timing(self):
        a, b, n = C.c_double(), C.c_double(), C.c_int()
        _check(_lib.cgfd_last_run_timing(self.h, C.byref(a), C.byref(b), C.byref(n)))
        return a.value, b.value, n.value
