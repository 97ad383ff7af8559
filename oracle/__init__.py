"""The oracle: TEST INFRASTRUCTURE, never product code.

A CPU restatement of the reference's ``nufht_batch`` path (Ansal06/NUFHT
``src/fast_hankel_nufht.c``, ``src/expansions.c``, ``src/bounds.c``) in plain C
(``oracle/*.c``), plus an FP64 direct-sum reference.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this package, and only to check results or as the timed CPU baseline.

``load_ref_planner()`` returns the reference's own planner compiled from its
sources into ``oracle/_ref`` (see ``oracle/Makefile``), or ``None``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    out = subprocess.run(["make", "-C", _HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_besselj.restype = ctypes.c_double
        L.orc_besselj.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_besselj_wr.restype = ctypes.c_double
        L.orc_besselj_wr.argtypes = [ctypes.c_double] * 3
        L.orc_direct.argtypes = [ctypes.c_double, _DP, ctypes.c_int, _DP, ctypes.c_int,
                                 ctypes.c_int, _DP, ctypes.c_int, _DP, ctypes.c_int,
                                 ctypes.c_int]
        L.orc_direct_rows.argtypes = [ctypes.c_double, _DP, ctypes.c_int, _DP,
                                      ctypes.c_int, ctypes.c_int, _DP, _IP, ctypes.c_int,
                                      _DP, ctypes.c_int]
        L.orc_a_coef.restype = ctypes.c_double
        L.orc_a_coef.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_generate_tables.argtypes = [_DP, _DP, _IP]
        L.orc_fill_params.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_int]
        L.orc_generate_boxes.argtypes = [_DP, ctypes.c_int, _DP, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_void_p]
        L.orc_boxes_free.argtypes = [ctypes.c_void_p]
        L.orc_nufht_batch.argtypes = [ctypes.c_double, _DP, ctypes.c_int, _DP, ctypes.c_int,
                                      _DP, ctypes.c_int, _DP, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_set_plant.argtypes = [ctypes.c_int]
        L.orc_nufft3.argtypes = [ctypes.c_int, _DP, _DP, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, _DP, _DP, ctypes.c_int]
        L.orc_nufft3_direct.argtypes = [ctypes.c_int, _DP, _DP, ctypes.c_int, ctypes.c_int,
                                        _DP, _DP, ctypes.c_int]
        _LIB = L
    return _LIB


def _threads() -> int:
    """Worker threads for the direct sums: OMP_NUM_THREADS when set, else at
    most 16 (the host share of one GPU), never more than the cores."""
    env = os.environ.get("OMP_NUM_THREADS")
    n = int(env) if env and env.isdigit() else 16
    return max(1, min(n, os.cpu_count() or 1))


def _d(a):
    return a.ctypes.data_as(_DP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def besselj(nu: float, x) -> np.ndarray:
    L = lib()
    x = np.asarray(x, dtype=np.float64)
    return np.vectorize(lambda v: L.orc_besselj(float(nu), float(v)), otypes=[np.float64])(x)


def besselj_wr(nu: float, w, r) -> np.ndarray:
    L = lib()
    return np.vectorize(lambda a, b: L.orc_besselj_wr(float(nu), float(a), float(b)),
                        otypes=[np.float64])(np.asarray(w, np.float64), np.asarray(r, np.float64))


def direct(nu, r, c, w, nthreads=None) -> np.ndarray:
    """FP64 direct sum; c is (m, B); returns (n, B)."""
    L = lib()
    r, w = _f64(r), _f64(w)
    c = np.asarray(c, dtype=np.float64)
    if c.ndim == 1:
        c = c[:, None]
    cf = np.asfortranarray(c)
    m, B = cf.shape
    n = w.shape[0]
    g = np.zeros((n, B), dtype=np.float64, order="F")
    L.orc_direct(float(nu), _d(r), m, _d(cf), m, B, _d(w), n, _d(g), n,
                 nthreads or _threads())
    return np.ascontiguousarray(g)


def direct_rows(nu, r, c, w, rows, nthreads=None) -> np.ndarray:
    L = lib()
    r, w = _f64(r), _f64(w)
    c = np.asarray(c, dtype=np.float64)
    if c.ndim == 1:
        c = c[:, None]
    cf = np.asfortranarray(c)
    m, B = cf.shape
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    out = np.zeros((rows.shape[0], B), dtype=np.float64, order="F")
    L.orc_direct_rows(float(nu), _d(r), m, _d(cf), m, B, _d(w),
                      rows.ctypes.data_as(_IP), rows.shape[0], _d(out),
                      nthreads or _threads())
    return np.ascontiguousarray(out)


def tables():
    """(asy_coef[401,64], asy_z[401,32,16], wimp_k[401,32,16]) as bounds.c makes them."""
    L = lib()
    a = np.zeros((401, 64))
    z = np.zeros((401, 32, 16))
    k = np.zeros((401, 32, 16), dtype=np.int32)
    L.orc_generate_tables(_d(a), _d(z), k.ctypes.data_as(_IP))
    return a, z, k


class _Params(ctypes.Structure):
    _fields_ = [("nu", ctypes.c_double), ("tol", ctypes.c_double),
                ("z_split", ctypes.c_double), ("K_asy", ctypes.c_int),
                ("K_loc", ctypes.c_int), ("asy_coef", ctypes.c_double * 64),
                ("max_levels", ctypes.c_int), ("min_dim_prod", ctypes.c_int)]


class _Boxes(ctypes.Structure):
    _fields_ = [("loc", _IP), ("asy", _IP), ("dir", _IP),
                ("nloc", ctypes.c_int), ("nasy", ctypes.c_int), ("ndir", ctypes.c_int)]


def params(nu, tol=1e-8, z_split=float("nan"), K_asy=0, K_loc=0):
    """(rc, dict) from the restated fill_plan_parameters."""
    p = _Params()
    rc = lib().orc_fill_params(ctypes.byref(p), float(nu), float(tol), 0, 10000,
                               float(z_split), int(K_asy), int(K_loc))
    return rc, {"nu": p.nu, "tol": p.tol, "z_split": p.z_split, "K_asy": p.K_asy,
                "K_loc": p.K_loc, "asy_coef": np.array(p.asy_coef[:])}


def boxes(rs, ws, z_split, max_levels=0, min_dim_prod=10000, K_loc=0):
    """dict of (k,4) int arrays [i0,i1,j0,j1] from the restated generate_boxes."""
    rs, ws = _f64(rs), _f64(ws)
    b = _Boxes()
    L = lib()
    L.orc_generate_boxes(_d(rs), rs.shape[0], _d(ws), ws.shape[0], float(z_split),
                         int(max_levels), int(min_dim_prod), int(K_loc), ctypes.byref(b))
    out = {}
    for name in ("loc", "asy", "dir"):
        cnt = getattr(b, "n" + name)
        ptr = getattr(b, name)
        out[name] = (np.ctypeslib.as_array(ptr, shape=(cnt * 4,)).reshape(cnt, 4).copy()
                     if cnt else np.zeros((0, 4), dtype=np.int32))
    L.orc_boxes_free(ctypes.byref(b))
    return out


def nufht_batch(nu, rs, cs, ws, tol=1e-8, asy_mode=0, nthreads=1):
    """Ref-algorithm CPU nufht_batch: cs (m,B) -> (n,B).  asy_mode 0 = exact sums,
    1 = the oracle's CPU type-3 NUFFT."""
    rs, ws = _f64(rs), _f64(ws)
    c = np.asarray(cs, dtype=np.float64)
    if c.ndim == 1:
        c = c[:, None]
    cf = np.asfortranarray(c)
    m, B = cf.shape
    n = ws.shape[0]
    g = np.zeros((n, B), order="F")
    rc = lib().orc_nufht_batch(float(nu), _d(rs), m, _d(ws), n, _d(cf), m, _d(g), n, B,
                               float(tol), int(asy_mode), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"orc_nufht_batch returned {rc}")
    return np.ascontiguousarray(g)


def nufft3(x, c, s, iflag=1, eps=1e-8, exact=False):
    """c: complex (ntr, m) -> complex (ntr, n)."""
    L = lib()
    x, s = _f64(x), _f64(s)
    c = np.atleast_2d(np.asarray(c, dtype=np.complex128))
    ntr, m = c.shape
    ci = np.ascontiguousarray(c).view(np.float64)
    f = np.zeros((ntr, s.shape[0]), dtype=np.complex128)
    fv = f.view(np.float64)
    if exact:
        L.orc_nufft3_direct(m, _d(x), _d(ci), iflag, s.shape[0], _d(s), _d(fv), ntr)
    else:
        L.orc_nufft3(m, _d(x), _d(ci), iflag, float(eps), s.shape[0], _d(s), _d(fv), ntr)
    return f


# ---------------------------------------------------------------------------
# the reference's own planner (oracle/_ref), built from its sources
# ---------------------------------------------------------------------------
class RefBox(ctypes.Structure):
    _fields_ = [("i0", ctypes.c_int), ("i1", ctypes.c_int), ("j0", ctypes.c_int),
                ("j1", ctypes.c_int)]


class RefBoxVector(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(RefBox)), ("length", ctypes.c_size_t),
                ("capacity", ctypes.c_size_t)]


class RefBoxes(ctypes.Structure):
    _fields_ = [("loc", RefBoxVector), ("asy", RefBoxVector), ("dir", RefBoxVector)]


class RefPlan(ctypes.Structure):
    _fields_ = [("nu", ctypes.c_double), ("tol", ctypes.c_double),
                ("z_split", ctypes.c_double), ("K_asy", ctypes.c_int),
                ("K_loc", ctypes.c_int), ("asy_coef", ctypes.c_double * 64),
                ("m", ctypes.c_int), ("n", ctypes.c_int),
                ("rs", ctypes.c_void_p), ("ws", ctypes.c_void_p),
                ("max_levels", ctypes.c_int), ("min_dim_prod", ctypes.c_int),
                ("boxes", RefBoxes), ("tables_loaded", ctypes.c_int)]


class RefOptions(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_levels", ctypes.c_int),
                ("min_dim_prod", ctypes.c_int), ("z_split", ctypes.c_double),
                ("K_asy", ctypes.c_int), ("K_loc", ctypes.c_int)]


def load_ref_planner():
    """The reference planner compiled from its own sources, or None.

    It reads its plan tables from $NUFHT_TABLE_CACHE; callers point that at
    the reference's data/nufht_tables.cache before the first plan."""
    path = os.path.join(_HERE, "_ref", "libnufht_refplan.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    L.nufht_default_options.restype = RefOptions
    L.nufht_plan_init.argtypes = [ctypes.POINTER(RefPlan), ctypes.c_double, _DP,
                                  ctypes.c_int, _DP, ctypes.c_int,
                                  ctypes.POINTER(RefOptions)]
    L.nufht_plan_free.argtypes = [ctypes.POINTER(RefPlan)]
    L.generate_boxes.restype = RefBoxes
    L.generate_boxes.argtypes = [_DP, ctypes.c_int, _DP, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.boxes_free.argtypes = [ctypes.POINTER(RefBoxes)]
    return L


def ref_boxes_to_arrays(bx: RefBoxes):
    out = {}
    for name in ("loc", "asy", "dir"):
        v = getattr(bx, name)
        out[name] = np.array([[v.data[i].i0, v.data[i].i1, v.data[i].j0, v.data[i].j1]
                              for i in range(v.length)], dtype=np.int32).reshape(-1, 4)
    return out
