"""MI355X-native batched nonuniform fast Hankel transform.

    g[i, t] = sum_j cs[j, t] * J_nu(k_i * r_j)

The drop-in surfaces are the C ABI in ``include/fast_hankel_nufht.h`` (the
reference's ``nufht_batch`` plan/execute API) and the ``_pynufht`` CPython
extension built next to this file.  This module is the Python host layer on
top of the same C ABI: plans over device-resident torch tensors, adjoints,
node gradients, uneven multi-transform batches and per-stage timing.

The HIP library must be built (``python -m nufht_amd.build`` or
``__graft_entry__.build()``); importing the bindings without it raises.
"""
from __future__ import annotations

import ctypes
import math
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_DP = ctypes.POINTER(ctypes.c_double)

PREC_FP64, PREC_AUTO, PREC_MIXED = 0, 1, 2
_PREC = {"fp64": PREC_FP64, "auto": PREC_AUTO, "mixed": PREC_MIXED}


class NufhtError(RuntimeError):
    pass


# ---------------------------------------------------------------------------
# C types (mirror include/fast_hankel_nufht.h and include/nufht_amd.h)
class Options(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_levels", ctypes.c_int),
                ("min_dim_prod", ctypes.c_int), ("z_split", ctypes.c_double),
                ("K_asy", ctypes.c_int), ("K_loc", ctypes.c_int)]


class Box(ctypes.Structure):
    _fields_ = [("i0", ctypes.c_int), ("i1", ctypes.c_int), ("j0", ctypes.c_int),
                ("j1", ctypes.c_int)]


class BoxVector(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(Box)), ("length", ctypes.c_size_t),
                ("capacity", ctypes.c_size_t)]


class Boxes(ctypes.Structure):
    _fields_ = [("loc", BoxVector), ("asy", BoxVector), ("dir", BoxVector)]


class CPlan(ctypes.Structure):
    _fields_ = [("nu", ctypes.c_double), ("tol", ctypes.c_double),
                ("z_split", ctypes.c_double), ("K_asy", ctypes.c_int),
                ("K_loc", ctypes.c_int), ("asy_coef", ctypes.c_double * 64),
                ("m", ctypes.c_int), ("n", ctypes.c_int),
                ("rs", ctypes.c_void_p), ("ws", ctypes.c_void_p),
                ("max_levels", ctypes.c_int), ("min_dim_prod", ctypes.c_int),
                ("boxes", Boxes), ("tables_loaded", ctypes.c_int),
                ("impl", ctypes.c_void_p)]


class CScratch(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("in_buffer", "out_buffer", "real_buffer_1",
                                                "real_buffer_2", "cheb_buffer",
                                                "bessel_buffer_1", "bessel_buffer_2")] + \
               [("m_alloc", ctypes.c_int), ("n_alloc", ctypes.c_int),
                ("Kloc_alloc", ctypes.c_int), ("extra_alloc", ctypes.c_int),
                ("impl", ctypes.c_void_p)]


class PlanStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_longlong) for n in (
        "n_loc", "n_asy", "n_dir", "pairs_dir", "loc_src", "loc_tgt", "asy_src", "asy_tgt",
        "asy_grid", "asy_grid_points", "asy_channels")] + \
        [("kernel_width", ctypes.c_int), ("precision", ctypes.c_int),
         ("K_asy", ctypes.c_int), ("K_loc", ctypes.c_int), ("z_split", ctypes.c_double),
         ("fft_classes", ctypes.c_longlong)] + \
        [(n, ctypes.c_longlong) for n in (
            "spread_sc_valu", "spread_sc_mfma", "interp_tc_valu", "interp_tc_mfma")] + \
        [(n, ctypes.c_double) for n in ("asy_fine", "asy_fine_nlogn", "fft_bytes", "model_f64",
                                        "model_f32", "model_bytes", "model_dir", "model_loc")] + \
        [(n, ctypes.c_longlong) for n in ("fused_boxes", "fused_sc", "fused_tc", "loc_eval_multi")]


class Knobs(ctypes.Structure):
    """NufhtAmdKnobs: launch-geometry choices (and the precision mode)."""
    _fields_ = [("dense_cell", ctypes.c_double), ("dense_split", ctypes.c_int),
                ("spread_split", ctypes.c_int), ("interp_mfma", ctypes.c_int),
                ("own_fft", ctypes.c_int), ("precision", ctypes.c_int),
                ("spread_sched", ctypes.c_int), ("fused_max", ctypes.c_int)]


def _knobs_dict(k: Knobs) -> Dict[str, float]:
    return {f: getattr(k, f) for f, _ in Knobs._fields_}


# Every exported symbol of the C ABI, as declared in include/*.h.
EXPORTS = [
    "nufht_default_options", "boxes_free", "nufht_plan_init", "nufht_plan_free",
    "nufht_scratch_init", "nufht_scratch_free", "generate_boxes", "split_box",
    "nufht_batch", "nufht_inplace", "nufht",
    "nufht_amd_plan_set_precision", "nufht_amd_plan_precision", "nufht_batch_device",
    "nufht_adjoint_device", "nufht_adjoint", "nufht_grad_nodes_device",
    "nufht_multi_create", "nufht_multi_execute", "nufht_multi_free",
    "nufht_amd_profile", "nufht_amd_profile_read", "nufht_amd_profile_reset",
    "nufht_amd_plan_stats", "nufht_amd_autotune", "nufht_amd_version",
    "nufht_amd_last_error", "nufht_amd_get_knobs", "nufht_amd_set_knobs",
    "nufht_multi_autotune", "nufht_multi_get_knobs", "nufht_amd_plan_cache",
    "nufht_amd_write_table_cache", "nufht_batch_device_noalloc", "nufht_multi_stats",
    "nufht_amd_stream_fork",
]


def lib_path() -> str:
    return os.path.join(_HERE, "libnufht_amd.so")


def lib():
    """The loaded C-ABI library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path()
    if not os.path.exists(path):
        raise NufhtError(f"{path} is missing: build it with `python -m nufht_amd.build` "
                         "(the transform has no CPU fallback)")
    L = ctypes.CDLL(path)
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    PP = ctypes.POINTER(CPlan)
    L.nufht_default_options.restype = Options
    L.nufht_plan_init.argtypes = [PP, d, _DP, i, _DP, i, ctypes.POINTER(Options)]
    L.nufht_plan_free.argtypes = [PP]
    L.nufht_scratch_init.argtypes = [ctypes.POINTER(CScratch), PP]
    L.nufht_scratch_free.argtypes = [ctypes.POINTER(CScratch)]
    L.generate_boxes.restype = Boxes
    L.generate_boxes.argtypes = [_DP, i, _DP, i, d, i, i, i]
    L.split_box.argtypes = [_DP, i, _DP, i, ctypes.POINTER(Box), d,
                            ctypes.POINTER(i), ctypes.POINTER(i)]
    L.boxes_free.argtypes = [ctypes.POINTER(Boxes)]
    L.nufht_batch.argtypes = [PP, _DP, i, _DP, i, i, ctypes.POINTER(CScratch)]
    L.nufht_inplace.argtypes = [_DP, d, _DP, i, _DP, _DP, i, ctypes.POINTER(Options)]
    L.nufht.restype = _DP
    L.nufht.argtypes = [d, _DP, i, _DP, _DP, i, ctypes.POINTER(Options)]
    L.nufht_amd_plan_set_precision.argtypes = [PP, i]
    L.nufht_amd_plan_precision.argtypes = [PP]
    L.nufht_batch_device.argtypes = [PP, vp, i, vp, i, i, vp]
    L.nufht_batch_device_noalloc.argtypes = [PP, vp, i, vp, i, i, vp]
    L.nufht_adjoint_device.argtypes = [PP, vp, i, vp, i, i, vp]
    L.nufht_adjoint.argtypes = [PP, _DP, i, _DP, i, i]
    L.nufht_grad_nodes_device.argtypes = [PP, vp, i, vp, i, i, vp, vp, vp]
    L.nufht_multi_create.argtypes = [ctypes.POINTER(vp), i, ctypes.POINTER(PP),
                                     ctypes.POINTER(i)]
    L.nufht_multi_execute.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(i),
                                      ctypes.POINTER(vp), ctypes.POINTER(i), vp, i]
    L.nufht_multi_free.argtypes = [vp]
    L.nufht_amd_profile.argtypes = [i]
    L.nufht_amd_stream_fork.argtypes = [i]
    L.nufht_amd_profile_read.argtypes = [ctypes.POINTER(d), ctypes.POINTER(ctypes.c_longlong),
                                         i, ctypes.POINTER(ctypes.c_char_p)]
    L.nufht_amd_plan_stats.argtypes = [PP, i, ctypes.POINTER(PlanStats)]
    L.nufht_amd_autotune.argtypes = [PP, i, vp]
    L.nufht_amd_get_knobs.argtypes = [PP, ctypes.POINTER(Knobs)]
    L.nufht_amd_set_knobs.argtypes = [PP, ctypes.POINTER(Knobs)]
    L.nufht_multi_autotune.argtypes = [vp, vp]
    L.nufht_multi_get_knobs.argtypes = [vp, ctypes.POINTER(Knobs)]
    L.nufht_multi_stats.argtypes = [vp, ctypes.POINTER(PlanStats)]
    L.nufht_amd_plan_cache.argtypes = [i, ctypes.POINTER(ctypes.c_longlong)]
    L.nufht_amd_write_table_cache.argtypes = [ctypes.c_char_p]
    L.nufht_amd_version.restype = ctypes.c_char_p
    L.nufht_amd_last_error.restype = ctypes.c_char_p
    _LIB = L
    return L


def plan_cache(capacity: int = -1) -> int:
    """Device-state cache of freed plans: sets its capacity when capacity >= 0
    (0 empties and disables it) and returns the number of reuses so far."""
    hits = ctypes.c_longlong()
    _check(lib().nufht_amd_plan_cache(int(capacity), ctypes.byref(hits)), "nufht_amd_plan_cache")
    return hits.value


def last_error() -> str:
    return lib().nufht_amd_last_error().decode()


def _check(rc: int, what: str):
    if rc != 0:
        raise NufhtError(f"{what} returned {rc}: {last_error()}")


def default_options() -> Options:
    return lib().nufht_default_options()


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _boxes_to_arrays(bx: Boxes) -> Dict[str, np.ndarray]:
    out = {}
    for name in ("loc", "asy", "dir"):
        v = getattr(bx, name)
        out[name] = np.array([[v.data[q].i0, v.data[q].i1, v.data[q].j0, v.data[q].j1]
                              for q in range(v.length)], dtype=np.int64).reshape(-1, 4)
    return out


def generate_boxes(rs, ws, z_split, max_levels=0, min_dim_prod=10000, K_loc=0):
    """The C ABI's generate_boxes (reference src/fast_hankel_nufht.c:578)."""
    L = lib()
    r, w = _f64(rs), _f64(ws)
    bx = L.generate_boxes(r.ctypes.data_as(_DP), r.size, w.ctypes.data_as(_DP), w.size,
                          float(z_split), int(max_levels), int(min_dim_prod), int(K_loc))
    out = _boxes_to_arrays(bx)
    L.boxes_free(ctypes.byref(bx))
    return out


def _stream_handle(stream=None) -> Optional[int]:
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


def _check_out(t, shape, device, what: str):
    """A destination buffer: float64, on `device`, contiguous, of `shape`
    (the column-major (rows, B) data viewed as a (B, rows) tensor)."""
    import torch
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda:
        raise ValueError(f"{what} must be a float64 CUDA (HIP) tensor")
    if t.device != device:
        raise ValueError(f"{what} is on {t.device}, the inputs on {device}")
    if tuple(t.shape) != tuple(shape) or not t.is_contiguous():
        raise ValueError(f"{what} must be a contiguous tensor of shape {tuple(shape)}, "
                         f"got {tuple(t.shape)}{'' if t.is_contiguous() else ' (strided)'}")


def _colmajor(t, rows: int):
    """(rows, B) torch tensor -> (contiguous (B, rows) tensor, ld) without copying when
    the caller already holds column-major data."""
    import torch
    if t.dim() == 1:
        t = t[:, None]
    if t.dtype != torch.float64:
        raise TypeError("coefficients must be float64")
    if not t.is_cuda:
        raise TypeError("device entry points take CUDA (HIP) tensors")
    if t.shape[0] != rows:
        raise ValueError(f"expected {rows} rows, got {t.shape[0]}")
    tt = t.t()
    if not tt.is_contiguous():
        tt = tt.contiguous()
    return tt, rows


class Plan:
    """A NUFHT plan: nufht_plan_init(nu, r, k, options) plus device execution.

    ``forward(cs)`` computes g = H cs for cs of shape (m, B) (torch CUDA float64,
    or numpy through the host path); ``adjoint(y)`` computes H^T y;
    ``grad_nodes(cs, rho)`` returns (dL/dr, dL/dk) for rho = dL/dg.
    """

    def __init__(self, nu: float, r, k, tol: float = 1e-8, precision: str = "auto",
                 max_levels: int = 0, min_dim_prod: int = 10000, z_split: float = float("nan"),
                 K_asy: int = 0, K_loc: int = 0):
        L = lib()
        self._r = _f64(r)
        self._k = _f64(k)
        self.m, self.n = self._r.size, self._k.size
        opt = L.nufht_default_options()
        opt.tol, opt.max_levels, opt.min_dim_prod = tol, max_levels, min_dim_prod
        opt.z_split, opt.K_asy, opt.K_loc = z_split, K_asy, K_loc
        self._plan = CPlan()
        self._live = False
        rc = L.nufht_plan_init(ctypes.byref(self._plan), float(nu),
                               self._r.ctypes.data_as(_DP), self.m,
                               self._k.ctypes.data_as(_DP), self.n, ctypes.byref(opt))
        _check(rc, "nufht_plan_init")
        self._live = True
        self.nu, self.tol = float(nu), float(tol)
        self.set_precision(precision)

    # -- parameters ----------------------------------------------------------
    @property
    def z_split(self) -> float:
        return self._plan.z_split

    @property
    def K_asy(self) -> int:
        return self._plan.K_asy

    @property
    def K_loc(self) -> int:
        return self._plan.K_loc

    @property
    def boxes(self) -> Dict[str, np.ndarray]:
        return _boxes_to_arrays(self._plan.boxes)

    def set_precision(self, precision: str):
        _check(lib().nufht_amd_plan_set_precision(ctypes.byref(self._plan), _PREC[precision]),
               "nufht_amd_plan_set_precision")

    @property
    def precision(self) -> str:
        return {0: "fp64", 2: "mixed"}[lib().nufht_amd_plan_precision(ctypes.byref(self._plan))]

    def stats(self, batch: int) -> Dict[str, float]:
        st = PlanStats()
        _check(lib().nufht_amd_plan_stats(ctypes.byref(self._plan), int(batch),
                                          ctypes.byref(st)), "nufht_amd_plan_stats")
        return {f: getattr(st, f) for f, _ in PlanStats._fields_}

    # -- execution -----------------------------------------------------------
    def forward(self, cs, out=None, stream=None):
        """g = H cs.  torch CUDA (m, B) -> torch (n, B) (column-major view);
        numpy (m, B) -> numpy (n, B) via the host-pointer nufht_batch."""
        if isinstance(cs, np.ndarray):
            return self._forward_host(cs)
        import torch
        c, ldc = _colmajor(cs, self.m)
        B = c.shape[0]
        if out is not None:
            _check_out(out, (B, self.n), c.device, "out")
        g = out if out is not None else torch.empty((B, self.n), dtype=torch.float64,
                                                    device=c.device)
        rc = lib().nufht_batch_device(ctypes.byref(self._plan), c.data_ptr(), ldc, g.data_ptr(),
                                      self.n, B, _stream_handle(stream))
        _check(rc, "nufht_batch_device")
        return g.t()

    def _forward_host(self, cs: np.ndarray) -> np.ndarray:
        c = np.asfortranarray(np.asarray(cs, dtype=np.float64).reshape(self.m, -1))
        B = c.shape[1]
        g = np.zeros((self.n, B), order="F")
        rc = lib().nufht_batch(ctypes.byref(self._plan), c.ctypes.data_as(_DP), self.m,
                               g.ctypes.data_as(_DP), self.n, B, None)
        _check(rc, "nufht_batch")
        return np.ascontiguousarray(g)

    def adjoint(self, ys, stream=None):
        """x = H^T y.  torch CUDA (n, B) -> torch (m, B)."""
        if isinstance(ys, np.ndarray):
            y = np.asfortranarray(np.asarray(ys, dtype=np.float64).reshape(self.n, -1))
            B = y.shape[1]
            x = np.zeros((self.m, B), order="F")
            _check(lib().nufht_adjoint(ctypes.byref(self._plan), y.ctypes.data_as(_DP), self.n,
                                       x.ctypes.data_as(_DP), self.m, B), "nufht_adjoint")
            return np.ascontiguousarray(x)
        import torch
        y, ldy = _colmajor(ys, self.n)
        B = y.shape[0]
        x = torch.empty((B, self.m), dtype=torch.float64, device=y.device)
        rc = lib().nufht_adjoint_device(ctypes.byref(self._plan), y.data_ptr(), ldy, x.data_ptr(),
                                        self.m, B, _stream_handle(stream))
        _check(rc, "nufht_adjoint_device")
        return x.t()

    def grad_nodes(self, cs, rho, want_r=True, want_k=True, stream=None):
        """(dL/dr, dL/dk) given rho = dL/dg; torch CUDA tensors."""
        import torch
        c, ldc = _colmajor(cs, self.m)
        p, ldp = _colmajor(rho, self.n)
        B = c.shape[0]
        gr = torch.empty(self.m, dtype=torch.float64, device=c.device) if want_r else None
        gk = torch.empty(self.n, dtype=torch.float64, device=c.device) if want_k else None
        rc = lib().nufht_grad_nodes_device(ctypes.byref(self._plan), c.data_ptr(), ldc,
                                           p.data_ptr(), ldp, B,
                                           gr.data_ptr() if gr is not None else None,
                                           gk.data_ptr() if gk is not None else None,
                                           _stream_handle(stream))
        _check(rc, "nufht_grad_nodes_device")
        return gr, gk

    def autotune(self, batch: int, stream=None) -> Dict[str, float]:
        """Times the launch-geometry knobs (and, at tol >= 1e-6, the precision
        mode) for this plan and batch, keeps the fastest and returns them."""
        _check(lib().nufht_amd_autotune(ctypes.byref(self._plan), int(batch),
                                        _stream_handle(stream)), "nufht_amd_autotune")
        return self.knobs()

    def knobs(self) -> Dict[str, float]:
        k = Knobs()
        _check(lib().nufht_amd_get_knobs(ctypes.byref(self._plan), ctypes.byref(k)),
               "nufht_amd_get_knobs")
        return _knobs_dict(k)

    def set_knobs(self, **kw):
        """Overrides knobs by name (dense_cell, dense_split, spread_split,
        interp_mfma, own_fft, precision); the next call rebuilds."""
        k = Knobs()
        _check(lib().nufht_amd_get_knobs(ctypes.byref(self._plan), ctypes.byref(k)),
               "nufht_amd_get_knobs")
        for name, v in kw.items():
            if name not in dict(Knobs._fields_):
                raise KeyError(name)
            setattr(k, name, v)
        _check(lib().nufht_amd_set_knobs(ctypes.byref(self._plan), ctypes.byref(k)),
               "nufht_amd_set_knobs")

    def close(self):
        if self._live:
            lib().nufht_plan_free(ctypes.byref(self._plan))
            self._live = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MultiBatch:
    """Uneven batch of independent transforms run with one launch per stage
    (optionally replayed as a hipGraph).  ``plans`` must stay alive."""

    def __init__(self, plans: Sequence[Plan], batches: Sequence[int]):
        L = lib()
        self.plans = list(plans)
        self.batches = [int(b) for b in batches]
        arr = (ctypes.POINTER(CPlan) * len(plans))(*[ctypes.pointer(p._plan) for p in plans])
        bs = (ctypes.c_int * len(plans))(*self.batches)
        self._h = ctypes.c_void_p()
        _check(L.nufht_multi_create(ctypes.byref(self._h), len(plans), arr, bs),
               "nufht_multi_create")

    def execute(self, cs_list, gs_list, use_graph=True, stream=None):
        """cs_list[k]: contiguous (B_k, m_k) tensors (column-major data);
        gs_list[k]: contiguous (B_k, n_k) outputs."""
        L = lib()
        K = len(self.plans)
        if len(cs_list) != K or len(gs_list) != K:
            raise ValueError(f"expected {K} coefficient and {K} output tensors, got "
                             f"{len(cs_list)} and {len(gs_list)}")
        dev = getattr(cs_list[0], "device", None) if K else None
        for q, (p, b, c, g) in enumerate(zip(self.plans, self.batches, cs_list, gs_list)):
            _check_out(c, (b, p.m), dev, f"cs_list[{q}]")
            _check_out(g, (b, p.n), dev, f"gs_list[{q}]")
        cp = (ctypes.c_void_p * K)(*[c.data_ptr() for c in cs_list])
        gp = (ctypes.c_void_p * K)(*[g.data_ptr() for g in gs_list])
        lc = (ctypes.c_int * K)(*[p.m for p in self.plans])
        lg = (ctypes.c_int * K)(*[p.n for p in self.plans])
        _check(L.nufht_multi_execute(self._h, cp, lc, gp, lg, _stream_handle(stream),
                                     int(bool(use_graph))), "nufht_multi_execute")

    def autotune(self, stream=None) -> Dict[str, float]:
        """Times the launch-geometry knobs for the whole batch, keeps the
        fastest (the program is rebuilt) and returns them."""
        _check(lib().nufht_multi_autotune(self._h, _stream_handle(stream)),
               "nufht_multi_autotune")
        return self.knobs()

    def stats(self) -> Dict[str, float]:
        """Plan statistics summed over the batch's transforms (see
        ``Plan.stats``)."""
        st = PlanStats()
        _check(lib().nufht_multi_stats(self._h, ctypes.byref(st)), "nufht_multi_stats")
        return {f: getattr(st, f) for f, _ in PlanStats._fields_}

    def knobs(self) -> Dict[str, float]:
        k = Knobs()
        _check(lib().nufht_multi_get_knobs(self._h, ctypes.byref(k)), "nufht_multi_get_knobs")
        d = _knobs_dict(k)
        d.pop("precision")
        return d

    def close(self):
        if self._h:
            lib().nufht_multi_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# per-stage device timing (HIP events around each stage launch)
STAGES = ["zero", "permute", "direct", "loc_project", "loc_eval", "asy_grid_zero",
          "asy_spread", "asy_spread_mfma", "asy_fft", "asy_interp", "asy_interp_mfma",
          "grad_combine", "asy_fused"]


def stream_fork(at: int = -1) -> int:
    """Where calls fork the near-field stages onto a second stream (1 before
    the spreads, 2 before the FFTs, 0 never); at < 0 only queries.  Returns
    the previous setting."""
    return int(lib().nufht_amd_stream_fork(int(at)))


def profile(enable: bool = True):
    lib().nufht_amd_profile(int(bool(enable)))


def profile_reset():
    lib().nufht_amd_profile_reset()


def profile_read() -> Dict[str, Dict[str, float]]:
    n = len(STAGES)
    ms = (ctypes.c_double * n)()
    cnt = (ctypes.c_longlong * n)()
    names = (ctypes.c_char_p * n)()
    m = min(n, int(lib().nufht_amd_profile_read(ms, cnt, n, names)))
    return {names[q].decode(): {"ms": ms[q], "launches": cnt[q]} for q in range(m)}


def version() -> str:
    return lib().nufht_amd_version().decode()


def load_pynufht():
    """Import the drop-in ``_pynufht`` extension built next to this package."""
    import importlib.util
    import sysconfig
    path = os.path.join(_HERE, "_pynufht" + (sysconfig.get_config_var("EXT_SUFFIX") or ".so"))
    if not os.path.exists(path):
        raise NufhtError(f"{path} is missing: build it with `python -m nufht_amd.build`")
    spec = importlib.util.spec_from_file_location("_pynufht", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod
