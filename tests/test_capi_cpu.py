"""The C-ABI library on the CPU (no device calls): symbols, planner parity
with the reference planner, error codes, and the _pynufht argument checks."""
import ctypes
import json
import math
import os
import re

import numpy as np
import pytest

import nufht_amd as na
from nufht_amd import workloads as W
from tests.conftest import GOLDEN, ROOT


def _declared_functions():
    names = set()
    for hdr in ("fast_hankel_nufht.h", "nufht_amd.h"):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(",
                             text, flags=re.M):
            if m.group(1) not in ("if", "while", "for", "sizeof", "return"):
                names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = na.lib()
    declared = _declared_functions()
    assert declared == set(na.EXPORTS), declared ^ set(na.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name


def test_default_options_match_reference():
    o = na.default_options()
    assert (o.tol, o.max_levels, o.min_dim_prod, o.K_asy, o.K_loc) == (1e-8, 0, 10000, 0, 0)
    assert math.isnan(o.z_split)


def _grids(c):
    gen = c["gen"]
    if gen[0] == "gaussian_gl":
        g = W.gaussian_gl()
        return g["r"], g["k"]
    s = W.Stream(gen[1], 77)
    if gen[0] == "uni":
        return np.sort(s.uniform(gen[2], 0, gen[4])), np.sort(s.uniform(gen[3], 0, gen[5]))
    return np.sort(s.uniform(gen[2], 100.0, 200.0)), np.sort(s.uniform(gen[3], 100.0, 200.0))


def test_plan_parameters_and_boxes_match_reference_planner():
    d = json.load(open(os.path.join(GOLDEN, "ref_boxes.json")))
    for c in d["cases"]:
        r, k = _grids(c)
        p = na.Plan(c["nu"], r, k, tol=c["tol"])
        if c["name"] == "uni_nu40":
            # reference table hole (K_asy = 10 reads an unfilled slot): filled here
            assert c["z_split"] == 0.0 and p.z_split > 100.0 and p.K_loc > 0
            p.close()
            continue
        assert p.K_asy == c["K_asy"] and p.K_loc == c["K_loc"]
        if math.isnan(c["z_split"]):
            assert math.isnan(p.z_split)
        else:
            assert abs(p.z_split - c["z_split"]) <= 3e-11 * c["z_split"]
        bx = p.boxes
        for t in ("loc", "asy", "dir"):
            np.testing.assert_array_equal(bx[t], np.array(c["boxes"][t]).reshape(-1, 4))
        p.close()


def test_plan_error_codes_match_reference():
    d = json.load(open(os.path.join(GOLDEN, "ref_boxes.json")))
    L = na.lib()
    r = np.array([1.0, 2.0])
    for e in d["params"]:
        opt = L.nufht_default_options()
        opt.tol = e["tol"]
        plan = na.CPlan()
        rc = L.nufht_plan_init(ctypes.byref(plan), float(e["nu"]),
                               r.ctypes.data_as(na._DP), 2, r.ctypes.data_as(na._DP), 2,
                               ctypes.byref(opt))
        assert rc == e["rc"], e
        if rc == 0:
            assert (plan.K_asy, plan.K_loc if plan.K_asy < 10 else e["K_loc"]) == \
                (e["K_asy"], e["K_loc"]), e
            L.nufht_plan_free(ctypes.byref(plan))
    plan = na.CPlan()
    assert L.nufht_plan_init(ctypes.byref(plan), 0.0, None, 0, None, 0, None) == -200


def test_batch_argument_errors():
    L = na.lib()
    r = np.linspace(0.1, 1.0, 10)
    p = na.Plan(0.0, r, r)
    c = np.zeros(10)
    g = np.zeros(10)
    PP = ctypes.byref(p._plan)
    assert L.nufht_batch(PP, c.ctypes.data_as(na._DP), 10, g.ctypes.data_as(na._DP), 10, 0, None) == -400
    assert L.nufht_batch(PP, None, 10, g.ctypes.data_as(na._DP), 10, 1, None) == -400
    assert L.nufht_batch(PP, c.ctypes.data_as(na._DP), 9, g.ctypes.data_as(na._DP), 10, 1, None) == -401
    assert L.nufht_batch(PP, c.ctypes.data_as(na._DP), 10, g.ctypes.data_as(na._DP), 9, 1, None) == -401
    p.close()


def test_generate_boxes_and_split_box_abi():
    s = W.Stream(4, 4)
    r = np.sort(s.uniform(2000, 0, 200.0))
    k = np.sort(s.uniform(1500, 0, 100.0))
    import oracle
    mine = na.generate_boxes(r, k, 10.9, K_loc=14)
    ref = oracle.boxes(r, k, 10.9, K_loc=14)
    for t in ("loc", "asy", "dir"):
        np.testing.assert_array_equal(mine[t], ref[t])
    # unsorted grids: the drop-in keeps the reference's linear-scan semantics
    perm = np.argsort(s.uniform(r.size))
    mine_u = na.generate_boxes(r[perm], k, 10.9, K_loc=14)
    ref_u = oracle.boxes(r[perm], k, 10.9, K_loc=14)
    for t in ("loc", "asy", "dir"):
        np.testing.assert_array_equal(mine_u[t], ref_u[t])
    L = na.lib()
    box = na.Box(0, k.size - 1, 0, r.size - 1)
    i_out, j_out = ctypes.c_int(), ctypes.c_int()
    L.split_box(r.ctypes.data_as(na._DP), r.size, k.ctypes.data_as(na._DP), k.size,
                ctypes.byref(box), 10.9, ctypes.byref(i_out), ctypes.byref(j_out))
    oi, oj = ctypes.c_int(), ctypes.c_int()
    ob = oracle.lib()
    ob.orc_split_box.argtypes = [oracle._DP, ctypes.c_int, oracle._DP, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
    b4 = (ctypes.c_int * 4)(0, k.size - 1, 0, r.size - 1)
    ob.orc_split_box(r.ctypes.data_as(oracle._DP), r.size, k.ctypes.data_as(oracle._DP),
                     k.size, ctypes.cast(b4, ctypes.c_void_p), 10.9, ctypes.byref(oi),
                     ctypes.byref(oj))
    assert (i_out.value, j_out.value) == (oi.value, oj.value)


def test_half_integer_execution_partition_is_stable():
    """For half-integer orders the public plan mirrors the reference (one loc
    box, K_loc = -1, z_split NaN) while execution uses a split partition."""
    r = np.linspace(0.01, 60.0, 3000)
    k = np.linspace(0.01, 60.0, 2000)
    p = na.Plan(1.5, r, k, tol=1e-8)
    assert math.isnan(p.z_split) and p.K_loc == -1 and p.K_asy == 1
    assert len(p.boxes["loc"]) == 1
    p.close()


def test_pynufht_argument_errors_without_device():
    pn = na.load_pynufht()
    r = np.linspace(0.1, 5.0, 20)
    k = np.linspace(0.1, 5.0, 10)
    with pytest.raises(ValueError):
        pn.nufht_batch(0.0, r, np.zeros(20), k)
    with pytest.raises(ValueError):
        pn.nufht_batch(0.0, r, np.zeros((19, 2)), k)
    with pytest.raises(ValueError):
        pn.nufht_batch_loop(0.0, r, np.zeros((19, 2)), k)
    with pytest.raises(RuntimeError):
        pn.nufht_batch(0.3, r, np.zeros((20, 2)), k)
    # as in the reference, the single-column path reports and returns zeros
    out = pn.nufht(0.3, r, np.zeros(20), k)
    assert out.shape == (10,) and np.all(out == 0.0)


def test_negative_half_integer_orders_refused():
    """J_{-(n+1/2)} is not J_{n+1/2} (the reference computes the latter):
    such orders are refused at plan time with -2."""
    L = na.lib()
    r = np.array([1.0, 2.0])
    for nu in (-0.5, -1.5, -9.5):
        plan = na.CPlan()
        rc = L.nufht_plan_init(ctypes.byref(plan), nu, r.ctypes.data_as(na._DP), 2,
                               r.ctypes.data_as(na._DP), 2, None)
        assert rc == -2, nu
        assert "negative half-integer" in na.last_error()
    plan = na.CPlan()
    assert L.nufht_plan_init(ctypes.byref(plan), 1.5, r.ctypes.data_as(na._DP), 2,
                             r.ctypes.data_as(na._DP), 2, None) == 0
    L.nufht_plan_free(ctypes.byref(plan))


def test_asymptotic_grid_size_guard():
    """Uniform grids of asymptotic boxes: L = ceil(X/h) + kw/2 + 2, FFT length
    the smallest 2^a, 3 2^a (or 9 2^a above 4096) of at least 2(2L + 1);
    lengths beyond 2^30 are refused (-541) instead of overflowing."""
    L = na.lib()
    L.nufht_asy_grid.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    Lh, Nf = ctypes.c_int(), ctypes.c_int()

    def ok_len(n):
        for f in (1, 3, 9):
            q = n // f
            if n % f == 0 and q & (q - 1) == 0 and (f < 9 or n > 4096):
                return True
        return False

    for X, S, kw in [(500.0, 500.0, 10), (0.5, 1e-3, 4), (12.0, 700.0, 12), (1e4, 1e4, 16)]:
        h = np.pi / (2.0 * S)
        assert L.nufht_asy_grid(X, h, kw, ctypes.byref(Lh), ctypes.byref(Nf)) == 0
        assert Lh.value == int(np.ceil(X / h)) + kw // 2 + 2
        need = max(2 * (2 * Lh.value + 1), 2 * kw + 8)
        assert Nf.value >= need and ok_len(Nf.value)
        # the next smaller admissible length would not hold the grid
        assert not any(ok_len(v) for v in range(need, Nf.value))
    for X, S in [(1e6, 1e6), (3e5, 1e6), (1e9, 1.0)]:
        assert L.nufht_asy_grid(X, np.pi / (2.0 * S), 10, ctypes.byref(Lh),
                                ctypes.byref(Nf)) == -541
        assert "above the supported" in na.last_error()
