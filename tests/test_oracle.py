"""Pins of the oracle (CPU, no GPU): Bessel values against mpmath vectors,
identities and closed forms; plan tables against the reference's own cache;
boxes against the reference planner; the CPU NUFFT against brute-force
sums; the ref-algorithm port against the direct sum and closed forms."""
import json
import os

import numpy as np
import pytest

import oracle
from nufht_amd import workloads as W
from tests.conftest import GOLDEN, rel_l2


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


# ---------------------------------------------------------------- Bessel
def test_bessel_against_mpmath_vectors():
    d = np.load(os.path.join(GOLDEN, "bessel_golden.npz"))
    nu, x, j = d["nu"], d["x"], d["j"]
    got = np.array([oracle.lib().orc_besselj(float(a), float(b)) for a, b in zip(nu, x)])
    # error relative to the local amplitude max(|J|, sqrt(2/(pi x)) for x > nu)
    env = np.maximum(np.abs(j), np.where(x > nu, np.sqrt(2 / (np.pi * np.maximum(x, 1e-300))), 0))
    err = np.abs(got - j) / np.maximum(env, 1e-300)
    small = nu <= 10
    assert np.max(err[small]) < 5e-15
    assert np.max(err) < 5e-14


def test_bessel_recurrence_and_closed_forms():
    x = np.concatenate([np.linspace(0.05, 40, 400), np.logspace(2, 6, 50)])
    for n in [1, 2, 5, 9]:
        lhs = oracle.besselj(n - 1, x) + oracle.besselj(n + 1, x)
        rhs = 2 * n / x * oracle.besselj(n, x)
        assert np.max(np.abs(lhs - rhs)) < 1e-13
    j12 = oracle.besselj(0.5, x)
    assert np.max(np.abs(j12 - np.sqrt(2 / (np.pi * x)) * np.sin(x))) < 1e-14
    j32 = oracle.besselj(1.5, x)
    ref = np.sqrt(2 / (np.pi * x)) * (np.sin(x) / x - np.cos(x))
    assert np.max(np.abs(j32 - ref)) < 1e-14


def test_bessel_two_product_phase():
    """J_0(w r) for w*r not representable: the two-product keeps the phase.
    Checked against J_0 of the exact product evaluated in higher precision."""
    import mpmath
    mpmath.mp.dps = 40
    s = W.Stream(3, 3)
    w = s.uniform(20, 500.0, 1000.0)
    r = s.uniform(20, 500.0, 1000.0)
    for a, b in zip(w, r):
        exact = float(mpmath.besselj(0, mpmath.mpf(float(a)) * mpmath.mpf(float(b))))
        got = oracle.lib().orc_besselj_wr(0.0, float(a), float(b))
        assert abs(got - exact) < 1e-15 * 50


# ---------------------------------------------------------------- tables
def test_tables_against_reference_cache():
    d = np.load(os.path.join(GOLDEN, "ref_tables_slice.npz"))
    a, z, k = oracle.tables()
    assert list(d["header"]) == [0x4E5546, 1, 401, 401, 32, 16]
    assert np.array_equal(a[:81], d["asy_coef"])
    ref_z, ref_k = d["asy_z"][:, :, :10], d["wimp_k"][:, :, :10]
    np.testing.assert_array_equal(k[:41, :12, :10], ref_k)
    assert np.max(np.abs(z[:41, :12, :10] - ref_z) / ref_z) < 5e-11
    # the reference leaves slot K = 10 empty (src/bounds.c:202)
    assert np.all(d["asy_z"][:, :, 10] == 0)


def test_hankel_coefficients_closed_form():
    L = oracle.lib()
    for nu in [0.0, 1.0, 2.5, 7.0]:
        assert L.orc_a_coef(0, nu) == 1.0
        assert abs(L.orc_a_coef(1, nu) - (4 * nu * nu - 1) / 8) < 1e-12 * max(1, nu * nu)
        assert abs(L.orc_a_coef(2, nu) - (4 * nu * nu - 1) * (4 * nu * nu - 9) / 128) < 1e-10 * max(1, nu ** 4)
    # the series terminates for half-integer orders: a_k(n+1/2) = 0 for k > n
    for n in range(5):
        for kk in range(n + 1, n + 4):
            assert L.orc_a_coef(kk, n + 0.5) == 0.0


# ---------------------------------------------------------------- planner
def _case_grids(c):
    gen = c["gen"]
    if gen[0] == "gaussian_gl":
        g = W.gaussian_gl()
        return g["r"], g["k"]
    s = W.Stream(gen[1], 77)
    if gen[0] == "uni":
        return np.sort(s.uniform(gen[2], 0, gen[4])), np.sort(s.uniform(gen[3], 0, gen[5]))
    return np.sort(s.uniform(gen[2], 100.0, 200.0)), np.sort(s.uniform(gen[3], 100.0, 200.0))


def test_oracle_planner_matches_reference_boxes():
    d = json.load(open(os.path.join(GOLDEN, "ref_boxes.json")))
    for c in d["cases"]:
        r, k = _case_grids(c)
        rc, p = oracle.params(c["nu"], c["tol"])
        assert rc == c["rc"]
        if c["name"] == "uni_nu40":
            assert p["z_split"] == 0.0 and p["K_loc"] == 0   # the reference's table hole
        assert abs(p["z_split"] - c["z_split"]) <= 3e-11 * abs(c["z_split"]) or \
            (np.isnan(p["z_split"]) and np.isnan(c["z_split"]))
        assert (p["K_asy"], p["K_loc"]) == (c["K_asy"], c["K_loc"])
        bx = oracle.boxes(r, k, c["z_split"], K_loc=max(0, c["K_loc"]))
        for t in ("loc", "asy", "dir"):
            np.testing.assert_array_equal(bx[t], np.array(c["boxes"][t]).reshape(-1, 4))
    for e in d["params"]:
        rc, p = oracle.params(e["nu"], e["tol"])
        assert rc == e["rc"], e
        if rc == 0:
            assert (p["K_asy"], p["K_loc"]) == (e["K_asy"], e["K_loc"])


def test_box_invariants():
    """Exact disjoint tiling; loc boxes have w_max r_max <= z; asy boxes
    w_min r_min > z (the regime rule of src/fast_hankel_nufht.c:578-744)."""
    s = W.Stream(8, 8)
    r = np.sort(s.uniform(3000, 0, 400.0))
    k = np.sort(s.uniform(2000, 0, 300.0))
    z = 11.0
    bx = oracle.boxes(r, k, z, K_loc=14)
    cover = np.zeros((k.size, r.size), dtype=np.int8)
    for t in ("loc", "asy", "dir"):
        for i0, i1, j0, j1 in bx[t]:
            cover[i0:i1 + 1, j0:j1 + 1] += 1
    assert np.all(cover == 1)
    for i0, i1, j0, j1 in bx["loc"]:
        assert k[i1] * r[j1] <= z
    for i0, i1, j0, j1 in bx["asy"]:
        assert k[i0] * r[j0] > z


def test_live_reference_planner_when_built():
    L = oracle.load_ref_planner()
    cache = "/root/reference/data/nufht_tables.cache"
    if L is None or not os.path.exists(cache):
        pytest.skip("reference planner not built here")
    import ctypes
    os.environ["NUFHT_TABLE_CACHE"] = cache
    s = W.Stream(31, 31)
    for nu, tol, R, K in [(0.0, 1e-8, 200.0, 150.0), (3.0, 1e-10, 90.0, 300.0), (1.0, 1e-6, 500.0, 60.0)]:
        r = np.sort(s.uniform(4000, 0, R))
        k = np.sort(s.uniform(3500, 0, K))
        opt = L.nufht_default_options()
        opt.tol = tol
        plan = oracle.RefPlan()
        rc = L.nufht_plan_init(ctypes.byref(plan), nu, r.ctypes.data_as(oracle._DP), r.size,
                               k.ctypes.data_as(oracle._DP), k.size, ctypes.byref(opt))
        assert rc == 0
        ref = oracle.ref_boxes_to_arrays(plan.boxes)
        mine = oracle.boxes(r, k, plan.z_split, K_loc=plan.K_loc)
        for t in ("loc", "asy", "dir"):
            np.testing.assert_array_equal(mine[t], ref[t])
        L.nufht_plan_free(ctypes.byref(plan))


# ---------------------------------------------------------------- NUFFT
@pytest.mark.parametrize("X,S,eps", [(50, 30, 1e-6), (200, 10, 1e-8), (1.0, 300, 1e-10),
                                     (1e-3, 1e3, 1e-8), (400, 400, 1e-10), (30, 30, 1e-12)])
def test_cpu_nufft_against_brute_force(X, S, eps):
    """FP64 phases round to ~1.1e-16 * X * S, so eps is only reachable above
    that floor (X * S = 1.6e5 -> 2e-11)."""
    s = W.Stream(int(X * 7 + S), 4)
    x = np.sort(s.uniform(500, 100.0, 100.0 + 2 * X))
    w = np.sort(s.uniform(400, 5.0, 5.0 + 2 * S))
    c = s.normal(1000).reshape(2, 500) + 1j * s.normal(1000).reshape(2, 500)
    for iflag in (1, -1):
        f = oracle.nufft3(x, c, w, iflag, eps)
        fe = oracle.nufft3(x, c, w, iflag, exact=True)
        assert rel_l2(f, fe) < eps


# ---------------------------------------------------------------- port
def test_port_gaussian_closed_form():
    g = W.gaussian_gl(n_gl=2000, rmax=12.0, nk=60)
    out = oracle.nufht_batch(0.0, g["r"], g["cs"][:, :1], g["k"], tol=1e-10, asy_mode=1)
    assert np.max(np.abs(out[:, 0] - np.exp(-0.5 * g["k"] ** 2))) < 1e-10


@pytest.mark.parametrize("nu,tol", [(0.0, 1e-8), (1.0, 1e-10), (4.0, 1e-8), (-3.0, 1e-8)])
def test_port_against_direct_sum(nu, tol):
    s = W.Stream(int(abs(nu) * 3 + 1), 6)
    r = np.sort(s.uniform(1500, 0.0, 120.0))
    k = np.sort(s.uniform(1200, 0.0, 90.0))
    cs = s.normal(1500 * 2).reshape(1500, 2)
    ref = oracle.direct(nu, r, cs, k)
    exact_sums = oracle.nufht_batch(nu, r, cs, k, tol=tol, asy_mode=0)
    fast = oracle.nufht_batch(nu, r, cs, k, tol=tol, asy_mode=1)
    assert rel_l2(exact_sums, ref) < tol
    assert rel_l2(fast, ref) < 2 * tol


@pytest.mark.parametrize("plant,what", [(1, "odd Hankel terms' sign flipped"),
                                         (2, "odd terms' source weight exponent off by one"),
                                         (3, "Wimp expansion's alternating sign dropped"),
                                         (4, "Hankel phase's quarter-turn sign flipped")])
def test_port_planted_mistake_is_caught(plant, what):
    """A plausible one-symbol mistake planted in the oracle's restatement of
    the reference's expansions (src/expansions.c:549-557, :523-531, :301-318)
    must fail the direct-sum comparison the oracle is pinned by, while the
    same case passes without it."""
    lib = oracle.lib()
    s = W.Stream(12, 6)
    r = np.sort(s.uniform(3000, 0.0, 100.0))
    k = np.sort(s.uniform(2500, 0.0, 100.0))
    cs = s.normal(3000)[:, None]
    prm = oracle.params(4.0)[1]
    bx = oracle.boxes(r, k, prm["z_split"], K_loc=prm["K_loc"])
    assert len(bx["loc"]) and len(bx["asy"])      # both expansions contribute
    ref = oracle.direct(4.0, r, cs, k)
    good = oracle.nufht_batch(4.0, r, cs, k, tol=1e-8, asy_mode=0)
    assert rel_l2(good, ref) < 2e-8
    lib.orc_set_plant(plant)
    try:
        bad = oracle.nufht_batch(4.0, r, cs, k, tol=1e-8, asy_mode=0)
    finally:
        lib.orc_set_plant(0)
    assert rel_l2(bad, ref) > 1e-6, what
