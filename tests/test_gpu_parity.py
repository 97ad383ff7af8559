"""Parity of the HIP path against the oracle (FP64 direct sums, closed forms
and the reference's recorded notebook output).  Marked gpu."""
import json
import os

import numpy as np
import pytest

import oracle
from nufht_amd import workloads as W
from tests.conftest import GOLDEN, rel_l2, say

pytestmark = pytest.mark.gpu


def _plan(nu, r, k, tol, precision="fp64"):
    import nufht_amd as na
    return na.Plan(nu, r, k, tol=tol, precision=precision)


def _fwd(plan, cs, device):
    import torch
    c = torch.as_tensor(np.asarray(cs, dtype=np.float64).reshape(plan.m, -1), device=device)
    g = plan.forward(c)
    torch.cuda.synchronize()
    return g.cpu().numpy()


def _uniform(seed, m, n, R, K, B):
    s = W.Stream(seed, 501)
    r = np.sort(s.uniform(m, 0.0, R))
    k = np.sort(s.uniform(n, 0.0, K))
    cs = s.normal(m * B).reshape(m, B)
    return r, k, cs


def test_gaussian_gl_closed_form_and_oracle(device):
    """GL10000 on [0, 50], nu=0, B=4: column 0 against exp(-k^2/2), all
    columns against the FP64 direct sum (test_nufht / notebook shape)."""
    w = W.gaussian_gl()
    say("gaussian_gl nu=0 M=10000 N=200 B=4 tol=1e-8")
    p = _plan(0.0, w["r"], w["k"], 1e-8)
    g = _fwd(p, w["cs"], device)
    closed = np.exp(-0.5 * w["k"] ** 2)
    # tol = 1e-8 on an O(1) output: require an order of magnitude better
    assert np.max(np.abs(g[:, 0] - closed)) < 1e-9
    ref = oracle.direct(0.0, w["r"], w["cs"], w["k"])
    assert rel_l2(g, ref) < 1e-8
    p.close()


def test_reference_notebook_output(device):
    """The reference's only recorded nufht_batch output
    (notebooks/nufht_batch_example.ipynb:148-218): GL1000 on [0,10], nu=0,
    B=2 (scales 1, 0.5): max abs error 9.96e-14 / 4.98e-14 vs exp(-k^2/2)."""
    nb = json.load(open(os.path.join(GOLDEN, "ref_notebook.json")))
    x, wq = W.gauss_legendre(1000)
    r = 5.0 * (x + 1.0)
    wq = 5.0 * wq
    k = 10.0 * np.arange(1, 101) / 100.0
    base = r * np.exp(-0.5 * r * r) * wq
    cs = np.stack([base, 0.5 * base], axis=1)
    say("notebook case nu=0 GL1000 B=2")
    p = _plan(0.0, r, k, 1e-8)
    g = _fwd(p, cs, device)
    analytic = np.exp(-0.5 * k ** 2)[:, None] * np.array([1.0, 0.5])[None, :]
    err = np.max(np.abs(g - analytic), axis=0)
    ref_err = np.array(nb["max_abs_err_per_column"])
    assert np.all(err <= 2.0 * ref_err), (err, ref_err)
    assert np.allclose(g[:5], np.array(nb["g_first5"]), atol=5e-9)
    p.close()


CASES = [
    # nu, m, n, R, K, B, tol
    (0.0, 3000, 2500, 100.0, 100.0, 3, 1e-8),
    (1.0, 4000, 3000, 300.0, 200.0, 2, 1e-8),
    (2.0, 2500, 4000, 50.0, 400.0, 4, 1e-10),
    (5.0, 4000, 3000, 300.0, 200.0, 16, 1e-8),
    (-3.0, 2000, 2000, 200.0, 100.0, 2, 1e-8),
    (10.0, 3000, 3000, 200.0, 200.0, 1, 1e-10),
    (0.5, 2000, 1500, 60.0, 60.0, 2, 1e-8),
    (1.5, 2500, 2000, 100.0, 100.0, 3, 1e-8),
    (2.5, 2000, 2500, 80.0, 120.0, 2, 1e-10),
    (3.0, 50, 100, 10.0, 10.0, 2, 1e-8),        # m*n < min_dim_prod: all direct
]


@pytest.mark.parametrize("nu,m,n,R,K,B,tol", CASES)
def test_random_against_direct_sum(device, nu, m, n, R, K, B, tol):
    r, k, cs = _uniform(int(abs(nu) * 10 + m), m, n, R, K, B)
    say(f"random nu={nu} m={m} n={n} R={R} K={K} B={B} tol={tol}")
    p = _plan(nu, r, k, tol)
    g = _fwd(p, cs, device)
    ref = oracle.direct(nu, r, cs, k)
    err = rel_l2(g, ref)
    assert err < 2 * tol, err
    p.close()


def test_mixed_precision_matches_tolerance(device):
    r, k, cs = _uniform(77, 5000, 5000, 500.0, 100.0, 4)
    say("mixed precision nu=1 tol=1e-6")
    p = _plan(1.0, r, k, 1e-6, precision="mixed")
    assert p.precision == "mixed"
    g = _fwd(p, cs, device)
    ref = oracle.direct(1.0, r, cs, k)
    assert rel_l2(g, ref) < 2e-6
    p.close()


def test_unsorted_inputs_are_permuted(device):
    r, k, cs = _uniform(5, 3000, 2000, 100.0, 100.0, 2)
    s = W.Stream(9, 9)
    pr = np.argsort(s.uniform(r.size))
    pk = np.argsort(s.uniform(k.size))
    say("unsorted nu=2")
    p = _plan(2.0, r[pr], k[pk], 1e-8)
    g = _fwd(p, cs[pr], device)
    ref = oracle.direct(2.0, r, cs, k)
    assert rel_l2(g, ref[pk]) < 2e-8
    p.close()


DENSE_CASES = [
    # nu, m, n, R, K, B, tol, precision: wide r with narrow k (and the reverse)
    # makes source-dense and target-dense asymptotic boxes (matrix-core paths)
    (0.0, 40000, 30000, 1000.0, 40.0, 16, 1e-8, "fp64"),
    (5.0, 30000, 40000, 40.0, 1000.0, 16, 1e-8, "fp64"),
    (1.0, 40000, 40000, 500.0, 200.0, 8, 1e-10, "fp64"),
    (2.0, 30000, 30000, 800.0, 60.0, 3, 1e-8, "fp64"),
    (1.0, 40000, 30000, 1000.0, 50.0, 16, 1e-6, "mixed"),
    (3.0, 30000, 30000, 300.0, 300.0, 5, 1e-6, "mixed"),
    # kernel width 13 (tol 1e-11): the widest the matrix-core spread takes
    (2.0, 30000, 30000, 800.0, 60.0, 4, 1e-11, "fp64"),
    # more than 16 columns: several column chunks, the last one partial
    (1.0, 30000, 30000, 1000.0, 50.0, 20, 1e-8, "fp64"),
    (2.0, 30000, 30000, 800.0, 60.0, 33, 1e-6, "mixed"),
]


@pytest.mark.parametrize("nu,m,n,R,K,B,tol,prec", DENSE_CASES)
def test_dense_boxes_against_direct_sum(device, nu, m, n, R, K, B, tol, prec):
    r, k, cs = _uniform(int(m / 1000 + K), m, n, R, K, B)
    say(f"dense nu={nu} m={m} n={n} R={R} K={K} B={B} tol={tol} {prec}")
    p = _plan(nu, r, k, tol, precision=prec)
    g = _fwd(p, cs, device)
    rows = np.arange(0, n, max(1, n // 400))
    ref = oracle.direct_rows(nu, r, cs, k, rows)
    err = rel_l2(g[rows], ref)
    assert err < 2 * tol, err
    p.close()
