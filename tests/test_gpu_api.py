"""GPU tests of the rest of the surface: the drop-in _pynufht entry points,
the host-pointer C ABI, the adjoint, node gradients against central finite
differences, uneven multi-transform batches and hipGraph replays."""
import numpy as np
import pytest

import oracle
from nufht_amd import workloads as W
from tests.conftest import rel_l2, say

pytestmark = pytest.mark.gpu


def _uniform(seed, m, n, R, K, B):
    s = W.Stream(seed, 601)
    return (np.sort(s.uniform(m, 0.0, R)), np.sort(s.uniform(n, 0.0, K)),
            s.normal(m * B).reshape(m, B))


def test_pynufht_entry_points(device):
    import nufht_amd as na
    pn = na.load_pynufht()
    r, k, cs = _uniform(1, 3000, 2000, 150.0, 100.0, 3)
    ref = oracle.direct(1.0, r, cs, k)
    say("_pynufht.nufht_batch nu=1")
    g = pn.nufht_batch(1.0, r, cs, k, tol=1e-8)
    assert g.shape == (2000, 3) and g.flags.c_contiguous
    assert rel_l2(g, ref) < 2e-8
    say("_pynufht.nufht_batch_loop nu=1")
    gl = pn.nufht_batch_loop(1.0, r, np.ascontiguousarray(cs), k, tol=1e-8)
    assert rel_l2(gl, ref) < 2e-8
    say("_pynufht.nufht nu=1")
    g1 = pn.nufht(1.0, r, np.ascontiguousarray(cs[:, 0]), k, tol=1e-8)
    assert g1.shape == (2000,)
    assert rel_l2(g1, ref[:, 0]) < 2e-8


def test_host_pointer_batch_with_leading_dimensions(device):
    import ctypes
    import nufht_amd as na
    r, k, cs = _uniform(2, 2500, 1800, 120.0, 90.0, 2)
    m, n = r.size, k.size
    p = na.Plan(3.0, r, k, tol=1e-9)
    ldc, ldg = m + 7, n + 5
    cbuf = np.zeros(ldc * 2)
    for t in range(2):
        cbuf[t * ldc:t * ldc + m] = cs[:, t]
    gbuf = np.full(ldg * 2, 123.0)
    say("host nufht_batch nu=3 with ld_cs > m, ld_gs > n")
    rc = na.lib().nufht_batch(ctypes.byref(p._plan), cbuf.ctypes.data_as(na._DP), ldc,
                              gbuf.ctypes.data_as(na._DP), ldg, 2, None)
    assert rc == 0
    g = np.stack([gbuf[t * ldg:t * ldg + n] for t in range(2)], axis=1)
    assert rel_l2(g, oracle.direct(3.0, r, cs, k)) < 2e-9
    assert np.all(gbuf[n:ldg] == 123.0)      # padding between columns untouched
    p.close()


@pytest.mark.parametrize("nu", [0.0, 2.0, -1.0, 1.5])
def test_adjoint_against_swapped_direct_sum(device, nu):
    import torch
    import nufht_amd as na
    r, k, _ = _uniform(3, 3000, 2500, 150.0, 120.0, 1)
    y = W.Stream(4, 4).normal(2500 * 3).reshape(2500, 3)
    p = na.Plan(nu, r, k, tol=1e-9)
    say(f"adjoint nu={nu}")
    x = p.adjoint(torch.as_tensor(y, device=device))
    torch.cuda.synchronize()
    ref = oracle.direct(nu, k, y, r)      # sum_i y_i J_nu(k_i r_j): r and k swapped
    assert rel_l2(x.cpu().numpy(), ref) < 3e-9
    # <H c, y> == <c, H^T y>
    c = W.Stream(5, 5).normal(3000 * 3).reshape(3000, 3)
    g = p.forward(torch.as_tensor(c, device=device)).cpu().numpy()
    lhs, rhs = float(np.sum(g * y)), float(np.sum(c * x.cpu().numpy()))
    assert abs(lhs - rhs) < 1e-8 * (abs(lhs) + np.linalg.norm(g) * np.linalg.norm(y))
    p.close()


def _loss_fd_r(nu, r, k, c, R, j, h):
    """Central difference of L = ||H(c; r) - y||^2 in r_j with exact column
    deltas: changing r_j changes only column j of H."""
    J0 = oracle.besselj_wr(abs(nu), k, np.full_like(k, r[j]))
    Jp = oracle.besselj_wr(abs(nu), k, np.full_like(k, r[j] + h))
    Jm = oracle.besselj_wr(abs(nu), k, np.full_like(k, r[j] - h))
    sgn = -1.0 if (nu < 0 and int(round(nu)) % 2) else 1.0
    dp = sgn * np.outer(Jp - J0, c[j])
    dm = sgn * np.outer(Jm - J0, c[j])
    return (2 * np.sum(R * (dp - dm)) + np.sum(dp * dp) - np.sum(dm * dm)) / (2 * h)


def _exact_dk(nu, r, c, Rrow, kval):
    """dL/dk_i from FP64 direct sums: sum_t 2 R[i,t] sum_j c[j,t] r_j J'_nu(k_i r_j),
    J'_nu(x) = (nu/x) J_nu(x) - J_{nu+1}(x)."""
    x = kval * r
    Jn = oracle.besselj_wr(abs(nu), np.full_like(r, kval), r)
    Jn1 = oracle.besselj_wr(abs(nu) + 1, np.full_like(r, kval), r)
    dJ = np.where(x > 0, abs(nu) / np.where(x > 0, x, 1.0) * Jn, 0.0) - Jn1
    sgn = -1.0 if (nu < 0 and int(round(nu)) % 2) else 1.0
    return sgn * np.dot(2 * Rrow, (r * dJ) @ c)


@pytest.mark.parametrize("nu", [1.0, 0.0, 3.0])
def test_node_gradients_against_finite_differences(device, nu):
    """Gradient workload shape (r, k ~ U[0,300), B=8, tol 1e-10), reduced to
    M=N=2e4: y = H(c*) + 1e-3 N(0,1); L = ||H(c) - y||^2.  Finite differences
    (h = 1e-6, exact column/row deltas) carry ~1e-5 absolute noise from
    Bessel roundoff / h, so they are compared relative to the largest sampled
    gradient; the exact analytic gradient (FP64 direct sums) is compared per
    coordinate."""
    import torch
    from nufht_amd.autograd import NufhtOperator
    wl = W.gradient(m=20000, n=20000)
    r, k, c, cstar = wl["r"], wl["k"], wl["c0"], wl["cstar"]
    op = NufhtOperator(nu, tol=1e-10, precision="fp64")
    rt = torch.tensor(r, device=device, requires_grad=True)
    kt = torch.tensor(k, device=device, requires_grad=True)
    ct = torch.tensor(c, device=device, requires_grad=True)
    say(f"gradient nu={nu} M=N=2e4 B=8 through autograd")
    with torch.no_grad():
        y = op(rt, torch.as_tensor(cstar, device=device), kt) + \
            torch.as_tensor(1e-3 * wl["noise"], device=device)
    H = op(rt, ct, kt)
    loss = ((H - y) ** 2).sum()
    loss.backward()
    R = (H - y).detach().cpu().numpy()
    gr, gk, dc = rt.grad.cpu().numpy(), kt.grad.cpu().numpy(), ct.grad.cpu().numpy()
    # the autograd layer and the plan's own entry point agree
    pr, pk = op._plan.grad_nodes(ct.detach(), torch.as_tensor(2 * R, device=device))
    assert rel_l2(pr.cpu().numpy(), gr) < 1e-13 and rel_l2(pk.cpu().numpy(), gk) < 1e-13
    s = W.Stream(52, 5)
    js = np.unique((s.uniform(50) * r.size).astype(int))
    fd = np.array([_loss_fd_r(nu, r, k, c, R, j, 1e-6) for j in js])
    assert np.max(np.abs(gr[js] - fd)) < 1e-6 * np.max(np.abs(fd))
    # dL/dc is exactly linear: compare with the direct adjoint at the same coords
    jc = js[:20]
    ref_dc = oracle.direct(nu, k, 2 * R, r[jc])
    assert rel_l2(dc[jc], ref_dc) < 1e-9
    # dL/dk: exact row deltas, and the exact analytic value
    iks = np.unique((s.uniform(20) * k.size).astype(int))
    fdk, exk = [], []
    for i in iks:
        h = 1e-6
        J0 = oracle.besselj_wr(abs(nu), np.full_like(r, k[i]), r)
        Jp = oracle.besselj_wr(abs(nu), np.full_like(r, k[i] + h), r)
        Jm = oracle.besselj_wr(abs(nu), np.full_like(r, k[i] - h), r)
        sgn = -1.0 if (nu < 0 and int(round(nu)) % 2) else 1.0
        dp, dm = sgn * (Jp - J0) @ c, sgn * (Jm - J0) @ c
        fdk.append((2 * np.dot(R[i], dp - dm) + dp @ dp - dm @ dm) / (2 * h))
        exk.append(_exact_dk(nu, r, c, R[i], k[i]))
    fdk, exk = np.array(fdk), np.array(exk)
    assert np.max(np.abs(gk[iks] - fdk)) < 1e-6 * np.max(np.abs(fdk))
    assert np.max(np.abs(gk[iks] - exk) / (np.abs(exk) + 1e-6 * np.max(np.abs(exk)))) < 1e-8
    op.close()


def test_uneven_multi_batch_and_graph_replays(device):
    import torch
    import nufht_amd as na
    ts = W.uneven_batch(count=24, mmin=500, mmax=4000)
    plans, cs_d, gs_d, refs = [], [], [], []
    for t in ts:
        p = na.Plan(t["nu"], t["r"], t["k"], tol=1e-10)
        plans.append(p)
        cs_d.append(torch.as_tensor(np.asfortranarray(t["cs"]).T.copy(), device=device))
        gs_d.append(torch.empty((t["cs"].shape[1], t["k"].size), dtype=torch.float64, device=device))
        refs.append(oracle.direct(t["nu"], t["r"], t["cs"], t["k"]))
    say("multi-batch 24 transforms, eager")
    mb = na.MultiBatch(plans, [t["cs"].shape[1] for t in ts])
    mb.execute(cs_d, gs_d, use_graph=False)
    torch.cuda.synchronize()
    for g, ref in zip(gs_d, refs):
        assert rel_l2(g.t().cpu().numpy(), ref) < 5e-10
    eager = [g.clone() for g in gs_d]
    say("multi-batch 24 transforms, graph replays")
    for rep in range(3):
        for g in gs_d:
            g.fill_(7.0)
        mb.execute(cs_d, gs_d, use_graph=True)
        torch.cuda.synchronize()
        for g, e in zip(gs_d, eager):
            assert rel_l2(g.cpu().numpy(), e.cpu().numpy()) < 1e-13
        _ = torch.ones(10, device=device).sum().item()   # eager work between replays
    mb.close()
    for p in plans:
        p.close()


def test_plan_cache_reuses_device_state(device):
    import time
    import nufht_amd as na
    pn = na.load_pynufht()
    r, k, cs = _uniform(3, 20000, 20000, 200.0, 200.0, 2)
    rows = np.arange(0, 20000, 397)
    ref = oracle.direct_rows(2.0, r, cs, k, rows)
    h0 = na.plan_cache()
    say("_pynufht.nufht_batch twice on the same nodes")
    t0 = time.perf_counter()
    g1 = pn.nufht_batch(2.0, r, cs, k, tol=1e-8)
    t1 = time.perf_counter()
    g2 = pn.nufht_batch(2.0, r, cs, k, tol=1e-8)
    t2 = time.perf_counter()
    say(f"  first call {t1 - t0:.3f} s, second {t2 - t1:.3f} s")
    assert na.plan_cache() == h0 + 1
    assert rel_l2(g1[rows], ref) < 2e-8
    assert rel_l2(g2, g1) < 1e-13
    say("other nodes or order: no reuse")
    r2 = r.copy()
    r2[5] = 0.5 * (r2[4] + r2[5])
    g3 = pn.nufht_batch(2.0, r2, cs, k, tol=1e-8)
    g4 = pn.nufht_batch(3.0, r, cs, k, tol=1e-8)
    assert na.plan_cache() == h0 + 1
    assert rel_l2(g3[rows], oracle.direct_rows(2.0, r2, cs, k, rows)) < 2e-8
    assert rel_l2(g4[rows], oracle.direct_rows(3.0, r, cs, k, rows)) < 2e-8
    na.plan_cache(0)
    g5 = pn.nufht_batch(2.0, r, cs, k, tol=1e-8)
    assert na.plan_cache(2) == h0 + 1
    assert rel_l2(g5, g1) < 1e-13
