"""Parity at the benchmark workloads' full sizes, on the paths bench.py
times: all 256 transforms of the uneven batch through the autotuned,
graph-replayed MultiBatch (forward and adjoint, tol 1e-10 and 1e-6), the
gradient workload at M = N = 2e5 through autograd, and the clustered
workload's adjoint in FP64 and in the FP32-mixed mode.  Each output is
checked on sampled rows against the oracle's FP64 direct sums."""
import numpy as np
import pytest

import oracle
from nufht_amd import workloads as W
from tests.conftest import rel_l2, say
from tests.test_gpu_api import _exact_dk, _loss_fd_r

pytestmark = pytest.mark.gpu


def _rows(n, count, seed):
    return np.unique((W.Stream(seed, 9).uniform(count) * n).astype(np.int32))


@pytest.mark.parametrize("tol", [1e-10, 1e-6])
def test_uneven_256_autotuned_graph_forward_and_adjoint(device, tol):
    """bench.py's uneven-batch leg: 256 forward plans and 256 adjoint plans
    (r and k swapped), each set one MultiBatch, autotuned, executed as a
    replayed hipGraph; every transform checked on 24 sampled output rows
    (forward) and 16 sampled output rows (adjoint)."""
    import torch
    import nufht_amd as na
    ts = W.uneven_batch()
    say(f"uneven batch: 256 transforms tol={tol}, autotuned graph forward + adjoint")
    pf = [na.Plan(t["nu"], t["r"], t["k"], tol=tol, precision="fp64") for t in ts]
    pa = [na.Plan(t["nu"], t["k"], t["r"], tol=tol, precision="fp64") for t in ts]
    Bs = [t["cs"].shape[1] for t in ts]
    cs = [torch.as_tensor(np.ascontiguousarray(t["cs"].T), device=device) for t in ts]
    gs = [torch.empty((b, t["k"].size), dtype=torch.float64, device=device) for b, t in zip(Bs, ts)]
    xs = [torch.empty((b, t["r"].size), dtype=torch.float64, device=device) for b, t in zip(Bs, ts)]
    mf, ma = na.MultiBatch(pf, Bs), na.MultiBatch(pa, Bs)
    knobs = (mf.autotune(), ma.autotune())
    say(f"  knobs forward {knobs[0]}, adjoint {knobs[1]}")
    for rep in range(2):
        for g in gs + xs:
            g.fill_(3.0)
        mf.execute(cs, gs, use_graph=True)
        ma.execute(gs, xs, use_graph=True)
        torch.cuda.synchronize()
        if rep == 0:
            g0 = [g.clone() for g in gs]
            x0 = [x.clone() for x in xs]
    for a, b in zip(gs + xs, g0 + x0):
        assert torch.equal(a, b) or rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1e-13
    worst_f = worst_a = 0.0
    for q, (t, g, x) in enumerate(zip(ts, gs, xs)):
        rows = _rows(t["k"].size, 24, 1000 + q)
        ref = oracle.direct_rows(t["nu"], t["r"], t["cs"], t["k"], rows)
        ef = rel_l2(g.t().cpu().numpy()[rows], ref)
        gh = np.ascontiguousarray(g.t().cpu().numpy())
        cols = _rows(t["r"].size, 16, 2000 + q)
        ea = rel_l2(x.t().cpu().numpy()[cols], oracle.direct_rows(t["nu"], t["k"], gh, t["r"], cols))
        worst_f, worst_a = max(worst_f, ef), max(worst_a, ea)
        assert ef < 3 * tol and ea < 3 * tol, (q, t["nu"], t["r"].size, t["k"].size, ef, ea)
    say(f"  worst rel l2: forward {worst_f:.2e}, adjoint {worst_a:.2e}")
    mf.close()
    ma.close()
    for p in pf + pa:
        p.close()


def test_gradient_workload_full_size(device):
    """nu = 1, M = N = 2e5, r, k ~ U[0,300), 8 columns, tol 1e-10:
    L = ||H(c; r, k) - y||^2 through autograd.  dL/dr on 50 coordinates and
    dL/dk on 20 against central finite differences (h = 1e-6, exact
    column/row deltas) to 1e-6 of the largest, dL/dk also against its exact
    value from FP64 direct sums, dL/dc on 20 coordinates against the direct
    adjoint."""
    import torch
    from nufht_amd.autograd import NufhtOperator
    nu = 1.0
    wl = W.gradient()
    r, k, c, cstar = wl["r"], wl["k"], wl["c0"], wl["cstar"]
    op = NufhtOperator(nu, tol=1e-10, precision="fp64")
    rt = torch.tensor(r, device=device, requires_grad=True)
    kt = torch.tensor(k, device=device, requires_grad=True)
    ct = torch.tensor(c, device=device, requires_grad=True)
    say("gradient workload nu=1 M=N=2e5 B=8 tol=1e-10 through autograd")
    with torch.no_grad():
        y = op(rt, torch.as_tensor(cstar, device=device), kt) + \
            torch.as_tensor(1e-3 * wl["noise"], device=device)
    H = op(rt, ct, kt)
    loss = ((H - y) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    R = (H - y).detach().cpu().numpy()
    gr, gk, dc = rt.grad.cpu().numpy(), kt.grad.cpu().numpy(), ct.grad.cpu().numpy()
    s = W.Stream(52, 5)
    js = np.unique((s.uniform(50) * r.size).astype(int))
    assert js.size >= 45
    fd = np.array([_loss_fd_r(nu, r, k, c, R, j, 1e-6) for j in js])
    err_r = np.max(np.abs(gr[js] - fd)) / np.max(np.abs(fd))
    jc = js[:20]
    err_c = rel_l2(dc[jc], oracle.direct(nu, k, 2 * R, r[jc]))
    iks = np.unique((s.uniform(20) * k.size).astype(int))
    fdk, exk = [], []
    for i in iks:
        h = 1e-6
        J0 = oracle.besselj_wr(nu, np.full_like(r, k[i]), r)
        Jp = oracle.besselj_wr(nu, np.full_like(r, k[i] + h), r)
        Jm = oracle.besselj_wr(nu, np.full_like(r, k[i] - h), r)
        dp, dm = (Jp - J0) @ c, (Jm - J0) @ c
        fdk.append((2 * np.dot(R[i], dp - dm) + dp @ dp - dm @ dm) / (2 * h))
        exk.append(_exact_dk(nu, r, c, R[i], k[i]))
    fdk, exk = np.array(fdk), np.array(exk)
    err_k = np.max(np.abs(gk[iks] - fdk)) / np.max(np.abs(fdk))
    err_kx = np.max(np.abs(gk[iks] - exk) / (np.abs(exk) + 1e-6 * np.max(np.abs(exk))))
    say(f"  dL/dr vs FD {err_r:.2e}; dL/dk vs FD {err_k:.2e}, vs exact {err_kx:.2e}; "
        f"dL/dc vs direct {err_c:.2e}")
    assert err_r < 1e-6 and err_k < 1e-6 and err_kx < 1e-8 and err_c < 1e-9
    op.close()


def test_clustered_adjoint_fp64_and_mixed(device):
    """The clustered tol 1e-6 workload's adjoint x = H^T y (y = the FP64
    forward output, as bench.py's step feeds it) in both precisions, on 24
    sampled output rows against the oracle's direct sums."""
    import torch
    import nufht_amd as na
    wl = W.clustered()
    r, k, cs = wl["r"], wl["k"], wl["cs"]
    c = torch.as_tensor(np.ascontiguousarray(cs.T), device=device).t()
    p64 = na.Plan(1.0, r, k, tol=1e-6, precision="fp64")
    y = p64.forward(c)
    torch.cuda.synchronize()
    yh = np.ascontiguousarray(y.cpu().numpy())
    cols = _rows(r.size, 24, 77)
    ref = oracle.direct_rows(1.0, k, yh, r, cols)
    errs = {}
    for prec in ("fp64", "mixed"):
        say(f"clustered adjoint nu=1 M=N=1e6 32 columns tol=1e-6 {prec}")
        p = p64 if prec == "fp64" else na.Plan(1.0, r, k, tol=1e-6, precision=prec)
        x = p.adjoint(y)
        torch.cuda.synchronize()
        errs[prec] = rel_l2(x.cpu().numpy()[cols], ref)
        if p is not p64:
            p.close()
    say(f"  adjoint rel l2: fp64 {errs['fp64']:.2e}, mixed {errs['mixed']:.2e}")
    assert errs["fp64"] < 1e-6 and errs["mixed"] < 1e-6
    p64.close()
