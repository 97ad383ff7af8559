"""Orders and inputs outside the benchmark workloads: high integer orders
(where the local expansion needs more terms than the local kernels hold and
its boxes are summed directly, and the K_asy = 10 table slot the reference
leaves empty), negative odd orders with a single column, multi-transform
batches over unsorted grids, and autograd graphs over two node sets."""
import numpy as np
import pytest

import oracle
from nufht_amd import workloads as W
from tests.conftest import rel_l2, say

pytestmark = pytest.mark.gpu


def _uniform(seed, m, n, R, K, B):
    s = W.Stream(seed, 811)
    return (np.sort(s.uniform(m, 0.0, R)), np.sort(s.uniform(n, 0.0, K)),
            s.normal(m * B).reshape(m, B))


@pytest.mark.parametrize("nu,tol", [(40.0, 1e-8), (40.0, 1e-10), (60.0, 1e-8), (27.0, 1e-8)])
def test_high_orders(device, nu, tol):
    import torch
    import nufht_amd as na
    r, k, cs = _uniform(int(nu), 100000, 100000, 100.0, 100.0, 2)
    p = na.Plan(nu, r, k, tol=tol, precision="fp64")
    st = p.stats(2)
    say(f"nu={nu} tol={tol} M=N=1e5 on [0,100]^2: K_loc={p.K_loc} K_asy={p.K_asy} "
        f"boxes loc/asy/dir {st['n_loc']}/{st['n_asy']}/{st['n_dir']}")
    g = p.forward(torch.as_tensor(np.ascontiguousarray(cs.T), device=device).t())
    torch.cuda.synchronize()
    rows = np.unique((W.Stream(int(nu), 12).uniform(48) * k.size).astype(np.int32))
    # include the rows of the smallest k, where J_nu(k r) is tiny everywhere
    rows = np.union1d(rows, np.arange(0, 100000, 4001, dtype=np.int32))
    ref = oracle.direct_rows(nu, r, cs, k, rows)
    err = rel_l2(g.cpu().numpy()[rows], ref)
    say(f"  rel l2 {err:.2e}")
    assert err < 2 * tol
    p.close()


def test_negative_odd_order_single_column(device):
    """nu = -3 with one column: J_{-3} = -J_3, applied once (the reference
    negates twice on its single-column path)."""
    import torch
    import nufht_amd as na
    r, k, cs = _uniform(3, 4000, 3000, 150.0, 120.0, 1)
    ref = oracle.direct(-3.0, r, cs, k)
    ref3 = oracle.direct(3.0, r, cs, k)
    assert rel_l2(ref, -ref3) < 1e-14
    p = na.Plan(-3.0, r, k, tol=1e-9)
    say("nu=-3 B=1: device, host nufht_batch, _pynufht.nufht")
    g = p.forward(torch.as_tensor(cs, device=device)).cpu().numpy()
    assert rel_l2(g, ref) < 2e-9
    assert rel_l2(p.forward(cs), ref) < 2e-9
    pn = na.load_pynufht()
    g1 = pn.nufht(-3.0, r, np.ascontiguousarray(cs[:, 0]), k, tol=1e-9)
    assert rel_l2(g1, ref[:, 0]) < 2e-9
    p.close()


def test_multi_batch_unsorted_grids(device):
    """A multi-transform batch whose transforms have unsorted nodes: rows are
    gathered into sorted order and scattered back inside the launch sequence
    (eager and replayed graph)."""
    import torch
    import nufht_amd as na
    ts = W.uneven_batch(count=8, mmin=600, mmax=5000)
    plans, cs_d, gs_d, refs = [], [], [], []
    for q, t in enumerate(ts):
        s = W.Stream(900 + q, 3)
        pr = np.argsort(s.uniform(t["r"].size)) if q % 2 == 0 else np.arange(t["r"].size)
        pk = np.argsort(s.uniform(t["k"].size)) if q % 3 != 2 else np.arange(t["k"].size)
        r, k, c = t["r"][pr], t["k"][pk], t["cs"][pr]
        plans.append(na.Plan(t["nu"], r, k, tol=1e-9))
        cs_d.append(torch.as_tensor(np.ascontiguousarray(c.T), device=device))
        gs_d.append(torch.empty((c.shape[1], k.size), dtype=torch.float64, device=device))
        refs.append(oracle.direct(t["nu"], r, c, k))
    mb = na.MultiBatch(plans, [c.shape[0] for c in cs_d])
    for use_graph in (False, True, True):
        say(f"multi-batch 8 transforms, unsorted grids, graph={use_graph}")
        for g in gs_d:
            g.fill_(5.0)
        mb.execute(cs_d, gs_d, use_graph=use_graph)
        torch.cuda.synchronize()
        for g, ref in zip(gs_d, refs):
            assert rel_l2(g.t().cpu().numpy(), ref) < 3e-9
    mb.close()
    for p in plans:
        p.close()


def test_autograd_two_node_sets_in_one_graph(device):
    """loss = ||op(r1, c, k)||^2 + ||op(r2, c, k)||^2: the operator re-plans for
    r2 while the first node still needs the r1 plan in backward."""
    import torch
    from nufht_amd.autograd import NufhtOperator
    r1, k, cs = _uniform(5, 3000, 2000, 100.0, 80.0, 2)
    r2 = np.sort(W.Stream(6, 1).uniform(3000, 0.0, 90.0))
    op = NufhtOperator(2.0, tol=1e-10, precision="fp64")
    say("autograd: two node sets in one graph")
    ct = torch.tensor(cs, device=device, requires_grad=True)
    rt1 = torch.tensor(r1, device=device, requires_grad=True)
    rt2 = torch.tensor(r2, device=device, requires_grad=True)
    kt = torch.tensor(k, device=device)
    loss = (op(rt1, ct, kt) ** 2).sum() + (op(rt2, ct, kt) ** 2).sum()
    loss.backward()
    g1 = oracle.direct(2.0, r1, cs, k)
    g2 = oracle.direct(2.0, r2, cs, k)
    ref_dc = oracle.direct(2.0, k, 2 * g1, r1) + oracle.direct(2.0, k, 2 * g2, r2)
    assert rel_l2(ct.grad.cpu().numpy(), ref_dc) < 1e-9
    assert rt1.grad is not None and rt2.grad is not None
    assert np.all(np.isfinite(rt1.grad.cpu().numpy()))
    op.close()


@pytest.mark.parametrize("nu,B,tol,prec", [
    (1.5, 12, 1e-10, "fp64"),   # two passes of 8 columns over the Chebyshev nodes
    (2.5, 3, 1e-6, "mixed"),
    (0.5, 9, 1e-6, "mixed"),
])
def test_half_integer_local_expansions(device, nu, B, tol, prec):
    """Half-integer orders: local boxes evaluate j_n at the Chebyshev nodes
    against moments folded with the interpolation matrix."""
    import torch
    import nufht_amd as na
    r, k, cs = _uniform(int(10 * nu) + B, 100000, 100000, 100.0, 100.0, B)
    p = na.Plan(nu, r, k, tol=tol, precision=prec)
    st = p.stats(B)
    say(f"nu={nu} B={B} tol={tol} {prec} M=N=1e5 on [0,100]^2: K_loc={p.K_loc} "
        f"boxes loc/asy/dir {st['n_loc']}/{st['n_asy']}/{st['n_dir']}")
    assert st["n_loc"] > 0
    g = p.forward(torch.as_tensor(np.ascontiguousarray(cs.T), device=device).t())
    torch.cuda.synchronize()
    rows = np.unique((W.Stream(int(10 * nu), 13).uniform(64) * k.size).astype(np.int32))
    ref = oracle.direct_rows(nu, r, cs, k, rows)
    err = rel_l2(g.cpu().numpy()[rows], ref)
    say(f"  rel l2 {err:.2e}")
    assert err < (2 * tol if prec == "fp64" else 1e-6)
    p.close()


def test_half_integer_multibox_local_evaluation(device):
    """The clustered nodes at nu = 1.5 with 32 columns: local evaluation
    items that loop over every box covering their targets, half-integer
    path (checked to be in use)."""
    import torch
    import nufht_amd as na
    wl = W.clustered()
    r, k, cs = wl["r"], wl["k"], wl["cs"]
    p = na.Plan(1.5, r, k, tol=1e-6, precision="fp64")
    st = p.stats(cs.shape[1])
    say(f"clustered nu=1.5 32 columns tol=1e-6: multi-box local items {st['loc_eval_multi']}")
    assert st["loc_eval_multi"] > 0
    g = p.forward(torch.as_tensor(np.asfortranarray(cs).T.copy(), device=device).t())
    torch.cuda.synchronize()
    rows = np.unique((W.Stream(15, 9).uniform(48) * k.size).astype(np.int32))
    ref = oracle.direct_rows(1.5, r, cs, k, rows)
    err = rel_l2(g.cpu().numpy()[rows], ref)
    say(f"  rel l2 {err:.2e}")
    assert err < 1e-6
    p.close()
