"""The benchmark workloads at their full sizes, checked on sampled output
rows against the oracle's FP64 direct sums (the full direct sums are 1e12
pairs): the headline transform and its adjoint, the clustered tol 1e-6
workload in FP64 and in the FP32-mixed mode (relative l2 errors of both are
printed; both must be below the requested tolerance 1e-6), and transforms of the uneven batch
run through one multi-transform launch."""
import numpy as np
import pytest

import oracle
from nufht_amd import workloads as W
from tests.conftest import rel_l2, say

pytestmark = pytest.mark.gpu


def _rows(n, count, seed):
    return np.unique((W.Stream(seed, 9).uniform(count) * n).astype(np.int32))


def test_headline_forward_and_adjoint_sampled_rows(device):
    import torch
    import nufht_amd as na
    wl = W.large_uniform()
    r, k, cs = wl["r"], wl["k"], wl["cs"]
    say("nu=5 M=N=1e6 B=16 tol=1e-8: forward and adjoint")
    p = na.Plan(5.0, r, k, tol=1e-8, precision="fp64")
    c = torch.as_tensor(np.asfortranarray(cs).T.copy(), device=device).t()
    g = p.forward(c)
    x = p.adjoint(g)
    torch.cuda.synchronize()
    rows = _rows(k.size, 48, 20)
    ref = oracle.direct_rows(5.0, r, cs, k, rows)
    err_f = rel_l2(g.cpu().numpy()[rows], ref)
    gh = np.ascontiguousarray(g.cpu().numpy())
    cols = _rows(r.size, 24, 21)
    ref_a = oracle.direct_rows(5.0, k, gh, r, cols)
    err_a = rel_l2(x.cpu().numpy()[cols], ref_a)
    say(f"  forward rel l2 {err_f:.2e}, adjoint rel l2 {err_a:.2e}")
    assert err_f < 2e-8 and err_a < 2e-8
    p.close()


def test_clustered_mixed_and_fp64_meet_tolerance(device):
    import torch
    import nufht_amd as na
    wl = W.clustered()
    r, k, cs = wl["r"], wl["k"], wl["cs"]
    c = torch.as_tensor(np.asfortranarray(cs).T.copy(), device=device).t()
    rows = _rows(k.size, 48, 40)
    ref = oracle.direct_rows(1.0, r, cs, k, rows)
    errs = {}
    for prec in ("fp64", "mixed"):
        say(f"clustered nu=1 M=N=1e6 32 columns tol=1e-6 {prec}")
        p = na.Plan(1.0, r, k, tol=1e-6, precision=prec)
        g = p.forward(c)
        torch.cuda.synchronize()
        errs[prec] = rel_l2(g.cpu().numpy()[rows], ref)
        p.close()
    say(f"  rel l2: fp64 {errs['fp64']:.2e}, mixed {errs['mixed']:.2e}")
    assert errs["fp64"] < 1e-6 and errs["mixed"] < 1e-6


@pytest.mark.parametrize("tol", [1e-10, 1e-6])
def test_uneven_batch_subset_one_launch(device, tol):
    """Every 16th transform of the 256-transform uneven batch (16 transforms),
    one multi-transform launch; each checked on up to 64 sampled rows."""
    import torch
    import nufht_amd as na
    ts = W.uneven_batch()[::16]
    plans = [na.Plan(t["nu"], t["r"], t["k"], tol=tol, precision="fp64") for t in ts]
    cs = [torch.as_tensor(np.asfortranarray(t["cs"]).T.copy(), device=device) for t in ts]
    gs = [torch.empty((t["cs"].shape[1], t["k"].size), dtype=torch.float64, device=device)
          for t in ts]
    say(f"uneven batch subset: {len(ts)} transforms tol={tol}")
    mb = na.MultiBatch(plans, [t["cs"].shape[1] for t in ts])
    mb.execute(cs, gs, use_graph=True)
    torch.cuda.synchronize()
    for q, (t, g) in enumerate(zip(ts, gs)):
        rows = _rows(t["k"].size, 64, 60 + q)
        ref = oracle.direct_rows(t["nu"], t["r"], t["cs"], t["k"], rows)
        err = rel_l2(g.t().cpu().numpy()[rows], ref)
        assert err < 3 * tol, (q, t["nu"], t["r"].size, t["k"].size, err)
    mb.close()
    for p in plans:
        p.close()
