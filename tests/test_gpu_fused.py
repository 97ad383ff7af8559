"""GPU tests of the fused small-box kernel (knob fused_max): asymptotic boxes
with small uniform grids spread, transformed and interpolated with the grid
in LDS, in one persistent launch over a work queue of units.  The fused
transform is checked against the oracle's FP64 direct sums on sampled rows
and against the staged kernels (fused_max = 0): one column (the one-column
lane layout), five columns (column pairs), mixed precision and the
adjoint."""
import numpy as np
import pytest

import oracle
from nufht_amd import workloads as W
from tests.conftest import rel_l2, say

pytestmark = pytest.mark.gpu


def _case(batch, seed):
    # M = N = 2e4 on [0, 200]^2: most asymptotic boxes have uniform grids of
    # at most 256 points and a few thousand sources and targets
    s = W.Stream(seed, 733)
    r = np.sort(s.uniform(20_000, 0.0, 200.0))
    k = np.sort(s.uniform(20_000, 0.0, 200.0))
    cs = s.normal(20_000 * batch).reshape(batch, 20_000).T.copy(order="F")
    rows = np.unique((W.Stream(seed + 1, 733).uniform(80) * k.size).astype(np.int32))
    return r, k, cs, rows


def _cols(cs, device):
    import torch
    return torch.as_tensor(np.asfortranarray(cs).T.copy(), device=device).t()


@pytest.mark.parametrize("nu,batch,tol,prec", [
    (3.0, 1, 1e-8, "fp64"),
    (3.0, 5, 1e-10, "fp64"),
    (0.5, 4, 1e-8, "fp64"),
    (1.0, 8, 1e-6, "mixed"),
])
def test_fused_matches_oracle_and_staged(device, nu, batch, tol, prec):
    import nufht_amd as na
    r, k, cs, rows = _case(batch, 11 + batch)
    say(f"fused nu={nu} M=N=2e4 B={batch} tol={tol} {prec}")
    ref = oracle.direct_rows(nu, r, cs, k, rows)
    c = _cols(cs, device)
    p = na.Plan(nu, r, k, tol=tol, precision=prec)
    p.set_knobs(fused_max=0)
    staged = p.forward(c).cpu().numpy()
    assert p.stats(batch)["fused_boxes"] == 0
    p.set_knobs(fused_max=4096)
    st = p.stats(batch)
    say(f"  fused boxes {st['fused_boxes']} of {st['n_asy']}")
    assert st["fused_boxes"] > 0
    fused = p.forward(c).cpu().numpy()
    p.close()
    bound = 2.0 * tol if prec == "fp64" else 1e-6
    err = rel_l2(fused[rows], ref)
    say(f"  rel l2 vs oracle {err:.2e}, vs staged {rel_l2(fused, staged):.2e}")
    assert err < bound
    assert rel_l2(fused, staged) < bound


def test_fused_adjoint_matches_staged(device):
    import nufht_amd as na
    r, k, cs, rows = _case(3, 21)
    say("fused adjoint nu=2 M=N=2e4 B=3 tol=1e-8")
    y = _cols(np.asfortranarray(cs[: k.size]), device)
    p = na.Plan(2.0, r, k, tol=1e-8, precision="fp64")
    p.set_knobs(fused_max=0)
    staged = p.adjoint(y).cpu().numpy()
    p.set_knobs(fused_max=4096)
    fused = p.adjoint(y).cpu().numpy()
    p.close()
    ref = oracle.direct_rows(2.0, k, np.asfortranarray(cs[: k.size]), r, rows)
    err = rel_l2(fused[rows], ref)
    say(f"  rel l2 vs oracle {err:.2e}, vs staged {rel_l2(fused, staged):.2e}")
    assert err < 2e-8
    assert rel_l2(fused, staged) < 2e-8
