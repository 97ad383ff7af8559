"""GPU tests of column counts that are not powers of 2 on the matrix-core
interpolation paths.  A tile's 16 rows are (column, exponent) pairs: with B
columns (B < 16) each exponent group holds 16 // B exponents and the rows
past that are idle (B = 3, 5: row 15; B = 6: rows 12-15), and B = 21 adds a
second chunk of 5 columns.  The run kernel (interp_mfma = 1) and the
one-tile kernel (interp_mfma = 2) are checked against the oracle's FP64
direct sums on sampled rows and against the VALU interpolation on every
row."""
import numpy as np
import pytest

import oracle
from nufht_amd import workloads as W
from tests.conftest import rel_l2, say

pytestmark = pytest.mark.gpu


def _nodes():
    # M = N = 1e5 on [0, 300]^2: target tiles that fit the matrix-core windows
    s = W.Stream(7, 611)
    r = np.sort(s.uniform(100_000, 0.0, 300.0))
    k = np.sort(s.uniform(100_000, 0.0, 300.0))
    rows = np.unique((W.Stream(8, 611).uniform(64) * k.size).astype(np.int32))
    return r, k, rows


@pytest.mark.parametrize("batch", [3, 5, 6, 21])
def test_interpolation_column_counts(device, batch):
    import torch
    import nufht_amd as na
    r, k, rows = _nodes()
    cs = W.Stream(30 + batch, 611).normal(r.size * batch).reshape(batch, r.size).T.copy(order="F")
    ref = oracle.direct_rows(2.0, r, cs, k, rows)
    c = torch.as_tensor(np.asfortranarray(cs).T.copy(), device=device).t()
    p = na.Plan(2.0, r, k, tol=1e-8, precision="fp64")
    out = {}
    for mode in (0, 1, 2):
        say(f"nu=2 M=N=1e5 B={batch} tol=1e-8 interp_mfma={mode}")
        p.set_knobs(interp_mfma=mode)
        st = p.stats(batch)
        if mode == 0:
            assert st["interp_tc_mfma"] == 0
        else:
            assert st["interp_tc_mfma"] > 0
        g = p.forward(c).cpu().numpy()
        err = rel_l2(g[rows], ref)
        say(f"  rel l2 vs oracle {err:.2e}")
        assert err < 2e-8
        out[mode] = g
    p.close()
    for mode in (1, 2):
        d = rel_l2(out[mode], out[0])
        say(f"  interp_mfma={mode} vs VALU {d:.2e}")
        assert d < 2e-8
