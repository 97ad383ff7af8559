"""The allocation-free entry point nufht_batch_device_noalloc, called
through the C ABI, against nufht_batch_device on the same inputs: two calls
in a row on unsorted grids (row staging and local-expansion moments in the
static scratch array), the second smaller than the first, plus its refusals
(no launch program yet; staging beyond the scratch array)."""
import ctypes

import numpy as np
import pytest

import oracle
from nufht_amd import workloads as W
from tests.conftest import rel_l2, say

pytestmark = pytest.mark.gpu


def _case(seed, m, n, R, K, B):
    s = W.Stream(seed, 707)
    r = s.uniform(m, 0.0, R)          # not sorted: the plan keeps permutations
    k = s.uniform(n, 0.0, K)
    cs = s.normal(m * B).reshape(m, B)
    return r, k, cs


def _run(fn, p, c, g, B):
    import torch
    import nufht_amd as na
    return fn(ctypes.byref(p._plan), c.data_ptr(), p.m, g.data_ptr(), p.n, B,
              na._stream_handle(torch.cuda.current_stream()))


def test_noalloc_matches_batch_device_two_calls(device):
    import torch
    import nufht_amd as na
    L = na.lib()
    for q, (m, n, B, nu, R) in enumerate([(20000, 15000, 3, 2.0, 200.0),
                                          (8000, 6000, 2, 1.0, 100.0)]):
        r, k, cs = _case(30 + q, m, n, R, R, B)
        p = na.Plan(nu, r, k, tol=1e-9, precision="fp64")
        assert p.stats(B)["n_loc"] > 0          # moments in the scratch array
        c = torch.as_tensor(np.ascontiguousarray(cs.T), device=device)
        g_ref = torch.empty((B, n), dtype=torch.float64, device=device)
        g_na = torch.full((B, n), 7.0, dtype=torch.float64, device=device)
        say(f"noalloc call {q}: nu={nu} m={m} n={n} B={B} unsorted grids")
        assert _run(L.nufht_batch_device, p, c, g_ref, B) == 0
        assert _run(L.nufht_batch_device_noalloc, p, c, g_na, B) == 0, na.last_error()
        torch.cuda.synchronize()
        a, b = g_na.cpu().numpy(), g_ref.cpu().numpy()
        assert rel_l2(a, b) < 1e-13, rel_l2(a, b)
        rows = np.arange(0, n, 37)
        ref = oracle.direct_rows(nu, r, cs, k, rows)
        assert rel_l2(a.T[rows], ref) < 3e-9
        p.close()


def test_noalloc_refusals(device):
    import torch
    import nufht_amd as na
    L = na.lib()
    r, k, cs = _case(40, 800, 700, 60.0, 60.0, 2)
    p = na.Plan(0.0, r, k, tol=1e-8)
    c = torch.as_tensor(np.ascontiguousarray(cs.T), device=device)
    g = torch.empty((2, 700), dtype=torch.float64, device=device)
    say("noalloc refusals: no launch program, then staging beyond the scratch array")
    assert _run(L.nufht_batch_device_noalloc, p, c, g, 2) == -551
    assert "nufht_batch_device once" in na.last_error()
    assert _run(L.nufht_batch_device, p, c, g, 2) == 0
    assert _run(L.nufht_batch_device_noalloc, p, c, g, 2) == 0
    g1 = torch.empty((1, 700), dtype=torch.float64, device=device)
    assert _run(L.nufht_batch_device_noalloc, p, c[:1], g1, 1) == -551   # other batch
    p.close()
    # 600000 unsorted nodes on both sides, 16 columns: 1.92e7 staged values
    s = W.Stream(41, 708)
    rb, kb = s.uniform(600000, 0.0, 10.0), s.uniform(600000, 0.0, 10.0)
    pb = na.Plan(0.0, rb, kb, tol=1e-8)
    cb = torch.zeros((16, 600000), dtype=torch.float64, device=device)
    gb = torch.empty((16, 600000), dtype=torch.float64, device=device)
    assert _run(L.nufht_batch_device_noalloc, pb, cb, gb, 16) == -550
    assert "exceeds the scratch" in na.last_error()
    pb.close()
