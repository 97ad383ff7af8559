"""GPU tests of the launch-geometry knobs and the autotuner: every code path a
knob selects (VALU or matrix-core spread and interpolation, split sizes, the
two-pass FFT kernels or rocFFT) gives the transform to the requested
tolerance, checked against the oracle's FP64 direct sums on sampled rows;
the autotuned plan and the autotuned multi-transform batch (then replayed
from a hipGraph) agree with the defaults."""
import numpy as np
import pytest

import oracle
from nufht_amd import workloads as W
from tests.conftest import rel_l2, say

pytestmark = pytest.mark.gpu


def _case():
    # M = N = 1e5 on [0, 300]^2: asymptotic boxes with source-dense grids
    # (matrix-core spread), FFT lengths above 4096 (two-pass FFT kernels) and
    # target tiles that fit a 16-position window (matrix-core interpolation)
    s = W.Stream(7, 611)
    r = np.sort(s.uniform(100_000, 0.0, 300.0))
    k = np.sort(s.uniform(100_000, 0.0, 300.0))
    cs = s.normal(100_000 * 4).reshape(4, 100_000).T.copy(order="F")
    rows = np.unique((W.Stream(8, 611).uniform(96) * k.size).astype(np.int32))
    return r, k, cs, rows


ALTERNATIVES = [
    {},
    {"dense_cell": 1e300},          # VALU spread only
    {"dense_cell": 0.25},           # matrix-core spread for almost every box
    {"interp_mfma": 0},             # VALU interpolation only
    {"interp_mfma": 2},             # matrix-core interpolation, one tile per item
    {"own_fft": 0},                 # rocFFT for every length
    {"dense_split": 2048, "spread_split": 1024},
    {"dense_split": 8192, "spread_split": 4096},
    {"dense_split": 64},            # matrix-core spread: the smallest pieces (most boundaries)
    {"dense_split": 0},             # matrix-core spread: pieces sized from the kernel width
    {"spread_sched": 0},            # matrix-core spread: unified schedule
    {"spread_sched": 1},            # matrix-core spread: split schedule
]


def test_every_knob_path_meets_tolerance(device):
    import torch
    import nufht_amd as na
    r, k, cs, rows = _case()
    ref = oracle.direct_rows(3.0, r, cs, k, rows)
    c = torch.as_tensor(np.asfortranarray(cs).T.copy(), device=device).t()
    p = na.Plan(3.0, r, k, tol=1e-8, precision="fp64")
    base = None
    for alt in ALTERNATIVES:
        say(f"nu=3 M=N=1e5 B=4 tol=1e-8 knobs {alt or 'default'}")
        p.set_knobs(**{**{"dense_cell": 4.0, "dense_split": 4096, "spread_split": 2048,
                          "interp_mfma": 1, "own_fft": 1, "spread_sched": -1}, **alt})
        st = p.stats(4)
        if alt.get("dense_cell") == 1e300:
            assert st["spread_sc_mfma"] == 0
        if alt.get("interp_mfma") == 0:
            assert st["interp_tc_mfma"] == 0
        if alt.get("interp_mfma") == 2:
            assert st["interp_tc_mfma"] > 0
        if not alt:
            assert st["spread_sc_mfma"] > 0 and st["interp_tc_mfma"] > 0
        g = p.forward(c).cpu().numpy()
        err = rel_l2(g[rows], ref)
        say(f"  rel l2 {err:.2e}")
        assert err < 1e-8
        if base is None:
            base = g
        else:
            assert rel_l2(g, base) < 1e-8
    p.close()


def test_autotuned_plan_and_multi_batch(device):
    import torch
    import nufht_amd as na
    r, k, cs, rows = _case()
    ref = oracle.direct_rows(3.0, r, cs, k, rows)
    c = torch.as_tensor(np.asfortranarray(cs).T.copy(), device=device).t()
    say("autotune nu=3 M=N=1e5 B=4 tol=1e-8")
    p = na.Plan(3.0, r, k, tol=1e-8)
    kn = p.autotune(4)
    say(f"  chosen {kn}")
    assert kn["precision"] == na.PREC_AUTO          # tol < 1e-6: no precision choice
    g = p.forward(c).cpu().numpy()
    assert rel_l2(g[rows], ref) < 1e-8
    p.close()

    say("autotune mixed-eligible plan tol=1e-6")
    p = na.Plan(3.0, r, k, tol=1e-6)
    kn = p.autotune(4)
    assert kn["precision"] in (na.PREC_FP64, na.PREC_MIXED)
    g = p.forward(c).cpu().numpy()
    assert rel_l2(g[rows], ref) < 1e-6
    p.close()

    ts = W.uneven_batch(count=16, mmin=500, mmax=4000)
    plans, cs_d, gs_d, refs = [], [], [], []
    for t in ts:
        plans.append(na.Plan(t["nu"], t["r"], t["k"], tol=1e-10))
        cs_d.append(torch.as_tensor(np.asfortranarray(t["cs"]).T.copy(), device=device))
        gs_d.append(torch.empty((t["cs"].shape[1], t["k"].size), dtype=torch.float64,
                                device=device))
        refs.append(oracle.direct(t["nu"], t["r"], t["cs"], t["k"]))
    say("multi-batch 16 transforms: autotune, then graph replays")
    mb = na.MultiBatch(plans, [t["cs"].shape[1] for t in ts])
    mb.execute(cs_d, gs_d, use_graph=True)
    kn = mb.autotune()
    say(f"  chosen {kn}")
    for rep in range(2):
        for g in gs_d:
            g.fill_(7.0)
        mb.execute(cs_d, gs_d, use_graph=True)
        torch.cuda.synchronize()
        for g, ref in zip(gs_d, refs):
            assert rel_l2(g.t().cpu().numpy(), ref) < 5e-10
        _ = torch.ones(10, device=device).sum().item()
    mb.close()
    for p in plans:
        p.close()


@pytest.mark.parametrize("case", [
    # one asymptotic box of Nf = 9 * 2^10 (r in [100, 300], k in [100, 170])
    ("9216", 100.0, 300.0, 100.0, 170.0),
    # the largest class at Nf = 9 * 2^12 (r, k in [0, 250])
    ("36864", 0.0, 250.0, 0.0, 250.0),
])
@pytest.mark.parametrize("tol,precision", [(1e-8, "fp64"), (1e-6, "mixed")])
def test_fft_lengths_nine_times_power_of_two(device, case, tol, precision):
    """FFT lengths 9 * 2^a (the two-pass FFT kernels' radix-9 stage; the
    plan's largest asymptotic box gets that length for these node ranges)
    against the oracle's FP64 direct sums on sampled rows."""
    import torch
    import nufht_amd as na
    name, ra, rb, ka, kb = case
    s = W.Stream(7, 611)
    n = 50_000
    r = np.sort(s.uniform(n, ra, rb))
    k = np.sort(s.uniform(n, ka, kb))
    cs = s.normal(n * 3).reshape(3, n).T.copy(order="F")
    rows = np.unique((W.Stream(9, 611).uniform(64) * k.size).astype(np.int32))
    ref = oracle.direct_rows(3.0, r, cs, k, rows)
    c = torch.as_tensor(np.asfortranarray(cs).T.copy(), device=device).t()
    say(f"Nf {name} case: nu=3 M=N=5e4 B=3 tol={tol} {precision}")
    p = na.Plan(3.0, r, k, tol=tol, precision=precision)
    g = p.forward(c).cpu().numpy()
    err = rel_l2(g[rows], ref)
    say(f"  rel l2 {err:.2e}")
    assert err < tol
    p.close()
