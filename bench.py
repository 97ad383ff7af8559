"""Benchmark: NUFHT column-points/s, forward + adjoint, on one MI355X.

Headline workload: nu = 5, M = N = 1e6 nodes r, k ~ U[0, 1000) (sorted),
B = 16 N(0,1) coefficient columns, tol 1e-8, FP64.  One step is one forward
transform g = H c plus one adjoint x = H^T g on inputs already resident in
HBM; column_points per step = 2 * B * (M + N).

The JSON line carries the contract fields plus
  roofline      the dominant kernel (the stage with the most device time;
                every stage but asy_fft is one kernel), timed with HIP
                events on its launch stream inside the timed region, with
                its algorithmic work per launch from the plan statistics;
                traffic (HBM bytes per launch) is read from the committed
                rocprofv3 FETCH_SIZE / WRITE_SIZE summary of that kernel
                when one exists for this workload;
  cpu_baseline  the oracle's CPU restatement of the reference algorithm
                (boxes, Wimp and Hankel expansions, its own CPU type-3
                NUFFT) on a bounded sample of the same workload, 1 thread;
  plain         the same engine run one column at a time (the reference's
                nufht_batch_loop pattern);
  targets       the other workloads, each against its own comparator:
                uneven 256-transform batch (one hipGraph-replayed launch
                per stage vs a per-transform loop) at tol 1e-10 and 1e-6,
                FP32-mixed vs FP64 at tol 1e-6 on the clustered nu = 1
                workload, and the autograd forward + backward of the
                gradient workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                       [--no-targets] [--copy-calibration]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_FP64_TFLOPS = 78.6     # MI355X dense FP64; vector and matrix peaks are equal
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
METRIC = "NUFHT column-points/s at tol 1e-8/1e-6, batched vs plain GPU; fwd+adjoint"


def say(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_work(st: dict) -> dict:
    """Algorithmic work of one launch of each single-kernel stage, from the
    plan statistics of the program that ran it.

    A channel is one complex series (exponent e, column t); a box has
    C = 2 (K_asy + 1) B of them.  The per-unit figures are the algorithm's
    operation counts: 2 w flops per (source, channel) spread, 4 w flops per
    (target, channel) interpolated (w = kernel width), 2 B flops per direct
    (target, source) pair with the Bessel evaluation not counted, 2 (K_loc +
    1) B flops per local-expansion source (projection) and target
    (evaluation).  FFT: the bytes the two-pass kernels must move (pass 1
    reads the 2L + 1 spread positions of a channel and writes its Nf points,
    pass 2 reads Nf and writes the frequencies interpolation reads).
    """
    kw = st["kernel_width"]
    return {
        "asy_spread": ("flop", 2.0 * kw * st["spread_sc_valu"]),
        "asy_spread_mfma": ("flop", 2.0 * kw * st["spread_sc_mfma"]),
        "asy_interp": ("flop", 4.0 * kw * st["interp_tc_valu"]),
        "asy_interp_mfma": ("flop", 4.0 * kw * st["interp_tc_mfma"]),
        "direct": ("flop", st["model_dir"]),
        "loc_project": ("flop", st["model_loc"] * st["loc_src"] / max(1, st["loc_src"] + st["loc_tgt"])),
        "loc_eval": ("flop", st["model_loc"] * st["loc_tgt"] / max(1, st["loc_src"] + st["loc_tgt"])),
        "asy_fft": ("byte", st["fft_bytes"]),
    }


def path_model(stats: list, t_measured_ms: float) -> dict:
    """Whole-path roofline of one step: the algorithmic work of every
    application the step runs (FP64 and FP32 flops, HBM bytes; the plan
    statistics' model_* totals), its floor t_roof = max(F64 / 78.6 TFLOP/s +
    F32 / 157.3 TFLOP/s, bytes / 8 TB/s) and achieved = t_roof / t_measured."""
    f64 = sum(st["model_f64"] for st in stats)
    f32 = sum(st["model_f32"] for st in stats)
    by = sum(st["model_bytes"] for st in stats)
    t_flop = f64 / (PEAK_FP64_TFLOPS * 1e12) + f32 / (PEAK_FP32_TFLOPS * 1e12)
    t_byte = by / (PEAK_HBM_GBS * 1e9)
    t_roof = max(t_flop, t_byte) * 1e3
    return {"t_roof_ms": t_roof, "t_measured_ms": t_measured_ms,
            "achieved": t_roof / t_measured_ms, "bound": "flop" if t_flop >= t_byte else "hbm",
            "fp64_flops": f64, "fp32_flops": f32, "bytes": by, "applications": len(stats)}


def _traffic(stage: str, tag: str):
    """HBM bytes per launch of `stage` on workload `tag` from the committed
    rocprofv3 counter summary, or None."""
    path = os.path.join(HERE, "profiles", f"traffic_{stage}_{tag}.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"^bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), flags=re.M)
    return float(m.group(1)) if m else None


def roofline(prof: dict, works: list, mixed: bool, tag) -> dict:
    """Dominant kernel over a profiled region (the single-kernel stages; the
    FFT stage is two kernels per size class and is reported by fft_stage);
    works = stage_work dicts of the programs whose launches ran, each program
    launched equally often (forward and adjoint)."""
    cand = [s for s in ("direct", "loc_project", "loc_eval", "asy_spread", "asy_spread_mfma",
                        "asy_interp", "asy_interp_mfma")
            if prof.get(s, {}).get("launches", 0) > 0]
    dom = max(cand, key=lambda s: prof[s]["ms"])
    launches = prof[dom]["launches"]
    avg_ms = prof[dom]["ms"] / launches
    kind = works[0][dom][0]
    # launches of this stage per pass over the programs (a program without
    # work for the stage does not launch it)
    ran = [w[dom][1] for w in works if w[dom][1] > 0]
    per_launch = sum(ran) / max(1, len(ran))
    if kind == "byte":
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s"}
    else:
        peak = PEAK_FP32_TFLOPS if (mixed and dom.startswith("asy")) else PEAK_FP64_TFLOPS
        achieved = per_launch / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma" if "mfma" in dom or dom == "loc_project" else "valu",
                "achieved": achieved, "peak": peak, "unit": "TFLOP/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = _traffic(dom, tag) if tag else None
    roof.update({"kernel": dom, "avg_launch_ms": avg_ms, "launches": launches,
                 "algorithmic_per_launch": per_launch})
    return roof


def fft_stage(prof: dict, works: list):
    """The uniform-FFT stage against HBM: bytes the kernels must move (pruned)
    per stage launch over its measured time."""
    if prof.get("asy_fft", {}).get("launches", 0) == 0:
        return None
    f_ms = prof["asy_fft"]["ms"] / prof["asy_fft"]["launches"]
    ran = [w["asy_fft"][1] for w in works if w["asy_fft"][1] > 0]
    f_b = sum(ran) / max(1, len(ran))
    return {"bound": "hbm", "avg_stage_ms": f_ms, "algorithmic_bytes": f_b,
            "achieved_GBps": f_b / (f_ms * 1e-3) / 1e9,
            "frac": f_b / (f_ms * 1e-3) / 1e9 / PEAK_HBM_GBS}


def _profiled(fn, steps: int) -> dict:
    """Per-stage device times over `steps` calls of fn (HIP events around
    every stage launch, on the launch stream)."""
    import torch
    import nufht_amd as na
    torch.cuda.synchronize()
    na.profile(True)
    na.profile_reset()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    prof = na.profile_read()
    na.profile(False)
    return prof


def _stages(prof: dict, steps: int) -> dict:
    return {s: round(v["ms"] / steps, 4) for s, v in prof.items() if v["launches"] > 0}


def _dev_cols(a: np.ndarray, dev):
    """(m, B) host array -> (m, B) column-major device tensor."""
    import torch
    return torch.as_tensor(np.asfortranarray(a).T.copy(), device=dev).t()


# ---------------------------------------------------------------- headline
def run_headline(args, dev):
    import torch
    import nufht_amd as na
    from nufht_amd import workloads as W

    wl = W.large_uniform(batch=args.batch)
    r, k, cs, nu, tol = wl["r"], wl["k"], wl["cs"], wl["nu"], wl["tol"]
    M, N, B = r.size, k.size, cs.shape[1]
    say(f"headline nu={nu} M={M} N={N} B={B} tol={tol} precision={args.precision}")
    t0 = time.time()
    plan = na.Plan(nu, r, k, tol=tol, precision=args.precision)
    plan_s = time.time() - t0
    c = _dev_cols(cs, dev)

    def step():
        g = plan.forward(c)
        plan.adjoint(g)

    t0 = time.time()
    step()
    torch.cuda.synchronize()
    first_s = time.time() - t0
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    na.profile(True)
    na.profile_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = na.profile_read()
    na.profile(False)
    ms_step = (t1 - t0) * 1e3 / args.steps
    colpts = 2.0 * B * (M + N)
    value = colpts / ((t1 - t0) / args.steps)

    st_f = plan.stats(B)
    adj = na.Plan(nu, k, r, tol=tol, precision=args.precision)
    st_a = adj.stats(B)
    adj.close()
    mixed = st_f["precision"] == 2
    prec = "mixed" if mixed else "fp64"
    tag = f"nu5_M1e6_N1e6_B{B}_tol1e-8_{prec}"
    works = [stage_work(st_f), stage_work(st_a)]
    roof = roofline(prof, works, mixed, tag)
    path = path_model([st_f, st_a], ms_step)
    roof["achieved_path"] = path["achieved"]
    roof["path"] = path
    fft = fft_stage(prof, works)
    # the same kernel and FFT stage alone: a few steps with the near-field
    # stages on the calling stream (in the timed steps they run on a second
    # stream and share the GPU with the asymptotic kernels)
    prev = na.stream_fork(0)
    prof1 = _profiled(step, 3)
    na.stream_fork(prev)
    r1 = roofline(prof1, works, mixed, None)
    f1 = fft_stage(prof1, works)
    roof["alone"] = {"avg_launch_ms": r1["avg_launch_ms"], "achieved": r1["achieved"],
                     "frac": r1["frac"], "kernel": r1["kernel"],
                     "fft_stage_ms": f1["avg_stage_ms"] if f1 else None,
                     "fft_frac": f1["frac"] if f1 else None,
                     "stages_ms_per_step": _stages(prof1, 3)}

    plain = None
    if args.plain_steps > 0:
        say("plain per-column loop")
        c1 = [c[:, t:t + 1] for t in range(B)]
        for t in range(B):
            plan.adjoint(plan.forward(c1[t]))
        torch.cuda.synchronize()
        p0 = time.perf_counter()
        for _ in range(args.plain_steps):
            for t in range(B):
                plan.adjoint(plan.forward(c1[t]))
        torch.cuda.synchronize()
        p1 = time.perf_counter()
        pv = colpts / ((p1 - p0) / args.plain_steps)
        plain = {"value": pv, "ms_per_step": (p1 - p0) * 1e3 / args.plain_steps,
                 "speedup_batched_over_plain": value / pv}
    plan.close()
    stages = _stages(prof, args.steps)
    return {"value": value, "ms_per_step": ms_step, "M": M, "N": N, "B": B, "nu": nu,
            "tol": tol, "plan_s": plan_s, "first_call_s": first_s, "stages_ms": stages,
            "roofline": roof, "fft_stage": fft, "plain": plain, "precision": prec}


# ---------------------------------------------------------------- targets
def run_uneven(dev, tol: float, steps: int):
    """256 transforms of uneven size and order: MultiBatch (one launch per
    stage, replayed as a hipGraph) for the forwards and another for the
    adjoints (plans with r and k swapped), against the per-transform loop."""
    import torch
    import nufht_amd as na
    from nufht_amd import workloads as W
    ts = W.uneven_batch()
    say(f"uneven batch: {len(ts)} transforms tol={tol}")
    t0 = time.time()
    pf = [na.Plan(t["nu"], t["r"], t["k"], tol=tol, precision="fp64") for t in ts]
    pa = [na.Plan(t["nu"], t["k"], t["r"], tol=tol, precision="fp64") for t in ts]
    plan_s = time.time() - t0
    Bs = [t["cs"].shape[1] for t in ts]
    cs = [torch.as_tensor(np.asfortranarray(t["cs"]).T.copy(), device=dev) for t in ts]
    gs = [torch.empty((b, t["k"].size), dtype=torch.float64, device=dev) for b, t in zip(Bs, ts)]
    xs = [torch.empty((b, t["r"].size), dtype=torch.float64, device=dev) for b, t in zip(Bs, ts)]
    colpts = 2.0 * sum(b * (t["r"].size + t["k"].size) for b, t in zip(Bs, ts))
    mf, ma = na.MultiBatch(pf, Bs), na.MultiBatch(pa, Bs)
    t0 = time.time()
    knobs = {"forward": mf.autotune(), "adjoint": ma.autotune()}
    tune_s = time.time() - t0

    def batched():
        mf.execute(cs, gs, use_graph=True)
        ma.execute(gs, xs, use_graph=True)

    def plain():
        for p, c in zip(pf, cs):
            p.adjoint(p.forward(c.t()))

    out = {"transforms": len(ts), "tol": tol, "plan_s": plan_s, "autotune_s": tune_s,
           "knobs": knobs, "column_points": colpts}
    for name, fn in (("batched", batched), ("plain", plain)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        out[name] = {"value": colpts / dt, "ms_per_step": dt * 1e3}
    out["speedup_batched_over_plain"] = out["batched"]["value"] / out["plain"]["value"]
    # per-stage times from a separate profiled pass (the launches run eagerly
    # there: events cannot sit inside the replayed graph)
    st_f, st_a = mf.stats(), ma.stats()
    prof = _profiled(lambda: (mf.execute(cs, gs, use_graph=False),
                              ma.execute(gs, xs, use_graph=False)), steps)
    works = [stage_work(st_f), stage_work(st_a)]
    out["roofline"] = roofline(prof, works, False, None)
    out["roofline"]["path"] = path_model([st_f, st_a], out["batched"]["ms_per_step"])
    out["roofline"]["achieved_path"] = out["roofline"]["path"]["achieved"]
    out["fft_stage"] = fft_stage(prof, works)
    out["stages_ms_per_step_eager"] = _stages(prof, steps)
    mf.close()
    ma.close()
    for p in pf + pa:
        p.close()
    return out


def run_mixed(dev, steps: int):
    """Clustered nu = 1 workload, 16 complex = 32 real columns, tol 1e-6: the
    FP32-mixed path against the FP64 path (fwd + adjoint each), with the
    relative l2 distance of the mixed outputs from the FP64 outputs."""
    import torch
    import nufht_amd as na
    from nufht_amd import workloads as W
    wl = W.clustered()
    r, k, cs = wl["r"], wl["k"], wl["cs"]
    M, N, B = r.size, k.size, cs.shape[1]
    say(f"mixed vs fp64: nu=1 M={M} N={N} B={B} tol=1e-6")
    c = _dev_cols(cs, dev)
    colpts = 2.0 * B * (M + N)
    out, res = {"M": M, "N": N, "columns": B, "tol": 1e-6}, {}
    for prec in ("fp64", "mixed"):
        plan = na.Plan(1.0, r, k, tol=1e-6, precision=prec)
        g = plan.forward(c)
        x = plan.adjoint(g)
        torch.cuda.synchronize()
        res[prec] = (g.clone(), x.clone())
        t0 = time.perf_counter()
        for _ in range(steps):
            plan.adjoint(plan.forward(c))
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        out[prec] = {"value": colpts / dt, "ms_per_step": dt * 1e3}
        st_f = plan.stats(B)
        adj = na.Plan(1.0, k, r, tol=1e-6, precision=prec)
        st_a = adj.stats(B)
        adj.close()
        prof = _profiled(lambda: plan.adjoint(plan.forward(c)), steps)
        works = [stage_work(st_f), stage_work(st_a)]
        roof = roofline(prof, works, prec == "mixed", None)
        roof["path"] = path_model([st_f, st_a], dt * 1e3)
        roof["achieved_path"] = roof["path"]["achieved"]
        out[prec]["roofline"] = roof
        out[prec]["fft_stage"] = fft_stage(prof, works)
        out[prec]["stages_ms_per_step"] = _stages(prof, steps)
        plan.close()
    gf, xf = res["fp64"]
    gm, xm = res["mixed"]
    out["speedup_mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    out["mixed_vs_fp64_rel_l2_forward"] = float(torch.linalg.norm(gm - gf) / torch.linalg.norm(gf))
    out["mixed_vs_fp64_rel_l2_adjoint"] = float(torch.linalg.norm(xm - xf) / torch.linalg.norm(xf))
    return out


def run_gradient(dev, steps: int):
    """Gradient workload (nu = 1, M = N = 2e5, B = 8, tol 1e-10): loss
    ||H(c; r, k) - y||^2 through the autograd layer, forward + backward
    (dL/dc by the adjoint, dL/dr and dL/dk by the node-gradient transforms)."""
    import torch
    import nufht_amd as na
    from nufht_amd import workloads as W
    from nufht_amd.autograd import NufhtOperator
    wl = W.gradient()
    say("gradient: nu=1 M=N=2e5 B=8 tol=1e-10 through autograd")
    op = NufhtOperator(1.0, tol=1e-10, precision="fp64")
    rt = torch.tensor(wl["r"], device=dev, requires_grad=True)
    kt = torch.tensor(wl["k"], device=dev, requires_grad=True)
    ct = torch.tensor(wl["c0"], device=dev, requires_grad=True)
    with torch.no_grad():
        y = op(rt, torch.as_tensor(wl["cstar"], device=dev), kt) + \
            torch.as_tensor(1e-3 * wl["noise"], device=dev)

    def step():
        for t in (rt, kt, ct):
            t.grad = None
        loss = ((op(rt, ct, kt) - y) ** 2).sum()
        loss.backward()
        return loss

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    prof = _profiled(step, steps)
    op.close()
    # the step's applications: forward H, backward H^T (dL/dc), and for the
    # node gradients H^T and H^T of order nu + 1 (dL/dr), H and H of order
    # nu + 1 (dL/dk)
    r, k, B = wl["r"], wl["k"], 8
    sts = {}
    for name, (nu_, a, b) in {"fwd": (1.0, r, k), "adj": (1.0, k, r), "fwd_up": (2.0, r, k),
                              "adj_up": (2.0, k, r)}.items():
        p = na.Plan(nu_, a, b, tol=1e-10, precision="fp64")
        sts[name] = p.stats(B)
        p.close()
    apps = [sts["fwd"], sts["fwd"], sts["adj"], sts["adj"], sts["fwd_up"], sts["adj_up"]]
    works = [stage_work(st) for st in apps]
    roof = roofline(prof, works, False, None)
    roof["path"] = path_model(apps, dt * 1e3)
    roof["achieved_path"] = roof["path"]["achieved"]
    return {"M": wl["r"].size, "N": wl["k"].size, "columns": 8, "tol": 1e-10,
            "ms_forward_backward": dt * 1e3, "roofline": roof,
            "stages_ms_per_step": _stages(prof, steps)}


# ---------------------------------------------------------------- CPU baseline
def _cpu_info() -> dict:
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_available": avail}


def _time_port(nu, r, c, k, tol, threads):
    """Forward + adjoint through the oracle's CPU restatement of the
    reference's nufht_batch (boxes, Wimp and Hankel expansions, its own CPU
    type-3 NUFFT); column groups split over `threads` OpenMP threads."""
    import oracle
    t0 = time.perf_counter()
    g = oracle.nufht_batch(nu, r, c, k, tol=tol, asy_mode=1, nthreads=threads)
    oracle.nufht_batch(nu, k, g, r, tol=tol, asy_mode=1, nthreads=threads)
    return time.perf_counter() - t0


def run_cpu_baseline(sub: int = 8):
    """The oracle's CPU restatement of the reference algorithm on the GPU
    box's host: the headline workload on every sub-th node (M = N = 125000)
    with all 16 columns over all host threads available to this process (at
    most 16, one GPU's share; one column group per thread), and the same
    sample with one column on one thread (the reference's own threading:
    FINUFFT nthreads = 1, no OpenMP regions); plus the Gauss-Legendre nu = 0
    workload (10000 nodes, 200 k, 4 columns) in full at both."""
    from nufht_amd import workloads as W
    info = _cpu_info()
    threads = max(1, min(16, info["cpus_available"]))
    wl = W.large_uniform()
    r, k, cs = wl["r"][::sub], wl["k"][::sub], wl["cs"][::sub]
    B = cs.shape[1]
    say(f"cpu baseline: M=N={r.size} {B} columns, {threads} threads")
    dt_all = _time_port(wl["nu"], r, cs, k, wl["tol"], threads)
    say(f"cpu baseline: M=N={r.size} 1 column, 1 thread")
    dt_one = _time_port(wl["nu"], r, cs[:, :1], k, wl["tol"], 1)
    gl = W.gaussian_gl()
    dt_gl1 = _time_port(0.0, gl["r"], gl["cs"], gl["k"], 1e-8, 1)
    dt_glt = _time_port(0.0, gl["r"], gl["cs"], gl["k"], 1e-8, min(threads, 4))
    cp = 2.0 * (r.size + k.size)
    cp_gl = 2.0 * gl["cs"].shape[1] * (gl["r"].size + gl["k"].size)
    return {"value": B * cp / dt_all, "unit": "column-points/s", "cores": threads,
            "kind": "port",
            "sample": f"nu=5 tol=1e-8, every {sub}th node of the M=N=1e6 U[0,1000) grids "
                      f"(M=N={r.size}), {B} columns, forward+adjoint, {threads} threads "
                      f"(one column group each), {dt_all:.1f} s",
            **info,
            "one_thread": {"value": cp / dt_one, "cores": 1,
                           "sample": f"the same nodes, 1 column, forward+adjoint, {dt_one:.1f} s"},
            "gauss_legendre_full": {
                "workload": "nu=0 tol=1e-8, GL10000 on [0,50], k=10(j+1)/200, 4 columns, "
                            "forward+adjoint",
                "one_thread": cp_gl / dt_gl1, "threads": min(threads, 4),
                "all_threads": cp_gl / dt_glt, "seconds": [dt_gl1, dt_glt]}}


def copy_calibration(dev):
    """A plain streaming kernel over 1 GiB (b = 2 a: read 1 GiB, write 1 GiB),
    run three times,
    for scaling the HBM byte counters of a profiled run."""
    import torch
    a = torch.ones(1 << 27, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    say("copy calibration: 3 x 1 GiB device copies done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "auto", "mixed"])
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-targets", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): it "
              "measures one GPU's share of the transform; sharding is a design note.")
        sys.exit(2)
    import torch
    dev = torch.device("cuda:0")
    if args.copy_calibration:
        copy_calibration(dev)
    res = run_headline(args, dev)
    targets = None
    if not args.no_targets:
        targets = {"batched_vs_plain": res["plain"],
                   "uneven_tol1e-10": run_uneven(dev, 1e-10, 3),
                   "uneven_tol1e-6": run_uneven(dev, 1e-6, 3),
                   "mixed_vs_fp64": run_mixed(dev, 3),
                   "gradient": run_gradient(dev, 3)}
    cpu = None if args.no_cpu else run_cpu_baseline()
    out = {
        "metric": METRIC,
        "value": res["value"],
        "unit": "column-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if res["precision"] == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "nu=5 M=N=1e6 r,k~U[0,1000) B=16 tol=1e-8 fwd+adjoint",
                   "M": res["M"], "N": res["N"], "batch": res["B"], "nu": res["nu"],
                   "tol": res["tol"], "precision": res["precision"]},
        "roofline": res["roofline"],
        "fft_stage": res["fft_stage"],
        "cpu_baseline": cpu,
        "plain": res["plain"],
        "targets": targets,
        "stages_ms_per_step": res["stages_ms"],
        "plan_s": res["plan_s"],
        "first_call_s": res["first_call_s"],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
