"""Write the committed golden fixtures under tests/golden/.

Run in the build container (where the reference tree and mpmath exist):

    python tests/golden/make_golden.py

Fixtures written (all small, data only):

* ``bessel_golden.npz``     J_nu(x) from mpmath 1.3.0 at 30 digits for the
                            orders the configs use; stands in for GSL's
                            Jnu/Jn, which the reference calls at
                            src/expansions.c:71, :89, :161, :177, :285, :301.
* ``ref_tables_slice.npz``  rows nu = 0..40 of the reference's own plan-table
                            cache data/nufht_tables.cache (layout
                            src/fast_hankel_nufht.c:186-247).
* ``gl_sample.npz``         every 37th row of data/gl10000 and all of
                            data/gl1000 (Gauss-Legendre rules the reference
                            ships).
* ``ref_notebook.json``     the reference's recorded nufht_batch output,
                            notebooks/nufht_batch_example.ipynb:148-218.
* ``ref_boxes.json``        plan parameters and box partitions produced by the
                            reference's own planner (oracle/_ref, compiled from
                            src/fast_hankel_nufht.c + src/bounds.c) on inputs
                            regenerated from nufht_amd.workloads seeds.
"""
from __future__ import annotations

import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)

from nufht_amd import workloads as W  # noqa: E402

BESSEL_ORDERS = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 20, 0.5, 1.5, 2.5, 5.5, 9.5]


def bessel():
    import mpmath
    mpmath.mp.dps = 30
    s = W.Stream(99, 99)
    xs = np.concatenate([np.logspace(-6, 6, 120), s.uniform(80, 0.0, 60.0),
                         s.uniform(40, 60.0, 2000.0)])
    xs = np.sort(xs)
    nus, xv, jv = [], [], []
    for nu in BESSEL_ORDERS:
        for x in xs:
            nus.append(nu)
            xv.append(x)
            jv.append(float(mpmath.besselj(nu, mpmath.mpf(float(x)))))
    np.savez_compressed(os.path.join(HERE, "bessel_golden.npz"), nu=np.array(nus),
                        x=np.array(xv), j=np.array(jv))


def read_cache(path):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:24], dtype=np.uint32)
    off = 24
    a = np.frombuffer(raw[off:off + 401 * 64 * 8], np.float64).reshape(401, 64)
    off += 401 * 64 * 8
    z = np.frombuffer(raw[off:off + 401 * 32 * 16 * 8], np.float64).reshape(401, 32, 16)
    off += 401 * 32 * 16 * 8
    k = np.frombuffer(raw[off:off + 401 * 32 * 16 * 4], np.int32).reshape(401, 32, 16)
    return hdr, a, z, k


def tables():
    hdr, a, z, k = read_cache(os.path.join(REF, "data", "nufht_tables.cache"))
    np.savez_compressed(os.path.join(HERE, "ref_tables_slice.npz"), header=hdr,
                        asy_coef=a[:81], asy_z=z[:41, :12, :11], wimp_k=k[:41, :12, :11])


def gl():
    g10k = np.loadtxt(os.path.join(REF, "data", "gl10000"))
    g1k = np.loadtxt(os.path.join(REF, "data", "gl1000"))
    idx = np.arange(0, 10000, 37)
    np.savez_compressed(os.path.join(HERE, "gl_sample.npz"), gl10000_idx=idx,
                        gl10000=g10k[idx], gl1000=g1k)


def notebook():
    nb = json.load(open(os.path.join(REF, "notebooks", "nufht_batch_example.ipynb")))
    out = None
    for c in nb["cells"]:
        for o in c.get("outputs", []):
            if "text" in o and "max abs err per batch" in "".join(o["text"]):
                out = "".join(o["text"])
    lines = out.splitlines()
    abs_err = [float(v) for v in lines[0].split("[")[1].split("]")[0].split()]
    rows = []
    for ln in lines[3:8]:
        rows.append([float(v) for v in ln.replace("[", " ").replace("]", " ").split()])
    json.dump({"source": "notebooks/nufht_batch_example.ipynb:148-218",
               "workload": {"nu": 0.0, "gl": 1000, "rmax": 10.0, "nk": 100, "kmax": 10.0,
                            "scales": [1.0, 0.5], "tol": 1e-8,
                            "cs": "r*exp(-r^2/2)*w_gl*scale",
                            "closed_form": "exp(-k^2/2)*scale"},
               "max_abs_err_per_column": abs_err,
               "g_first5": rows},
              open(os.path.join(HERE, "ref_notebook.json"), "w"), indent=1)


def boxes():
    import oracle
    oracle.build()
    L = oracle.load_ref_planner()
    if L is None:
        raise SystemExit("oracle/_ref not built")
    os.environ["NUFHT_TABLE_CACHE"] = os.path.join(REF, "data", "nufht_tables.cache")
    cases = []
    # (name, nu, tol, generator)
    def uni(seed, m, n, R, K):
        s = W.Stream(seed, 77)
        return np.sort(s.uniform(m, 0, R)), np.sort(s.uniform(n, 0, K))

    g = W.gaussian_gl()
    specs = [("gaussian_gl", 0.0, 1e-8, ("gaussian_gl",)),
             ("uni_small_nu0", 0.0, 1e-8, ("uni", 11, 3000, 2000, 100.0, 100.0)),
             ("uni_small_nu5", 5.0, 1e-8, ("uni", 12, 4000, 3000, 300.0, 200.0)),
             ("uni_nu1_tol6", 1.0, 1e-6, ("uni", 13, 5000, 5000, 500.0, 100.0)),
             ("uni_nu2_tol10", 2.0, 1e-10, ("uni", 14, 2500, 6000, 50.0, 400.0)),
             ("uni_nu10_tol10", 10.0, 1e-10, ("uni", 15, 6000, 3000, 200.0, 200.0)),
             ("uni_nu3_tol12", 3.0, 1e-12, ("uni", 16, 3000, 3000, 100.0, 100.0)),
             ("uni_nu40", 40.0, 1e-8, ("uni", 17, 2000, 2000, 100.0, 100.0)),
             ("uni_half", 1.5, 1e-8, ("uni", 18, 1000, 800, 30.0, 30.0)),
             ("tiny_direct", 0.0, 1e-8, ("uni", 19, 50, 100, 10.0, 10.0)),
             ("all_loc", 0.0, 1e-8, ("uni", 20, 400, 300, 1.0, 1.0)),
             ("all_asy", 0.0, 1e-8, ("uni_off", 21, 400, 300)),
             ("big_nu5", 5.0, 1e-8, ("uni", 22, 200000, 200000, 1000.0, 1000.0))]
    for name, nu, tol, gen in specs:
        if gen[0] == "gaussian_gl":
            r, k = g["r"], g["k"]
        elif gen[0] == "uni":
            r, k = uni(*gen[1:])
        else:
            s = W.Stream(gen[1], 77)
            r = np.sort(s.uniform(gen[2], 100.0, 200.0))
            k = np.sort(s.uniform(gen[3], 100.0, 200.0))
        opt = L.nufht_default_options()
        opt.tol = tol
        plan = oracle.RefPlan()
        rc = L.nufht_plan_init(ctypes.byref(plan), nu, r.ctypes.data_as(oracle._DP), r.size,
                               k.ctypes.data_as(oracle._DP), k.size, ctypes.byref(opt))
        entry = {"name": name, "nu": nu, "tol": tol, "gen": list(gen), "rc": rc,
                 "m": int(r.size), "n": int(k.size)}
        if rc == 0:
            bx = oracle.ref_boxes_to_arrays(plan.boxes)
            entry.update(z_split=plan.z_split, K_asy=plan.K_asy, K_loc=plan.K_loc,
                         asy_coef=list(plan.asy_coef[:24]),
                         boxes={kk: v.tolist() for kk, v in bx.items()})
            L.nufht_plan_free(ctypes.byref(plan))
        cases.append(entry)
    # parameter-only cases, including the error codes
    params = []
    for nu, tol in [(0, 1e-8), (5, 1e-8), (1, 1e-6), (1, 1e-10), (0, 1e-10), (2, 1e-10),
                    (10, 1e-10), (2.5, 1e-6), (9.5, 1e-8), (-3, 1e-8), (27, 1e-8), (40, 1e-8),
                    (200, 1e-15), (0.3, 1e-8), (201, 1e-8), (10.5, 1e-8), (1, 1e-16),
                    (0, 1e-4), (0, 2e-3), (7, 3e-9)]:
        r = np.array([1.0, 2.0])
        opt = L.nufht_default_options()
        opt.tol = tol
        plan = oracle.RefPlan()
        rc = L.nufht_plan_init(ctypes.byref(plan), float(nu), r.ctypes.data_as(oracle._DP), 2,
                               r.ctypes.data_as(oracle._DP), 2, ctypes.byref(opt))
        e = {"nu": nu, "tol": tol, "rc": rc}
        if rc == 0:
            e.update(z_split=plan.z_split, K_asy=plan.K_asy, K_loc=plan.K_loc)
            L.nufht_plan_free(ctypes.byref(plan))
        params.append(e)
    json.dump({"source": "reference planner src/fast_hankel_nufht.c:322-414, :578-744 "
                         "(oracle/_ref, tables from data/nufht_tables.cache)",
               "cases": cases, "params": params},
              open(os.path.join(HERE, "ref_boxes.json"), "w"))


if __name__ == "__main__":
    bessel()
    tables()
    gl()
    notebook()
    boxes()
    print("fixtures written to", HERE)
