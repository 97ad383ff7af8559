"""Application examples of the transform, after the reference's notebook
(notebooks/nufht_batch_example.ipynb), run on the GPU library.

    python -m nufht_amd.examples fresnel [--loop]
        axisymmetric Fresnel lensing integral F(w, y) of a cored isothermal
        halo (notebook cells 12-13, arXiv:2210.05658 Fig. 2D): 100 values of
        w in [0.1, 10], each its own order-0 transform of a 1000-point
        Gauss-Legendre rule on [0, u_max(w)] to 150 values of y in [0, 1],
        real and imaginary parts as two columns.  The 100 transforms have
        different nodes, so they run as one multi-transform batch (one launch
        per stage); --loop runs the notebook's per-w _pynufht.nufht_batch loop.

    python -m nufht_amd.examples xi
        matter correlation function xi(r) from P(k) = k T(k)^2 with the BBKS
        transfer function, for five values of Omega_m (notebook cell 14):
        xi(r) = (2 pi)^(-3/2) r^(-1/2) int dln k k^(5/2) P(k) J_(1/2)(k r),
        a 10000-point Gauss-Legendre rule in ln k on [-9, 2], 200 values of r
        log-spaced on [0.1, 2000], the five spectra as five columns.

Both print the wall time of the transform stage and a few values.
"""
from __future__ import annotations

import argparse
import sys
import time
from typing import Dict

import numpy as np

from .workloads import gauss_legendre


# ------------------------------------------------------------- Fresnel integral
def halo_potential(x: np.ndarray, xc: float = 0.05) -> np.ndarray:
    """Lensing potential of the cored isothermal halo:
    psi(x) = sqrt(x^2 + xc^2) + xc ln(2 xc / (sqrt(x^2 + xc^2) + xc))."""
    q = np.sqrt(x * x + xc * xc)
    return q + xc * np.log(2.0 * xc / (q + xc))


def fresnel_columns(w: float, n_gl: int = 1000):
    """Nodes u_j on [0, u_max] (u_max = w R, R = sqrt(n_gl / (2 w))) and the two
    real columns (Re, Im) of c_j = window(u_j) xi_j u_j e^{i u_j^2 / 2w}
    (e^{-i w psi(u_j / w) taper(u_j / w)} - 1), as the notebook builds them."""
    x0, w0 = gauss_legendre(n_gl)
    R = np.sqrt(n_gl / (2.0 * w))
    umax = w * R
    u = 0.5 * umax * (x0 + 1.0)
    xi = 0.5 * umax * w0
    du = max(0.02 * umax, 1e-12)
    window = 0.5 * (1.0 - np.tanh((u - umax) / du))
    x = u / w
    taper = 0.5 * (1.0 - np.tanh(x - 0.75 * (umax / w)))
    psi = halo_potential(x) * taper
    c = window * xi * u * np.exp(1j * u * u / (2.0 * w)) * (np.exp(-1j * w * psi) - 1.0)
    return u, np.column_stack([c.real, c.imag])


def fresnel(w_grid=None, y_grid=None, n_gl: int = 1000, loop: bool = False,
            tol: float = 1e-8) -> Dict[str, object]:
    """F(w, y) = 1 + e^{i w y^2 / 2} / (i w) int_0^umax c(u) J_0(u y) du."""
    if w_grid is None:
        w_grid = np.linspace(0.1, 10.0, 100)
    if y_grid is None:
        y_grid = np.linspace(0.0, 1.0, 150)
    w_grid = np.asarray(w_grid, dtype=np.float64)
    y_grid = np.asarray(y_grid, dtype=np.float64)
    cols = [fresnel_columns(float(w), n_gl) for w in w_grid]
    G = np.empty((len(w_grid), len(y_grid)), dtype=np.complex128)
    if loop:
        from . import load_pynufht
        pn = load_pynufht()
        t0 = time.perf_counter()
        for q, (u, cs) in enumerate(cols):
            g = pn.nufht_batch(0.0, u, np.asfortranarray(cs), y_grid, tol=tol)
            G[q] = g[:, 0] + 1j * g[:, 1]
        dt = time.perf_counter() - t0
    else:
        import torch
        from . import MultiBatch, Plan
        dev = torch.device("cuda")
        plans = [Plan(0.0, u, y_grid, tol=tol) for u, _ in cols]
        cs_d = [torch.as_tensor(np.ascontiguousarray(cs.T), device=dev) for _, cs in cols]
        gs_d = [torch.empty((2, y_grid.size), dtype=torch.float64, device=dev) for _ in cols]
        mb = MultiBatch(plans, [2] * len(cols))
        mb.execute(cs_d, gs_d, use_graph=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mb.execute(cs_d, gs_d, use_graph=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        for q, g in enumerate(gs_d):
            gh = g.cpu().numpy()
            G[q] = gh[0] + 1j * gh[1]
        mb.close()
        for p in plans:
            p.close()
    pref = np.exp(1j * np.outer(w_grid, y_grid ** 2) / 2.0) / (1j * w_grid[:, None])
    return {"w": w_grid, "y": y_grid, "G": G, "F": 1.0 + pref * G, "seconds": dt,
            "nodes": [u for u, _ in cols], "columns": [cs for _, cs in cols]}


# ------------------------------------------------------- correlation function
def bbks_transfer(k: np.ndarray, omega_m: float) -> np.ndarray:
    """BBKS transfer function with q = k / Omega_m."""
    q = k / omega_m
    return (np.log(1.0 + 2.34 * q) / (2.34 * q)) * \
        (1.0 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25


def xi_columns(omega_m=(0.08, 0.09, 0.10, 0.11, 0.12), n_gl: int = 10000,
               lnk=(-9.0, 2.0)):
    """Nodes k_j (GL in ln k) and columns w_j k_j^(5/2) P(k_j), P = k T^2."""
    x0, w0 = gauss_legendre(n_gl)
    a, b = lnk
    s = 0.5 * (b - a) * (x0 + 1.0) + a
    ws = 0.5 * (b - a) * w0
    k = np.exp(s)
    cs = np.empty((n_gl, len(omega_m)))
    for i, om in enumerate(omega_m):
        T = bbks_transfer(k, om)
        cs[:, i] = ws * k ** 2.5 * (k * T * T)
    return k, cs


def correlation(r_grid=None, omega_m=(0.08, 0.09, 0.10, 0.11, 0.12), n_gl: int = 10000,
                tol: float = 1e-8) -> Dict[str, object]:
    """xi(r) for each Omega_m through one batched order-1/2 transform."""
    from . import load_pynufht
    if r_grid is None:
        r_grid = np.logspace(-1, np.log10(2000.0), 200)
    r_grid = np.asarray(r_grid, dtype=np.float64)
    k, cs = xi_columns(omega_m, n_gl)
    pn = load_pynufht()
    pn.nufht_batch(0.5, k, np.asfortranarray(cs), r_grid, tol=tol)   # device set-up
    t0 = time.perf_counter()
    G = pn.nufht_batch(0.5, k, np.asfortranarray(cs), r_grid, tol=tol)
    dt = time.perf_counter() - t0
    xi = G / np.sqrt(r_grid)[:, None] / (2.0 * np.pi) ** 1.5
    return {"r": r_grid, "omega_m": np.asarray(omega_m), "G": G, "xi": xi, "seconds": dt,
            "k": k, "cs": cs}


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m nufht_amd.examples")
    sub = ap.add_subparsers(dest="cmd", required=True)
    f = sub.add_parser("fresnel")
    f.add_argument("--loop", action="store_true")
    sub.add_parser("xi")
    args = ap.parse_args(argv)
    if args.cmd == "fresnel":
        out = fresnel(loop=args.loop)
        F = out["F"]
        print(f"F(w, y): {F.shape[0]} w x {F.shape[1]} y, transforms "
              f"{out['seconds'] * 1e3:.2f} ms ({'per-w loop' if args.loop else 'one batch'})")
        for iw in (0, 49, 99):
            print(f"  w={out['w'][iw]:.2f}: |F| at y=0, 0.5, 1: "
                  f"{abs(F[iw, 0]):.6f} {abs(F[iw, 75]):.6f} {abs(F[iw, -1]):.6f}")
    else:
        out = correlation()
        print(f"xi(r): {out['xi'].shape[0]} r x {out['xi'].shape[1]} Omega_m, transform "
              f"{out['seconds'] * 1e3:.2f} ms")
        for ir in (0, 100, 199):
            print(f"  r={out['r'][ir]:.3g}: xi = " +
                  " ".join(f"{v:.6e}" for v in out["xi"][ir]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
