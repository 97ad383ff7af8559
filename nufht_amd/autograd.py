"""torch.autograd layer over the device transform.

    g[i, t] = sum_j cs[j, t] * J_nu(k_i * r_j)

``nufht(nu, r, cs, k)`` is differentiable in all three tensors:

* dL/dcs = H^T (dL/dg): the adjoint transform (r and k swapped; J_nu(k r)
  is symmetric in its two node sets), ``nufht_adjoint_device``;
* dL/dr_j = sum_t cs[j, t] sum_i rho[i, t] k_i J'_nu(k_i r_j) and
  dL/dk_i = sum_t rho[i, t] sum_j cs[j, t] r_j J'_nu(k_i r_j), rho = dL/dg,
  from ``nufht_grad_nodes_device`` (J'_nu = (nu/x) J_nu - J_{nu+1}, the
  same derivative as (J_{nu-1} - J_{nu+1}) / 2, each term a fast transform).

The reference exposes no gradient (its transform is the forward sum of
src/fast_hankel_nufht.c:1009-1101 only); these are additions on the same
plan.  Plans are built on the host from the node values, so a change of r
or k re-plans (``NufhtOperator`` caches one plan per node set).

Complex coefficients (complex128 or complex64) are transformed as pairs of
real columns, the kernel being real.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import Plan


def _as_cols(cs: torch.Tensor):
    """(m,) / (m, B) real or complex -> ((m, B') float64 CUDA tensor, restore)."""
    if cs.dim() not in (1, 2):
        raise ValueError("coefficients must be (m,) or (m, B)")
    x = cs if cs.dim() == 2 else cs[:, None]
    squeeze = cs.dim() == 1
    if x.is_complex():
        cdt = x.dtype
        x = torch.view_as_real(x.to(torch.complex128)).reshape(x.shape[0], -1)

        def restore(y):
            y = torch.view_as_complex(y.reshape(y.shape[0], -1, 2).contiguous()).to(cdt)
            return y[:, 0] if squeeze else y
        return x, restore
    if x.dtype != torch.float64:
        raise TypeError("real coefficients must be float64")

    def restore(y):
        return y[:, 0] if squeeze else y
    return x, restore


class _NufhtFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cs2, r, k, plan):
        ctx.plan = plan
        ctx.save_for_backward(cs2)
        return plan.forward(cs2)

    @staticmethod
    def backward(ctx, rho):
        (cs2,) = ctx.saved_tensors
        plan = ctx.plan
        want_c, want_r, want_k = ctx.needs_input_grad[:3]
        rho = rho.to(torch.float64)
        dcs = plan.adjoint(rho) if want_c else None
        dr = dk = None
        if want_r or want_k:
            gr, gk = plan.grad_nodes(cs2, rho, want_r=want_r, want_k=want_k)
            dr, dk = gr, gk
        return dcs, dr, dk, None


class NufhtOperator:
    """The transform for fixed (nu, tol, precision) with a cached plan.

    ``op(r, cs, k)`` plans for the node values of r and k the first time (and
    whenever they change) and returns g = H(r, k) cs as a differentiable
    torch tensor on the device of ``cs``.
    """

    def __init__(self, nu: float, tol: float = 1e-8, precision: str = "auto", **plan_kw):
        self.nu, self.tol, self.precision = float(nu), float(tol), precision
        self.plan_kw = plan_kw
        self._plan: Optional[Plan] = None
        self._key = None

    def plan_for(self, r: torch.Tensor, k: torch.Tensor) -> Plan:
        rh = np.ascontiguousarray(r.detach().to("cpu", torch.float64).numpy())
        kh = np.ascontiguousarray(k.detach().to("cpu", torch.float64).numpy())
        key = (rh.tobytes(), kh.tobytes())
        if self._plan is None or key != self._key:
            # the previous plan is not closed here: a recorded graph node may
            # still hold it for its backward; it is freed when the last
            # reference (this operator's or a node's) goes away
            self._plan = Plan(self.nu, rh, kh, tol=self.tol, precision=self.precision,
                              **self.plan_kw)
            self._key = key
        return self._plan

    def __call__(self, r: torch.Tensor, cs: torch.Tensor, k: torch.Tensor) -> torch.Tensor:
        if r.dim() != 1 or k.dim() != 1:
            raise ValueError("r and k must be 1-D")
        if cs.shape[0] != r.shape[0]:
            raise ValueError(f"cs has {cs.shape[0]} rows, r has {r.shape[0]} nodes")
        if not cs.is_cuda:
            raise TypeError("cs must be a CUDA (HIP) tensor")
        plan = self.plan_for(r, k)
        x, restore = _as_cols(cs)
        rd = r if r.device == x.device else r.to(x.device)
        kd = k if k.device == x.device else k.to(x.device)
        g = _NufhtFn.apply(x, rd.to(torch.float64), kd.to(torch.float64), plan)
        return restore(g)

    def close(self):
        """Drops the operator's plan (graph nodes that still hold it keep it
        alive until their backward has run)."""
        self._plan = None
        self._key = None


def nufht(nu: float, r: torch.Tensor, cs: torch.Tensor, k: torch.Tensor, tol: float = 1e-8,
          precision: str = "auto") -> torch.Tensor:
    """Differentiable g = sum_j cs[j, :] J_nu(k r_j) (plans on every call; hold a
    ``NufhtOperator`` to reuse plans)."""
    return NufhtOperator(nu, tol, precision)(r, cs, k)
