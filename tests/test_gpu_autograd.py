"""The torch.autograd layer: forward against the oracle's direct sum and all
three gradients (coefficients, r nodes, k nodes) against exact analytic
values from FP64 direct sums, including unsorted nodes, negative odd order,
complex coefficients and plan reuse."""
import numpy as np
import pytest

import oracle
from nufht_amd import workloads as W
from tests.conftest import rel_l2, say

pytestmark = pytest.mark.gpu


def _bessel_outer(nu, k, r):
    """J_nu(k_i r_j) as an (n, m) matrix through the oracle's J (signed order)."""
    n, m = k.size, r.size
    J = oracle.besselj_wr(abs(nu), np.repeat(k, m), np.tile(r, n)).reshape(n, m)
    return -J if (nu < 0 and int(round(nu)) % 2) else J


def _exact_grads(nu, r, k, c, rho):
    """(dL/dc, dL/dr, dL/dk) for rho = dL/dg from direct sums, with
    J'_nu(x) = (|nu|/x) J_|nu|(x) - J_{|nu|+1}(x) times the order's sign."""
    n, m = k.size, r.size
    sgn = -1.0 if (nu < 0 and int(round(nu)) % 2) else 1.0
    kk, rr = np.repeat(k, m), np.tile(r, n)
    x = (kk * rr).reshape(n, m)
    Jn = oracle.besselj_wr(abs(nu), kk, rr).reshape(n, m)
    Jn1 = oracle.besselj_wr(abs(nu) + 1, kk, rr).reshape(n, m)
    dJ = sgn * (np.where(x > 0, abs(nu) / np.where(x > 0, x, 1.0), 0.0) * Jn - Jn1)
    dc = (sgn * Jn).T @ rho
    dr = np.einsum("jt,it,i,ij->j", c, rho, k, dJ)
    dk = np.einsum("it,jt,j,ij->i", rho, c, r, dJ)
    return dc, dr, dk


@pytest.mark.parametrize("nu", [2.0, -1.0, 0.5])
def test_autograd_matches_exact_gradients(device, nu):
    import torch
    from nufht_amd.autograd import NufhtOperator
    s = W.Stream(70 + int(2 * abs(nu)), 7)
    m, n, B = 1500, 1200, 3
    r = s.uniform(m, 0.0, 120.0)           # unsorted on purpose
    k = s.uniform(n, 0.0, 100.0)
    c = s.normal(m * B).reshape(m, B)
    y = s.normal(n * B).reshape(n, B)
    rt = torch.tensor(r, device=device, requires_grad=True)
    kt = torch.tensor(k, device=device, requires_grad=True)
    ct = torch.tensor(c, device=device, requires_grad=True)
    op = NufhtOperator(nu, tol=1e-10, precision="fp64")
    say(f"autograd nu={nu} m={m} n={n} B={B}")
    g = op(rt, ct, kt)
    assert g.shape == (n, B)
    ref = oracle.direct(nu, r, c, k)
    assert rel_l2(g.detach().cpu().numpy(), ref) < 1e-9
    loss = ((g - torch.as_tensor(y, device=device)) ** 2).sum()
    loss.backward()
    rho = 2.0 * (ref - y)
    dc, dr, dk = _exact_grads(nu, r, k, c, rho)
    assert rel_l2(ct.grad.cpu().numpy(), dc) < 1e-8
    assert rel_l2(rt.grad.cpu().numpy(), dr) < 1e-8
    assert rel_l2(kt.grad.cpu().numpy(), dk) < 1e-8
    # the cached plan is reused for the same nodes and rebuilt for new ones
    p0 = op._plan
    op(rt, ct, kt)
    assert op._plan is p0
    op(rt.detach() * 1.01, ct, kt)
    assert op._plan is not p0
    op.close()


def test_autograd_complex_coefficients(device):
    import torch
    from nufht_amd.autograd import nufht
    s = W.Stream(81, 8)
    m, n, B = 2000, 1500, 2
    r = np.sort(s.uniform(m, 0.0, 150.0))
    k = np.sort(s.uniform(n, 0.0, 80.0))
    cre, cim = s.normal(m * B).reshape(m, B), s.normal(m * B).reshape(m, B)
    c = torch.tensor(cre + 1j * cim, dtype=torch.complex64, device=device)
    say("autograd complex64 nu=1")
    g = nufht(1.0, torch.as_tensor(r), c, torch.as_tensor(k), tol=1e-8, precision="fp64")
    assert g.dtype == torch.complex64 and g.shape == (n, B)
    # the complex64 inputs carry their own rounding: compare with the direct
    # sum of the rounded coefficients
    c64 = c.cpu().numpy().astype(np.complex128)
    ref = oracle.direct(1.0, r, c64.real.copy(), k) + 1j * oracle.direct(1.0, r, c64.imag.copy(), k)
    assert rel_l2(g.cpu().numpy().astype(np.complex128), ref) < 1e-6
    # gradient of a real loss of a complex output: dL/dc = H^T conj-free, per part
    cc = c.clone().requires_grad_(True)
    g2 = nufht(1.0, torch.as_tensor(r), cc, torch.as_tensor(k), tol=1e-8, precision="fp64")
    (g2.real.sum() + 2.0 * g2.imag.sum()).backward()
    ones = np.ones((n, B))
    adj = oracle.direct(1.0, k, ones, r)
    exp = adj + 1j * 2.0 * adj
    assert rel_l2(cc.grad.cpu().numpy().astype(np.complex128), exp) < 1e-6
