"""Synthetic inputs for the benchmark and test workloads.

Everything is generated from a SplitMix64 stream (Steele, Lea & Flood 2014)
with Box-Muller normals, so the same seed gives the same inputs on any
machine.  Gauss-Legendre rules are computed by Newton's method on P_n (the
reference ships them as text tables ``data/gl*``; this module reproduces them
to rounding, which ``tests/test_workloads.py`` checks against a committed
sample of ``data/gl10000``).

Workloads (sizes as used by ``bench.py`` and the parity tests):

* ``gaussian_gl``     nu=0, GL10000 on [0, 50], k = 10(j+1)/200, B=4; column 0
                      is w r exp(-r^2/2) (closed form exp(-k^2/2)), columns 1-3
                      are w r f with f ~ N(0,1).
* ``large_uniform``   nu=5, M=N=1e6, r,k ~ U[0,1000), B=16, tol 1e-8.
* ``uneven_batch``    256 transforms, M,N ~ 10^U(3,5), R*K ~ 10^U(2,5)
                      (R = K = sqrt(RK)), nu in {0..10, 1/2, 3/2, 5/2},
                      B ~ U{1..8}.
* ``clustered``       nu=1, M=N=1e6, r = |N(0,400)| truncated to [0,2000],
                      k ~ U[0,500), 16 complex columns (32 real), tol 1e-6.
* ``gradient``        nu=1, M=N=2e5, r,k ~ U[0,300), B=8, tol 1e-10.
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class Stream:
    """SplitMix64: state advances by the golden gamma, outputs are mixed."""

    def __init__(self, seed: int, tag: int = 0):
        with np.errstate(over="ignore"):
            s = _mix(np.array([np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
                               ^ np.uint64((tag * 0x632BE59BD9B4E019) & 0xFFFFFFFFFFFFFFFF)]))
        self.state = s[0]

    def u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * _GOLDEN
            self.state = self.state + np.uint64(n) * _GOLDEN
            return _mix(z)

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        rad = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([rad * np.cos(2 * np.pi * u2), rad * np.sin(2 * np.pi * u2)])
        return z[:n]

    def integers(self, n: int, lo: int, hi: int) -> np.ndarray:
        """Uniform integers in [lo, hi]."""
        return lo + np.floor(self.uniform(n) * (hi - lo + 1)).astype(np.int64)


def gauss_legendre(n: int):
    """Nodes (ascending) and weights of the n-point rule on [-1, 1]."""
    i = np.arange(1, n + 1, dtype=np.float64)
    x = np.cos(np.pi * (i - 0.25) / (n + 0.5))
    for _ in range(100):
        p0 = np.ones_like(x)
        p1 = x.copy()
        for k in range(2, n + 1):
            p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
        dp = n * (x * p1 - p0) / (x * x - 1.0)
        dx = p1 / dp
        x = x - dx
        if np.max(np.abs(dx)) < 1e-16:
            break
    p0 = np.ones_like(x)
    p1 = x.copy()
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    dp = n * (x * p1 - p0) / (x * x - 1.0)
    w = 2.0 / ((1.0 - x * x) * dp * dp)
    order = np.argsort(x)
    return x[order], w[order]


def gaussian_gl(n_gl: int = 10000, rmax: float = 50.0, nk: int = 200, kmax: float = 10.0,
                batch: int = 4, seed: int = 1):
    x, wq = gauss_legendre(n_gl)
    r = 0.5 * rmax * (x + 1.0)
    wq = 0.5 * rmax * wq
    k = kmax * np.arange(1, nk + 1, dtype=np.float64) / nk
    s = Stream(seed, 1)
    cs = np.empty((n_gl, batch))
    cs[:, 0] = wq * r * np.exp(-0.5 * r * r)
    for t in range(1, batch):
        cs[:, t] = wq * r * s.normal(n_gl)
    return dict(nu=0.0, r=r, k=k, cs=cs, tol=1e-8, weights=wq)


def large_uniform(m: int = 1_000_000, n: int = 1_000_000, batch: int = 16, nu: float = 5.0,
                  rmax: float = 1000.0, kmax: float = 1000.0, tol: float = 1e-8, seed: int = 2):
    s = Stream(seed, 2)
    r = np.sort(s.uniform(m, 0.0, rmax))
    k = np.sort(s.uniform(n, 0.0, kmax))
    cs = s.normal(m * batch).reshape(batch, m).T.copy(order="F")
    return dict(nu=nu, r=r, k=k, cs=cs, tol=tol)


def uneven_batch(count: int = 256, seed: int = 3, mmin: float = 1e3, mmax: float = 1e5):
    s = Stream(seed, 3)
    orders = [float(v) for v in range(11)] + [0.5, 1.5, 2.5]
    out = []
    for _ in range(count):
        m = int(round(10 ** s.uniform(1, math.log10(mmin), math.log10(mmax))[0]))
        n = int(round(10 ** s.uniform(1, math.log10(mmin), math.log10(mmax))[0]))
        rk = 10 ** s.uniform(1, 2.0, 5.0)[0]
        R = K = math.sqrt(rk)
        nu = orders[int(s.integers(1, 0, len(orders) - 1)[0])]
        b = int(s.integers(1, 1, 8)[0])
        r = np.sort(s.uniform(m, 0.0, R))
        k = np.sort(s.uniform(n, 0.0, K))
        cs = s.normal(m * b).reshape(b, m).T.copy(order="F")
        out.append(dict(nu=nu, r=r, k=k, cs=cs))
    return out


def clustered(m: int = 1_000_000, n: int = 1_000_000, batch_complex: int = 16,
              seed: int = 4, rmax: float = 2000.0, sd: float = 400.0, kmax: float = 500.0):
    s = Stream(seed, 4)
    r = np.empty(0)
    while r.size < m:
        cand = np.abs(sd * s.normal(2 * m))
        r = np.concatenate([r, cand[cand <= rmax]])
    r = np.sort(r[:m])
    k = np.sort(s.uniform(n, 0.0, kmax))
    re = s.normal(m * batch_complex).reshape(batch_complex, m)
    im = s.normal(m * batch_complex).reshape(batch_complex, m)
    cs = np.empty((m, 2 * batch_complex), order="F")
    cs[:, 0::2] = re.T
    cs[:, 1::2] = im.T
    return dict(nu=1.0, r=r, k=k, cs=cs, tol=1e-6)


def gradient(m: int = 200_000, n: int = 200_000, batch: int = 8, seed: int = 5,
             rmax: float = 300.0, kmax: float = 300.0):
    s = Stream(seed, 5)
    r = np.sort(s.uniform(m, 0.0, rmax))
    k = np.sort(s.uniform(n, 0.0, kmax))
    cstar = s.normal(m * batch).reshape(batch, m).T.copy(order="F")
    noise = Stream(50, 5).normal(n * batch).reshape(batch, n).T.copy(order="F")
    c0 = Stream(51, 5).normal(m * batch).reshape(batch, m).T.copy(order="F")
    return dict(nu=1.0, r=r, k=k, cstar=cstar, noise=noise, c0=c0, tol=1e-10)
