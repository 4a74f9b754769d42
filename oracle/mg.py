"""CPU statement of rheon_amd's multigrid solver (TEST INFRASTRUCTURE ONLY).

Multigrid is not in the reference (it has only PCG, src/pressure.rs); this
numpy module states the algorithm rheon_amd implements so the device V-cycle
can be checked cycle by cycle, and its convergence studied, on the CPU:

* levels: every axis halves while all axes are even and the level has more
  than 64 cells (or a requested level count); spacing doubles; the operator
  is the box operator of src/operator.rs:58-97 rediscretised on each level
  (weights A_d/(rho h_d)), Neumann walls, applied as in :400-416.
* V(nu1, nu2); sweep k of a side is x <- x + tau_k (b - K x)/D with
  Chebyshev weights (default, V(3,1)): tau_k = 1/lambda_k, lambda_k the roots
  of the degree-nu Chebyshev polynomial mapped to [lo, 2] (lo = 0.3 for
  nu <= 2, else 0.25), or weighted Jacobi (tau_k = omega = 6/7, V(2,2));
  the coarse levels may take their own sweep counts (coarse_pre,
  coarse_post; default the finest level's);
  first pre-sweep from zero x = tau_1 b/D; restriction = sum of the 8
  children's residuals; trilinear cell-centred prolongation (1D weights 3/4,
  1/4; walls mirror the parent); coarsest level solved exactly on the
  mean-free right-hand side.
* outer loop: p += V-cycle(b - K p) in f64 with the reference's acceptance
  rule (l2 <= max(abs, rel ||b||) and max|r| dt/V <= divergence limit); the
  returned pressure is shifted so p[0] = 0.

Parity pinning: the operator is checked against the oracle's apply_full
(bit-identical formula), and the solver's converged pressure against the
oracle's jacobi-pcg-v1 at tight tolerance, and the Chebyshev weights against
the closed form of the Chebyshev polynomial's extrema (tests/test_mg.py).
"""
from __future__ import annotations

import numpy as np


class Level:
    def __init__(self, counts, spacing, density):
        self.n = tuple(int(c) for c in counts)  # (nx, ny, nz)
        h = spacing
        area = (h[1] * h[2], h[0] * h[2], h[0] * h[1])
        self.w = tuple(area[d] * (1.0 / (density * h[d])) for d in range(3))
        nx, ny, nz = self.n
        # diagonal = sum of existing-neighbour weights, shape (nz, ny, nx)
        D = np.zeros((nz, ny, nx))
        for ax, (m, wd) in enumerate(zip((nx, ny, nz), self.w)):
            cnt = np.full(m, 2.0)
            cnt[0] -= 1
            cnt[-1] -= 1
            shape = [1, 1, 1]
            shape[2 - ax] = m
            D = D + wd * cnt.reshape(shape)
        self.D = D

    def apply(self, x):
        """K x for an array shaped (nz, ny, nx)."""
        y = np.zeros_like(x)
        for ax, wd in enumerate(self.w):
            a = 2 - ax
            diff = np.diff(x, axis=a)  # x[i+1] - x[i]
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[a] = slice(0, -1)
            hi[a] = slice(1, None)
            y[tuple(lo)] -= wd * diff
            y[tuple(hi)] += wd * diff
        return y


def build_levels(counts, spacing, density=1.0, levels=0, coarsest_cells=64):
    lv = [Level(counts, spacing, density)]
    h = list(spacing)
    n = list(counts)
    while True:
        if levels and len(lv) >= levels:
            break
        even = all(m % 2 == 0 for m in n)
        if not even or min(n) < 4:
            break
        if not levels and n[0] * n[1] * n[2] <= coarsest_cells:
            break
        n = [m // 2 for m in n]
        h = [2 * x for x in h]
        lv.append(Level(n, h, density))
    return lv


def restrict(r):
    nz, ny, nx = r.shape
    return r.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2).sum(axis=(1, 3, 5))


def prolong(xc, fine_shape):
    """Trilinear cell-centred interpolation with mirrored walls."""
    out = xc
    for a in range(3):
        m = out.shape[a]
        idx = np.arange(2 * m)
        parent = idx // 2
        nb = parent + np.where(idx % 2 == 1, 1, -1)
        nb = np.where((nb < 0) | (nb >= m), parent, nb)
        out = 0.75 * np.take(out, parent, axis=a) + 0.25 * np.take(out, nb, axis=a)
    assert out.shape == fine_shape
    return out


def coarsest_solve(L: Level, b):
    n = b.size
    nz, ny, nx = b.shape
    A = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        A[:, i] = L.apply(e.reshape(b.shape)).ravel()
    rhs = b.ravel() - b.mean()
    x = np.linalg.lstsq(A, rhs, rcond=None)[0]
    return (x - x.mean()).reshape(b.shape)


def sweep_weights(n, smoother="chebyshev", omega=6.0 / 7.0):
    """Weights tau_1..tau_n of one side's sweeps."""
    if smoother == "jacobi":
        return [omega] * n
    lo, hi = (0.3 if n <= 2 else 0.25), 2.0
    c, h = 0.5 * (hi + lo), 0.5 * (hi - lo)
    return [1.0 / (c + h * np.cos(np.pi * (2 * k + 1) / (2 * n))) for k in range(n)]


def vcycle(lv, l, b, tau_pre, tau_post, dtype=np.float64, tau_cpre=None, tau_cpost=None):
    """V-cycle from level l; levels below the finest use (tau_cpre,
    tau_cpost) when given."""
    L = lv[l]
    if l == len(lv) - 1:
        return coarsest_solve(L, b.astype(np.float64)).astype(dtype)
    tp = tau_pre if l == 0 or tau_cpre is None else tau_cpre
    tq = tau_post if l == 0 or tau_cpost is None else tau_cpost
    x = (tp[0] * b / L.D).astype(dtype)
    for t in tp[1:]:
        x = (x + t * (b - L.apply(x)) / L.D).astype(dtype)
    bc = restrict(b - L.apply(x)).astype(dtype)
    xc = vcycle(lv, l + 1, bc, tau_pre, tau_post, dtype, tau_cpre, tau_cpost)
    x = (x + prolong(xc, x.shape)).astype(dtype)
    for t in tq:
        x = (x + t * (b - L.apply(x)) / L.D).astype(dtype)
    return x


def solve(counts, spacing, density, rhs, dt, rel, abs_tol, div_limit, max_cycles=100, levels=0,
          pre=0, post=0, omega=6.0 / 7.0, dtype=np.float64, history=None, smoother="chebyshev",
          coarse_pre=0, coarse_post=0):
    """Defect correction with V-cycles; returns (status, p, cycles, l2, rmax)."""
    jac = smoother == "jacobi"
    pre = pre or (2 if jac else 3)
    post = post or (2 if jac else 1)
    coarse_pre, coarse_post = coarse_pre or pre, coarse_post or post
    tau_pre, tau_post = sweep_weights(pre, smoother, omega), sweep_weights(post, smoother, omega)
    tau_cpre = sweep_weights(coarse_pre, smoother, omega)
    tau_cpost = sweep_weights(coarse_post, smoother, omega)
    lv = build_levels(counts, spacing, density, levels)
    shape = (counts[2], counts[1], counts[0])
    b = np.asarray(rhs, np.float64).reshape(shape)
    vol = spacing[0] * spacing[1] * spacing[2]
    threshold = max(abs_tol, rel * float(np.sqrt(np.sum(b * b))))
    p = np.zeros(shape)
    for cycle in range(0, max_cycles + 1):
        r = b - lv[0].apply(p)
        l2 = float(np.sqrt(np.sum(r * r)))
        rmax = float(np.max(np.abs(r)))
        if history is not None:
            history.append(l2)
        if l2 <= threshold and rmax * dt / vol <= div_limit:
            p = p - p.flat[0]
            return "ok", p.ravel(), cycle, l2, rmax
        if cycle == max_cycles:
            break
        p = p + vcycle(lv, 0, r.astype(dtype), tau_pre, tau_post, dtype, tau_cpre, tau_cpost).astype(np.float64)
    p = p - p.flat[0]
    return "iteration_limit", p.ravel(), max_cycles, l2, rmax
