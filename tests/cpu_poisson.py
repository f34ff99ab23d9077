"""TEST INFRASTRUCTURE (used by tests and by bench.py's CPU baseline): a CPU
solve of the oracle's terrain Poisson operator (oracle_poisson_apply) for
tiles too large to factor, for the oracle's SIMPLEC composite
(tests/stage_helpers.py, its `solve` argument): restarted GMRES (scipy) with
a z-line preconditioner, the exact tridiagonal part of each column of the
operator (its entries read off 27 coloured probes: a cell's 15-point
footprint holds at most one cell of each (i, j, k) mod 3 colour), solved by
a batched Thomas recurrence.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from oracle import oracle as orc
from tests import poisson_helpers as PH


def zline_coefficients(m, bc, coeffs):
    """(lower, diag, upper) of every column: A[p, p - k], A[p, p], A[p, p + k]
    of the oracle operator (nz, ny, nx each; lower[0] and upper[-1] unused)."""
    nx, ny, nz = m.geom.n
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    lower, diag, upper = np.zeros((nz, ny, nx)), np.zeros((nz, ny, nx)), np.zeros((nz, ny, nx))
    per = m.geom.periodic
    # across a periodic seam the last column neighbours the first: they must
    # differ in colour, (n - 1) mod 3 != 0
    if (per[0] and nx % 3 == 1) or (per[1] and ny % 3 == 1):
        raise ValueError("the probing needs periodic extents n with n mod 3 != 1")
    for ck in range(3):
        for cj in range(3):
            for ci in range(3):
                probe = ((I % 3 == ci) & (J % 3 == cj) & (K % 3 == ck)).astype(np.float64)
                y = PH.apply(m, probe, bc, coeffs)
                col = (I % 3 == ci) & (J % 3 == cj)
                diag[col & (K % 3 == ck)] = y[col & (K % 3 == ck)]
                sel = col & ((K - 1) % 3 == ck) & (K >= 1)     # the probe one level below
                lower[sel] = y[sel]
                sel = col & ((K + 1) % 3 == ck) & (K <= nz - 2)
                upper[sel] = y[sel]
    return lower, diag, upper


def thomas(lower, diag, upper, r):
    """Solves every column's tridiagonal system at once (k the first axis)."""
    nz = r.shape[0]
    c = np.zeros_like(r)
    d = np.zeros_like(r)
    c[0] = upper[0] / diag[0]
    d[0] = r[0] / diag[0]
    for k in range(1, nz):
        den = diag[k] - lower[k] * c[k - 1]
        c[k] = upper[k] / den
        d[k] = (r[k] - lower[k] * d[k - 1]) / den
    x = np.zeros_like(r)
    x[-1] = d[-1]
    for k in range(nz - 2, -1, -1):
        x[k] = d[k] - c[k] * x[k + 1]
    return x


class LineGMRES:
    """solve(rhs, coeffs, bc) -> phi to rtol (relative to |rhs|), the J-weighted
    mean removed in the singular case; .iterations holds the last count."""

    def __init__(self, mesh, rtol=1e-3, restart=40, maxiter=20):
        self.mesh, self.rtol, self.restart, self.maxiter = mesh, rtol, restart, maxiter
        self.iterations = 0

    def __call__(self, rhs, coeffs, bc):
        m = self.mesh
        shape = rhs.shape
        n = rhs.size
        lo, di, up = zline_coefficients(m, bc, coeffs)
        singular = all(b != orc.DIRICHLET for b in bc)
        J = m.detJ[1:-1, 1:-1, 1:-1]
        r = rhs.copy()
        if singular:
            r = r - np.sum(J * r) / np.sum(J)
        A = spla.LinearOperator((n, n), matvec=lambda x: PH.apply(m, x.reshape(shape), bc, coeffs).ravel())
        Mop = spla.LinearOperator((n, n), matvec=lambda x: thomas(lo, di, up, x.reshape(shape)).ravel())
        count = [0]

        def cb(_):
            count[0] += 1

        x, info = spla.gmres(A, r.ravel(), M=Mop, rtol=self.rtol, atol=0.0, restart=self.restart,
                             maxiter=self.maxiter, callback=cb, callback_type="pr_norm")
        self.iterations = count[0]
        phi = x.reshape(shape)
        if singular:
            phi = phi - np.sum(J * phi) / np.sum(J)
        return phi
