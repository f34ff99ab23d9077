"""Shared helpers of the line-sweep tests: seeded systems in the Array4 layout,
calls of the oracle on them, and the dense references that pin the oracle.

The random M-matrix recipe follows Tests/Unit_Tests/test_LineSweeps.cpp:76-115
(neighbour coefficients in [0.1, 1], the x pair scaled by xscale, coef_P their
sum, a_P = sum + extra, rhs in [-1, 1], the two copies of a periodic seam face
carrying one system).
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as orc

COEF_P, XM, XP, YM, YP, ZM, ZP, NC = range(8)


def valid_box(n, loc):
    nx, ny, nz = n
    return (0, 0, 0), (nx - 1 + (loc & 1), ny - 1 + ((loc >> 1) & 1), nz - 1 + ((loc >> 2) & 1))


def shape_of(lo, hi):
    return (hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)


def random_system(n, loc, periodic, seed, extra=0.3, xscale=3.0):
    """A, ap, rhs over the valid box at loc (no ghosts), float64."""
    lo, hi = valid_box(n, loc)
    nz, ny, nx = shape_of(lo, hi)
    rng = np.random.default_rng(seed)
    A = np.zeros((NC, nz, ny, nx))
    nb = rng.uniform(0.1, 1.0, size=(6, nz, ny, nx))
    nb[0:2] *= xscale
    A[1:] = nb
    A[COEF_P] = nb.sum(axis=0)
    ap = (A[COEF_P] + extra)[None].copy()
    rhs = rng.uniform(-1.0, 1.0, size=(1, nz, ny, nx))
    seam_consistent(A, loc, periodic)
    seam_consistent(ap, loc, periodic)
    seam_consistent(rhs, loc, periodic)
    return A, ap, rhs


def seam_consistent(arr, loc, periodic):
    """OverrideSync of a one-box domain: the high copy of a periodic seam face
    takes the low copy's value (AMReX's default owner mask)."""
    if (loc & 1) and periodic[0]:
        arr[..., -1] = arr[..., 0]
    if ((loc >> 1) & 1) and periodic[1]:
        arr[..., -1, :] = arr[..., 0, :]


def new_delta(n, loc):
    lo, hi = valid_box(n, loc)
    nz, ny, nx = shape_of(lo, hi)
    return np.zeros((1, nz + 2, ny + 2, nx + 2)), (-1, -1, -1)


def geom(n, periodic):
    return orc.geom_c((0, 0, 0), (n[0] - 1, n[1] - 1, n[2] - 1), (periodic[0], periodic[1], 0))


def oracle_jacobi(n, loc, periodic, A, ap, rhs, sweeps):
    d, db = new_delta(n, loc)
    ch = orc.zline_jacobi(loc, A, (0, 0, 0), ap, rhs, (0, 0, 0), sweeps, geom(n, periodic), d, db)
    return d[0, 1:-1, 1:-1, 1:-1].copy(), ch, d


def oracle_lusgs(n, loc, periodic, A, ap, rhs, sweeps, ghosts=0, overlap=0, order=0, alternate=0):
    d, db = new_delta(n, loc)
    ch = orc.zline_lusgs(loc, A, (0, 0, 0), ap, rhs, (0, 0, 0), sweeps, geom(n, periodic), d, db, ghosts, overlap,
                         order, alternate)
    return d[0, 1:-1, 1:-1, 1:-1].copy(), ch, d


# ---------------------------------------------------------------------------
# Dense references (independent of the oracle's code paths).

def neighbours(n, loc, periodic):
    """For every valid point p (flattened i fastest): list of (q, coef, wrapped)
    of its x/y neighbours (q None outside a non-periodic face)."""
    lo, hi = valid_box(n, loc)
    nz, ny, nx = shape_of(lo, hi)
    L = (n[0], n[1])
    out = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                row = []
                for (di, dj, c) in ((-1, 0, XM), (1, 0, XP), (0, -1, YM), (0, 1, YP)):
                    ii, jj, wrapped = i + di, j + dj, False
                    if ii < 0 or ii > nx - 1:
                        if periodic[0]:
                            ii += L[0] if ii < 0 else -L[0]
                            wrapped = True
                        else:
                            ii = None
                    if jj < 0 or jj > ny - 1:
                        if periodic[1]:
                            jj += L[1] if jj < 0 else -L[1]
                            wrapped = True
                        else:
                            jj = None
                    q = None if (ii is None or jj is None) else (k * ny + jj) * nx + ii
                    row.append((q, c, wrapped))
                out.append(row)
    return out, (nz, ny, nx)


def dense_matrices(n, loc, periodic, A, ap):
    """N (full system: ap on the diagonal, -a_nb off it, z couplings inside
    the column only) and its block-diagonal column part M."""
    nbs, (nz, ny, nx) = neighbours(n, loc, periodic)
    npt = nz * ny * nx
    M = np.zeros((npt, npt))
    Nfull = np.zeros((npt, npt))
    Af = A.reshape(NC, -1)
    apf = ap.reshape(-1)
    for p in range(npt):
        k = p // (nx * ny)
        M[p, p] = apf[p]
        if k > 0:
            M[p, p - nx * ny] = -Af[ZM, p]
        if k < nz - 1:
            M[p, p + nx * ny] = -Af[ZP, p]
    Nfull[:] = M
    for p, row in enumerate(nbs):
        for (q, c, _w) in row:
            if q is not None:
                Nfull[p, q] -= Af[c, p]
    return Nfull, M, nbs


def dense_sgs_pass(n, loc, periodic, A, ap, rhs, x_old, order_cols, snapshot):
    """One line Gauss-Seidel pass as a block-triangular dense solve: the
    columns in order_cols (list of (i, j)); a coupling to a column earlier in
    the order takes the new value, to a later one the old value x_old, and a
    coupling across a periodic seam the snapshot (the ghosts filled before
    the pass)."""
    _N, M, nbs = dense_matrices(n, loc, periodic, A, ap)
    lo, hi = valid_box(n, loc)
    nz, ny, nx = shape_of(lo, hi)
    rank = {}
    for r, (i, j) in enumerate(order_cols):
        rank[(i, j)] = r
    Af = A.reshape(NC, -1)
    K = M.copy()
    b = rhs.reshape(-1).copy()
    for p, row in enumerate(nbs):
        i, j = p % nx, (p // nx) % ny
        for (q, c, wrapped) in row:
            if q is None:
                continue
            qi, qj = q % nx, (q // nx) % ny
            if wrapped:
                b[p] += Af[c, p] * snapshot[q]
            elif rank[(qi, qj)] < rank[(i, j)]:
                K[p, q] -= Af[c, p]
            else:
                b[p] += Af[c, p] * x_old[q]
    return np.linalg.solve(K, b)


def pass_orders(nx, ny, forward, jrev):
    if not jrev:
        cols = [(i, j) for j in range(ny) for i in range(nx)]
    else:
        cols = [(i, j) for j in reversed(range(ny)) for i in range(nx)]
    return cols if forward else list(reversed(cols))
