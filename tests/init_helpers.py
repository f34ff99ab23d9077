"""The mass-consistent initializer (Source/Initializer.cpp:344-418) composed
from the oracle (TEST INFRASTRUCTURE): the first guess of
oracle/initializer.py, the velocity divergence and fluxes of the oracle's C
pieces, and the Poisson correction as an exact sparse solve of the oracle's
operator. Also the reference's Askervein initializer deck
(Tests/Initialization/Mass_Consistent_Askervein/inputs_init_askervein)."""
from __future__ import annotations

import os

import numpy as np
import scipy.sparse.linalg as spla

from ablsteadyrans_amd import mesh as M
from oracle import initializer as OI
from oracle import oracle as orc
from oracle import physics as P
from tests import physics_helpers as PHY
from tests import poisson_helpers as H

B1, B2 = (-1, -1, -1), (-2, -2, -2)
GOLD = H.GOLD


def askervein_tile():
    """30 x 30 cells of 200 m, 18 layers from 20 m growing by 1.08 (the
    computational top 440 m), ERF's Askervein grid, no smoothing."""
    d = np.load(os.path.join(GOLD, "askervein_grid.npz"))
    grid = M.ErfGrid(d["x"], d["y"], d["z"])
    lo, hi = (-3000.0, -3000.0, 0.0), (3000.0, 3000.0, 440.0)
    geom = M.geometry((30, 30, 18), lo, hi)
    zl = M.make_zlevels(18, 440.0, dz0=20.0, stretch_ratio=1.08)
    return M.build_mesh(geom, lo, hi, zl, M.Terrain(kind="erf_grid", grid=grid))


# inflow at xlo, outflow at xhi, slip walls at y, the surface layer and a slip top
ASKERVEIN_FACES = (P.INFLOW, P.SLIP_WALL, P.SURFACE_LAYER, P.OUTFLOW, P.SLIP_WALL, P.SLIP_WALL)


def askervein_profile():
    """The log law along x with u* = 0.6 m/s, z0 = 0.03 m; theta 300 K; the
    closure's length cap (turb.max_geom_lscale 30 m, the default)."""
    return OI.InflowProfile(kind=OI.LOG_LAW, ustar=0.6, z0=0.03, direction=0.0, theta_ref_z=(0.0,),
                            theta_ref=(300.0,))


def askervein_options():
    return OI.InitOptions(first_guess=1, alpha_h=1.0, alpha_v=0.5)


def coefficients(tile, opt: OI.InitOptions):
    """Sigma = diag(alpha_h^2, alpha_h^2, alpha_v^2) as the oracle takes
    it: (bx, by, bz) on the faces with one ghost and the split z-face part
    dcross = bz - z_cross (z_cross = alpha_h^2), or None for unit alphas."""
    if opt.alpha_h == 1 and opt.alpha_v == 1:
        return None, None
    ah2, av2 = opt.alpha_h * opt.alpha_h, opt.alpha_v * opt.alpha_v
    bx = np.full(tile.ax.shape, ah2)
    by = np.full(tile.ay.shape, ah2)
    bz = np.full(tile.az.shape, av2)
    dc = (bz - ah2) if opt.alpha_v != opt.alpha_h else None
    return (bx, by, bz), dc


def solve(tile, bc, rhs, coeffs, dcross):
    """The correction's potential: an exact sparse solve (the J-weighted
    mean removed when no face is Dirichlet)."""
    A = H.sparse_operator(tile, bc, coeffs, dcross).tocsr()
    J = tile.detJ[1:-1, 1:-1, 1:-1]
    singular = all(b != orc.DIRICHLET for b in bc)
    r = rhs.ravel().copy()
    if singular:
        r -= np.sum(J.ravel() * r) / np.sum(J)
        A = A.tolil()
        A[0, :] = 0
        A[0, 0] = 1.0
        r[0] = 0.0
        A = A.tocsr()
    phi = spla.spsolve(A.tocsc(), r).reshape(rhs.shape)
    if singular:
        phi -= np.sum(J * phi) / np.sum(J)
    return phi


def initialize_state(tile, q, prof: OI.InflowProfile, opt: OI.InitOptions, faces):
    """Source/Initializer.cpp:344-418 on the two-ghost state q (in place).
    Returns (div_before, div_after, lambda, obrien_top_residual)."""
    m = PHY.omesh(tile)
    g = m.geom
    ob = OI.first_guess_state(q, m, prof, opt)
    if opt.first_guess == 0:
        return 0.0, 0.0, None, ob
    div = orc.velocity_divergence(q["u"], q["v"], q["w"], B2, m.z, B1, m.ax, m.ay, m.az, B1, m.J, B1, g)
    before = float(np.sqrt(np.sum(div * div)))
    if not opt.projection:
        return before, before, None, ob
    bc = OI.poisson_bc(faces, tile.geom.periodic)
    coeffs, dc = coefficients(tile, opt)
    lam = solve(tile, bc, div, coeffs, dc)
    cf = coeffs if coeffs is not None else (None, None, None)
    F = orc.poisson_fluxes(np.pad(lam, 1)[None].copy(), B1, tile.z_nd, B1, H.ogeom(tile), bc, coeffs=cf,
                           coeff_begin=B1, dcross=dc)
    om, omb = orc.omega_from_w(q["u"], q["v"], q["w"], B2, m.z, B1, g, 0)
    q["u"][0, 2:-2, 2:-2, 2:-2] += F[0]
    q["v"][0, 2:-2, 2:-2, 2:-2] += F[1]
    om = om + F[2]
    OI.periodic_fill(q["u"], B2, 1, g)
    OI.periodic_fill(q["v"], B2, 2, g)
    q["w"][0, 2:-2, 2:-2, 2:-2] = orc.w_from_omega(q["u"], q["v"], om[None].copy(), B2, omb, m.z, B1, g)
    div = orc.velocity_divergence(q["u"], q["v"], q["w"], B2, m.z, B1, m.ax, m.ay, m.az, B1, m.J, B1, g)
    after = float(np.sqrt(np.sum(div * div)))
    return before, after, lam, ob


def cell_fields(q):
    """The plotfile's cell fields (Source/StateIO.cpp:47-71): face means of
    u, v, w and the cell values of p, theta, k, each (nz, ny, nx)."""
    u = q["u"][0, 2:-2, 2:-2, 2:-2]
    v = q["v"][0, 2:-2, 2:-2, 2:-2]
    w = q["w"][0, 2:-2, 2:-2, 2:-2]
    c = lambda a: a[0, 2:-2, 2:-2, 2:-2]
    return {"u": 0.5 * (u[:, :, :-1] + u[:, :, 1:]), "v": 0.5 * (v[:, :-1, :] + v[:, 1:, :]),
            "w": 0.5 * (w[:-1] + w[1:]), "p": c(q["p"]), "theta": c(q["theta"]), "tke": c(q["tke"])}


def fcompare(a, b):
    """amrex_fcompare's relative error of one variable: max |a - b| over
    max |b|."""
    s = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (s if s > 0 else 1.0))
