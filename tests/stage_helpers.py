"""One SIMPLEC iteration (Source/Simplec.cpp:452-506) composed from the
oracle (TEST INFRASTRUCTURE): the residual and the implicit rows of
oracle/physics.py, the momentum predictor, pressure correction (an exact
sparse solve) and scalar sweep composites of tests/simplec_helpers.py, in
the stage's order. The state has two ghost layers (begin -2); the
composites take one (begin -1)."""
from __future__ import annotations

import math

import numpy as np

from oracle import oracle as orc
from oracle import physics as P
from tests import simplec_helpers as SH

B1, B2 = (-1, -1, -1), (-2, -2, -2)
FIELDS = ("u", "v", "w", "p", "theta", "tke")


def one(a):
    return np.ascontiguousarray(a[:, 1:-1, 1:-1, 1:-1])


def put_valid(q, name, a1):
    """Copy the valid points of a one-ghost array into the two-ghost state."""
    q[name][0, 2:-2, 2:-2, 2:-2] = a1[0, 1:-1, 1:-1, 1:-1]


def coriolis_f(f: P.Forcing):
    if not f.coriolis:
        return 0.0
    return 4 * math.pi * math.sin(f.latitude * math.pi / 180) / f.rotation_period


def poisson_bc(m, phys: P.Physics):
    """The correction's boundary types (Source/BoundaryConditions.cpp:300-320,
    ERF's get_fft_bc): periodic where the geometry is, Dirichlet on outflow
    faces, Neumann elsewhere."""
    out = []
    for o, t in enumerate(phys.bc.faces()):
        d = o % 3
        if d < 2 and m.geom.periodic[d]:
            out.append(orc.PERIODIC)
        elif t == P.OUTFLOW:
            out.append(orc.DIRICHLET)
        else:
            out.append(orc.NEUMANN)
    return tuple(out)


def simplec_iteration(tile, m, q, phys: P.Physics, dtau, sweeps=2, alpha_p=1.0, tke_floor=0.0, relax=0.3,
                      w_vnn_cap=0.0, margin=4.0, solve=None):
    """q (dict of two-ghost arrays) is updated in place; returns the residual R(q).
    solve: the pressure step's linear solve (None: sparse direct)."""
    g = m.geom
    nz = g.hi[2] - g.lo[2] + 1
    P.fill_ghosts(q, m, phys)
    R, cf, ss = P.residual(q, m, phys)
    Fx, Fy, Fz, Fb = orc.transport_fluxes(q["u"], q["v"], q["w"], B2, m.z, B1, m.ax, m.ay, B1, g)
    F = [Fx[None], Fy[None], Fz[None]]
    dtf = orc.face_dtau(dtau, g)
    A = [P.implicit_rows(eq, F, Fb, m, phys, cf, ss) for eq in (P.EQ_U, P.EQ_V, P.EQ_W)]
    buoy = phys.forcing.gravity > 0
    At = P.implicit_rows(P.EQ_THETA, F, Fb, m, phys, cf, ss) if buoy else None
    masks = [P.dirichlet_rows(eq, m, phys)[0] for eq in (P.EQ_U, P.EQ_V, P.EQ_W)]
    vel = [one(q[n]) for n in ("u", "v", "w")]
    vel, aps, ds, _, _ = SH.momentum_predictor(tile, vel, [R["u"][0], R["v"][0], R["w"][0]], A, dtf, sweeps,
                                               mask=masks, w_vnn_cap=w_vnn_cap, margin=margin,
                                               coriolis_f=coriolis_f(phys.forcing), relax=relax,
                                               gravity=phys.forcing.gravity, theta0=phys.forcing.theta0,
                                               theta=one(q["theta"]), R_theta=R["theta"][0], A_theta=At)
    for n, a in zip(("u", "v", "w"), vel):
        put_valid(q, n, a)
    P.fill_ghosts(q, m, phys)
    bc = poisson_bc(m, phys)
    u, v, w, p, *_ = SH.pressure_correction(tile, one(q["u"]), one(q["v"]), one(q["w"]), one(q["p"]), ds, bc, alpha_p,
                                            solve=solve)
    for n, a in zip(("u", "v", "w", "p"), (u, v, w, p)):
        put_valid(q, n, a)
    conv0 = orc.convection(0, q["theta"], B2, F, Fb, m.J, B1, g, phys.theta_limiter, phys.theta_eps)
    P.fill_ghosts(q, m, phys)
    Fx, Fy, Fz, Fb = orc.transport_fluxes(q["u"], q["v"], q["w"], B2, m.z, B1, m.ax, m.ay, B1, g)
    F = [Fx[None], Fy[None], Fz[None]]
    conv1 = orc.convection(0, q["theta"], B2, F, Fb, m.J, B1, g, phys.theta_limiter, phys.theta_eps)
    Rt = R["theta"][0] + (conv1 - conv0)
    At = P.implicit_rows(P.EQ_THETA, F, Fb, m, phys, cf, ss)
    th, *_ = SH.scalar_sweep(tile, one(q["theta"]), Rt, At, dtau, sweeps)
    put_valid(q, "theta", th)
    Ak = P.implicit_rows(P.EQ_TKE, F, Fb, m, phys, cf, ss)
    k, *_ = SH.scalar_sweep(tile, one(q["tke"]), R["tke"][0], Ak, dtau, sweeps, floor=tke_floor, clip=True)
    put_valid(q, "tke", k)
    return R, cf
