"""The residual R(q) of Source/Residual.cpp:159-256 and the implicit rows of
Derived in part from ERF (BSD 3-Clause, Copyright (c) 2022 The Regents of the
University of California et al.); the full notice is in THIRD_PARTY_NOTICES.md
at the repository root.
Source/Residual.cpp:272-304, composed from the oracle's C pieces.

TEST INFRASTRUCTURE ONLY (see ``oracle.py``): used by tests/ and bench.py's
CPU leg as the checker of the HIP library.

Scope, as the HIP library's: periodic, inflow, outflow, slip or no-slip sides; the ground ``surface_layer``
(neutral, first-centre reference), ``slip_wall`` or ``no_slip_wall``; the top
``slip_wall``, ``no_slip_wall`` or ``geostrophic``; the k-L closure (or given
diffusivities); buoyancy, Coriolis, geostrophic and pressure-gradient
forcing. Arrays are float64 (1, nz', ny', nx'): the state with two ghosts
(begin -2), the mesh's z_nd, detJ, ax, ay, az with one (begin -1).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import oracle as orc

B1 = (-1, -1, -1)
B2 = (-2, -2, -2)

PERIODIC, INFLOW, OUTFLOW, SLIP_WALL, NO_SLIP_WALL, SURFACE_LAYER, GEOSTROPHIC = range(7)
EQ_U, EQ_V, EQ_W, EQ_THETA, EQ_TKE = range(5)
SS_USTAR, SS_SPEED, SS_D0, SS_TAU0, SS_TAU1, SS_TAU2, SS_DRAG, SS_KWALL, SS_AREA, SS_INVL, SS_N = range(11)
LOCS = (1, 2, 4, 0, 0, 0)


@dataclass
class ClosureOptions:
    """Source/Closure.H:78-105 (the defaults are ERF's)."""
    Cmu0: float = 0.5562
    Cb: float = 0.35
    Rt_crit: float = -1.0
    Rt_min: float = -3.0
    max_geom_lscale: float = 30.0
    tke_floor: float = 0.0
    sigma_k: float = 1.0
    nu_molecular: float = 0.0
    alpha_molecular: float = 0.0
    kappa: float = 0.41

    def c(self):
        return orc.OClosure(self.Cmu0, self.Cb, self.Rt_crit, self.Rt_min, self.max_geom_lscale, self.tke_floor,
                            self.sigma_k, self.nu_molecular, self.alpha_molecular, self.kappa)


@dataclass
class SurfaceOptions:
    """Source/SurfaceLayer.H:100-146: neutral, first-centre reference."""
    z0: float = 0.1
    Cmu0: float = 0.5562
    kappa: float = 0.41
    wall_k_dirichlet: bool = True
    floor_one: bool = False

    def c(self):
        return orc.OSurface(self.z0, self.Cmu0, self.kappa, int(self.wall_k_dirichlet), int(self.floor_one))


@dataclass
class Forcing:
    """Source/Forcing.H:40-96."""
    gravity: float = 9.81
    theta0: float = 300.0
    coriolis: bool = False
    three_d: bool = True
    geostrophic: bool = False
    latitude: float = 0.0
    rotation_period: float = 86400.0
    ug: float = 0.0
    vg: float = 0.0
    pressure_gradient: tuple = (0.0, 0.0, 0.0)

    def c(self):
        f = orc.OForcing(self.gravity, self.theta0, int(self.coriolis), int(self.three_d), int(self.geostrophic),
                         self.latitude, self.rotation_period, self.ug, self.vg)
        for d in range(3):
            f.pgrad[d] = self.pressure_gradient[d]
        return f

    def coriolis_parameter(self):
        return orc.lib().oracle_coriolis_parameter(int(self.coriolis), self.latitude, self.rotation_period)


@dataclass
class Boundary:
    """Source/BoundaryConditions.H:82-100: the face types (the ground and the
    top, and the sides where the geometry is not periodic), the top's theta
    gradient, the walls' velocities and the inflow profile an inflow face
    takes (oracle/initializer.py's InflowProfile)."""
    zlo: int = SURFACE_LAYER
    zhi: int = SLIP_WALL
    zhi_theta_grad: float = 0.0
    wall_velocity: tuple = ((0.0, 0.0, 0.0), (0.0, 0.0, 0.0))    # zlo, zhi
    xlo: int = PERIODIC
    xhi: int = PERIODIC
    ylo: int = PERIODIC
    yhi: int = PERIODIC
    side_wall_velocity: tuple = ((0.0,) * 3,) * 4                # xlo, ylo, xhi, yhi
    profile: object = None
    damping: object = None                                        # Damping (off when None)

    def faces(self):
        return (self.xlo, self.ylo, self.zlo, self.xhi, self.yhi, self.zhi)

    def c(self, forcing: Forcing):
        b = orc.OBc()
        for o, t in enumerate(self.faces()):
            b.face[o] = int(t)
        b.zhi_theta_grad = self.zhi_theta_grad
        for o, vel in zip((2, 5), self.wall_velocity):
            for d in range(3):
                b.wall_velocity[o][d] = vel[d]
        for o, vel in zip((0, 1, 3, 4), self.side_wall_velocity):
            for d in range(3):
                b.wall_velocity[o][d] = vel[d]
        b.ug, b.vg = forcing.ug, forcing.vg
        if self.damping is not None:
            d = self.damping
            b.damping.rayleigh_zdamp, b.damping.rayleigh_dampcoef = d.rayleigh_zdamp, d.rayleigh_dampcoef
            for n, name in enumerate(("u", "v", "w", "theta", "tke")):
                b.damping.rayleigh_fields[n] = 1 if name in d.fields else 0
            b.damping.rayleigh_above_ground = 1 if d.above_ground else 0
            b.damping.sponge_width, b.damping.sponge_strength = d.sponge_width, d.sponge_strength
            b.damping.ztop = d.ztop
            for k in range(2):
                b.damping.prob_lo[k], b.damping.prob_hi[k] = d.prob_lo[k], d.prob_hi[k]
        if self.profile is not None:
            from .initializer import c_profile
            pc, keep = c_profile(self.profile)
            b.profile = C.pointer(pc)
            b._keep = (pc, keep)
        return b


@dataclass
class Damping:
    """The Rayleigh and sponge layers (Source/BoundaryConditions.H:107-121)."""
    rayleigh_zdamp: float = 0.0
    rayleigh_dampcoef: float = 0.0
    fields: tuple = ("u", "v", "w", "theta", "tke")
    above_ground: bool = False
    sponge_width: float = 0.0
    sponge_strength: float = 0.0
    ztop: float = 0.0
    prob_lo: tuple = (0.0, 0.0)
    prob_hi: tuple = (0.0, 0.0)


@dataclass
class Physics:
    """Source/Residual.H:90-128 (ResidualInputs) restricted to this scope."""
    closure: ClosureOptions | None = field(default_factory=ClosureOptions)
    nu: np.ndarray | None = None            # given diffusivities (cells, one ghost) without a closure
    gamma_theta: np.ndarray | None = None
    gamma_tke: np.ndarray | None = None
    surface: SurfaceOptions | None = field(default_factory=SurfaceOptions)
    forcing: Forcing = field(default_factory=Forcing)
    bc: Boundary | None = field(default_factory=Boundary)   # None: no boundary rows (the ghosts are the caller's)
    theta_ref_z: tuple = (0.0,)     # the hydrostatic reference theta_ref(z) (K at heights in m)
    theta_ref: tuple = (300.0,)
    momentum_limiter: int = orc.VAN_LEER
    momentum_eps: float = 1e-5
    theta_limiter: int = orc.VAN_LEER
    theta_eps: float = 1e-5
    log_face_mean: bool = False


@dataclass
class Mesh:
    z: np.ndarray
    J: np.ndarray
    ax: np.ndarray
    ay: np.ndarray
    az: np.ndarray
    geom: object      # orc.OGeom


def _v(a, b):
    return orc._v(a, b)


def _shape(g, loc, grow):
    n = [g.hi[d] - g.lo[d] + 1 + ((loc >> d) & 1) + 2 * grow for d in range(3)]
    return (1, n[2], n[1], n[0])


def height_above_ground(mesh: Mesh):
    """Source/Closure.cpp:261-293: cells with one ghost (begin -1)."""
    d = np.zeros(_shape(mesh.geom, 0, 1))
    orc.lib().oracle_height_above_ground(_v(mesh.z, B1), mesh.geom, _v(d, B1))
    return d


def closure_evaluate(q, mesh: Mesh, opt: ClosureOptions, dist, z0, forcing: Forcing):
    """Source/Closure.cpp:366-466: dict of nu, gamma_theta, gamma_tke (one
    ghost, begin -1) and Ps, Pb, eps, dSdk (valid cells)."""
    g = mesh.geom
    f = {n: np.zeros(_shape(g, 0, 1)) for n in ("nu", "gamma_theta", "gamma_tke")}
    f.update({n: np.zeros(_shape(g, 0, 0)) for n in ("Ps", "Pb", "eps", "dSdk")})
    lo = tuple(g.lo)
    oc = opt.c()
    rc = orc.lib().oracle_closure_evaluate(
        _v(q["u"], B2), _v(q["v"], B2), _v(q["w"], B2), _v(q["theta"], B2), _v(q["tke"], B2), _v(mesh.z, B1),
        _v(dist, B1), float(z0), oc, forcing.gravity, forcing.theta0, g, _v(f["nu"], B1), _v(f["gamma_theta"], B1),
        _v(f["gamma_tke"], B1), _v(f["Ps"], lo), _v(f["Pb"], lo), _v(f["eps"], lo), _v(f["dSdk"], lo))
    if rc != 0:
        raise orc.OracleError("Closure::evaluate: the wall distance plus z0 must be > 0 at every cell")
    return f


def fill_ghosts(q, mesh: Mesh, phys: Physics):
    """Source/BoundaryConditions.cpp:597-666 in place; returns the surface
    state slab (SS_N, 1, ny+2, nx+2), begin (-1, -1, lo_z), or None."""
    g = mesh.geom
    ss = np.zeros((SS_N, 1, g.hi[1] - g.lo[1] + 3, g.hi[0] - g.lo[0] + 3))
    views = [orc.a4(q[n], B2) for n in ("u", "v", "w", "p", "theta", "tke")]
    arr = (C.POINTER(orc.OA4) * 6)(*[C.pointer(x) for x in views])
    so = (phys.surface or SurfaceOptions()).c()
    bcc = phys.bc.c(phys.forcing)
    rc = orc.lib().oracle_fill_ghosts(arr, _v(mesh.z, B1), bcc, so, g, _v(ss, (-1, -1, g.lo[2])))
    if rc == -3:
        raise orc.OracleError("fill_ghosts: an inflow face needs the inflow profile")
    if rc != 0:
        raise orc.OracleError("compute_surface_state: no positive, finite friction velocity")
    return ss if phys.bc.zlo == SURFACE_LAYER else None


def fill_ghosts_homogeneous(q, mesh: Mesh, phys: Physics):
    """Source/BoundaryConditions.cpp:418-540, 597-666 with homogeneous = true
    (the fill of a perturbation: every prescribed value, inflow profile and
    z gradient, the wall law's too, replaced by 0), in place."""
    g = mesh.geom
    views = [orc.a4(q[n], B2) for n in ("u", "v", "w", "p", "theta", "tke")]
    arr = (C.POINTER(orc.OA4) * 6)(*[C.pointer(x) for x in views])
    orc.lib().oracle_fill_ghosts_homogeneous(arr, _v(mesh.z, B1), phys.bc.c(phys.forcing), g)


def theta_ref_field(mesh: Mesh, phys: Physics):
    """Source/Forcing.cpp:88-125 on the cells and one ghost (begin -1)."""
    t = np.zeros(_shape(mesh.geom, 0, 1))
    n = len(phys.theta_ref_z)
    zs = (C.c_double * n)(*phys.theta_ref_z)
    ts = (C.c_double * n)(*phys.theta_ref)
    orc.lib().oracle_fill_theta_ref(n, zs, ts, _v(mesh.z, B1), _v(t, B1))
    return t


def residual(q, mesh: Mesh, phys: Physics):
    """Source/Residual.cpp:159-256 with the state's ghosts filled
    (fill_ghosts). Returns (R, closure fields or None, surface slab or None);
    R maps u, v, w, p, theta, tke to valid-point arrays (1, ...)."""
    g = mesh.geom
    lo = tuple(g.lo)
    logz = phys.log_face_mean
    ss = None
    wall = phys.surface is not None and phys.bc is not None and phys.bc.zlo == SURFACE_LAYER
    if wall:
        ss = np.zeros((SS_N, 1, g.hi[1] - g.lo[1] + 3, g.hi[0] - g.lo[0] + 3))
        so = phys.surface.c()
        if orc.lib().oracle_surface_state(_v(q["u"], B2), _v(q["v"], B2), _v(mesh.z, B1), so, g,
                                          _v(ss, (-1, -1, g.lo[2]))) != 0:
            raise orc.OracleError("compute_surface_state: no positive, finite friction velocity")
    cf = None
    if phys.closure is not None:
        dist = height_above_ground(mesh)
        z0 = phys.surface.z0 if wall else 0.0
        cf = closure_evaluate(q, mesh, phys.closure, dist, z0, phys.forcing)
        nu, gth, gk = cf["nu"], cf["gamma_theta"], cf["gamma_tke"]
    else:
        nu, gth, gk = phys.nu, phys.gamma_theta, phys.gamma_tke
    Fx, Fy, Fz, Fb = orc.transport_fluxes(q["u"], q["v"], q["w"], B2, mesh.z, B1, mesh.ax, mesh.ay, B1, g)
    F = [Fx[None], Fy[None], Fz[None]]
    vel = [q["u"], q["v"], q["w"]]
    R = {}
    # momentum: C - D + G p - S
    Gm = []
    for c in range(3):
        R["uvw"[c]] = orc.convection(1 << c, vel[c], B2, F, Fb, mesh.J, B1, g, phys.momentum_limiter,
                                     phys.momentum_eps)[None].copy()
        Gm.append(orc.momentum_diffusive_fluxes(c, vel, B2, nu, B1, mesh.z, B1, g, logz))
    Gt, bt = orc.scalar_diffusive_fluxes(q["theta"], B2, gth, B1, mesh.z, B1, g, logz)
    Gk, bk = orc.scalar_diffusive_fluxes(q["tke"], B2, gk, B1, mesh.z, B1, g, logz)
    if ss is not None:
        orc.lib().oracle_set_ground_fluxes(_v(ss, (-1, -1, g.lo[2])), g, _v(Gm[0][0][2], Gm[0][1][2]),
                                           _v(Gm[1][0][2], Gm[1][1][2]), _v(Gt[2], bt[2]), _v(Gk[2], bk[2]))
    for c in range(3):
        R["uvw"[c]] -= orc.flux_divergence(1 << c, Gm[c][0], Gm[c][1], mesh.J, B1, g)[None]
    Gp = orc.pressure_gradient(q["p"], B2, mesh.z, B1, mesh.J, B1, mesh.ax, mesh.ay, B1, g)
    for c in range(3):
        R["uvw"[c]] += Gp[c][None]
    tref = theta_ref_field(mesh, phys)
    fc = phys.forcing.c()
    orc.lib().oracle_forcing(_v(q["u"], B2), _v(q["v"], B2), _v(q["w"], B2), _v(q["theta"], B2), _v(tref, B1), fc, g,
                             _v(R["u"], lo), _v(R["v"], lo), _v(R["w"], lo))
    R["p"] = orc.velocity_divergence(q["u"], q["v"], q["w"], B2, mesh.z, B1, mesh.ax, mesh.ay, mesh.az, B1, mesh.J,
                                     B1, g)[None].copy()
    R["theta"] = orc.convection(0, q["theta"], B2, F, Fb, mesh.J, B1, g, phys.theta_limiter,
                                phys.theta_eps)[None].copy()
    R["theta"] -= orc.flux_divergence(0, Gt, bt, mesh.J, B1, g)[None]
    R["tke"] = orc.convection(0, q["tke"], B2, F, Fb, mesh.J, B1, g, orc.UPWIND, 1e-5, bounded=True)[None].copy()
    R["tke"] -= orc.flux_divergence(0, Gk, bk, mesh.J, B1, g)[None]
    if cf is not None:
        R["tke"] -= cf["Ps"]
        R["tke"] -= cf["Pb"]
        R["tke"] += cf["eps"]
    if ss is not None:
        orc.lib().oracle_apply_wall_k(_v(ss, (-1, -1, g.lo[2])), _v(q["tke"], B2),
                                      int(phys.surface.wall_k_dirichlet), g, _v(R["tke"], lo))
    if phys.bc is not None:
        names = ("u", "v", "w", "p", "theta", "tke")
        qv = (C.POINTER(orc.OA4) * 6)(*[C.pointer(orc.a4(q[n], B2)) for n in names])
        rviews = [orc.a4(R[n], lo) for n in names]
        rv = (C.POINTER(orc.OA4) * 6)(*[C.pointer(x) for x in rviews])
        orc.lib().oracle_add_damping(qv, _v(mesh.z, B1), phys.bc.c(phys.forcing), g, rv)
        orc.lib().oracle_apply_rows(_v(q["u"], B2), _v(q["v"], B2), _v(q["w"], B2), _v(mesh.z, B1),
                                    phys.bc.c(phys.forcing), g, _v(R["u"], lo), _v(R["v"], lo), _v(R["w"], lo))
    # OverrideSync: the hi faces of a periodic direction take the lo faces' values
    if g.periodic[0]:
        R["u"][0, :, :, -1] = R["u"][0, :, :, 0]
    if g.periodic[1]:
        R["v"][0, :, -1, :] = R["v"][0, :, 0, :]
    return R, cf, ss


def dirichlet_rows(eq, mesh: Mesh, phys: Physics):
    """Source/BoundaryConditions.cpp:849-876: 1 at the Dirichlet face points
    of eq at its location (valid points, begin lo), else 0."""
    g = mesh.geom
    loc = (1, 2, 4, 0, 0)[eq]
    m = np.zeros(_shape(g, loc, 0))
    orc.lib().oracle_dirichlet_rows(eq, phys.bc.c(phys.forcing), g, _v(m, tuple(g.lo)))
    return m


def implicit_rows(eq, F, Fb, mesh: Mesh, phys: Physics, cf=None, ss=None):
    """Source/Residual.cpp:272-304: the 7 implicit coefficients of eq at its
    location (valid points, begin lo), with dSdk, the wall model's rows and
    the z faces' folds."""
    g = mesh.geom
    loc = (1, 2, 4, 0, 0)[eq]
    if cf is not None:
        gdiff = cf["gamma_theta"] if eq == EQ_THETA else cf["gamma_tke"] if eq == EQ_TKE else cf["nu"]
    else:
        gdiff = phys.gamma_theta if eq == EQ_THETA else phys.gamma_tke if eq == EQ_TKE else phys.nu
    A = orc.implicit_coefficients(loc, F, Fb, gdiff, B1, mesh.z, B1, mesh.J, B1, g, logz=phys.log_face_mean)
    A = np.ascontiguousarray(A)
    so = (phys.surface or SurfaceOptions()).c()
    dS = cf["dSdk"] if (cf is not None and eq == EQ_TKE) else None
    sv = _v(ss, (-1, -1, g.lo[2])) if ss is not None else None
    bcc = (phys.bc or Boundary(zlo=PERIODIC, zhi=PERIODIC)).c(phys.forcing)
    if phys.bc is None or phys.surface is None:
        sv = None
        if phys.bc is not None and phys.bc.zlo == SURFACE_LAYER:
            raise orc.OracleError("implicit_rows: a surface_layer ground needs the surface options")
    orc.lib().oracle_implicit_folds(eq, sv, _v(dS, tuple(g.lo)), _v(mesh.J, B1), _v(mesh.z, B1), bcc, so, g,
                                    _v(A, tuple(g.lo)))
    return A
