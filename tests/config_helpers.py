"""Helpers of the configuration-scale GPU tests (TEST INFRASTRUCTURE).

* :func:`open_abl_physics` -- the oracle's Physics of decks.open_abl (the
  open-channel neutral ABL of the flat 64x64x48 and Gaussian-hill 200x200x64
  configurations): log-law inflow at xlo, outflow at xhi, y periodic, the
  surface layer, a slip top, Coriolis without geostrophic forcing.
* :func:`host_state` -- a device state as the oracle's dict of two-ghost arrays.
* :class:`RefinedSolve` -- the composite's pressure solve on tiles too large
  to factor: iterative refinement on the ORACLE operator (each residual
  rhs - A phi is the oracle's C apply), with a linear solver as the inner
  solve, until the oracle residual is below rtol times the right-hand side.
  The inner solver may be any approximate solver; the device GMRES is used
  here only as that black box, and the returned phi is accepted only by its
  oracle residual, so the composite's correction is the oracle operator's.
"""
from __future__ import annotations

import numpy as np

from oracle import initializer as OI
from oracle import oracle as orc
from oracle import physics as P
from tests import physics_helpers as PHY
from tests import poisson_helpers as PH

FIELDS = ("u", "v", "w", "p", "theta", "tke")


def open_abl_physics(deck) -> P.Physics:
    """The oracle Physics of a decks.open_abl deck (its inputs are read back
    so the two sides cannot drift apart)."""
    gi = deck.inputs
    prof = OI.InflowProfile(kind=OI.LOG_LAW, ustar=gi.inflow.ustar, z0=gi.inflow.z0, direction=gi.inflow.direction)
    damping = None
    if gi.damping is not None:
        d = gi.damping
        damping = P.Damping(rayleigh_zdamp=d.rayleigh_zdamp, rayleigh_dampcoef=d.rayleigh_dampcoef, fields=d.fields,
                            above_ground=d.above_ground, sponge_width=d.sponge_width,
                            sponge_strength=d.sponge_strength, ztop=d.ztop, prob_lo=d.prob_lo, prob_hi=d.prob_hi)
    f = gi.forcing
    faces = gi.faces
    bc = P.Boundary(zlo=faces[2], zhi=faces[5], zhi_theta_grad=gi.zhi_theta_grad, xlo=faces[0], ylo=faces[1],
                    xhi=faces[3], yhi=faces[4], profile=prof, damping=damping)
    c = gi.closure
    return P.Physics(closure=P.ClosureOptions(Cmu0=c.Cmu0, Cb=c.Cb, Rt_crit=c.Rt_crit, Rt_min=c.Rt_min,
                                              max_geom_lscale=c.max_geom_lscale, tke_floor=c.tke_floor,
                                              sigma_k=c.sigma_k, nu_molecular=c.nu_molecular,
                                              alpha_molecular=c.alpha_molecular, kappa=c.kappa),
                     bc=bc, surface=P.SurfaceOptions(z0=gi.surface.z0),
                     forcing=P.Forcing(gravity=f.gravity, theta0=f.theta0, coriolis=f.coriolis,
                                       three_d=f.coriolis_3d, geostrophic=f.geostrophic, latitude=f.latitude,
                                       rotation_period=f.rotation_period, ug=f.ug, vg=f.vg),
                     theta_ref_z=gi.theta_ref_z, theta_ref=gi.theta_ref)


def stage_norms_of(R, q, m, stage_opt) -> list:
    """The oracle's stage norms (Source/Simplec.cpp:385-409) of a residual
    dict from oracle.physics.residual, with the stage's default residual
    scales (velocity 10, theta 10, k 0.1) and its scale height."""
    res, _ = orc.stage_norms([R[n] for n in FIELDS], (0, 0, 0), q["u"], q["v"], q["w"], (-2, -2, -2), q["tke"],
                             (-2, -2, -2), 10.0, 10.0, 0.1, float(stage_opt.height), m.geom)
    return res


def host_state(ds) -> dict:
    return {n: ds[n].numpy().astype(np.float64).copy() for n in FIELDS}


def to_device(ds, q) -> None:
    import torch
    for n in FIELDS:
        ds[n].data.copy_(torch.from_numpy(q[n]))


def omesh(tile):
    return PHY.omesh(tile)


class RefinedSolve:
    """phi with || rhs - A_oracle phi || <= rtol || rhs || (2-norm; the
    J-weighted means of rhs and phi removed in the singular case, as
    Source/Poisson.cpp:559-567 and 603-611), by iterative refinement on the
    oracle operator with the device Poisson solve (reltol inner_rtol) as the
    inner solver. Records the oracle residual of every solve."""

    def __init__(self, tile, dmesh, rtol=1e-11, inner_rtol=1e-10, max_refinements=6):
        from ablsteadyrans_amd import poisson as PS
        self.tile, self.dmesh, self.rtol = tile, dmesh, rtol
        self.solver = PS.PoissonSolver(dmesh, PS.PoissonOptions(reltol=inner_rtol, max_iter=400))
        self.max_refinements = max_refinements
        self.history = []

    def _device_solve(self, r, coeffs, bc):
        import torch
        from ablsteadyrans_amd import poisson as PS
        from ablsteadyrans_amd.fab import Fab
        nx, ny, nz = self.tile.geom.n
        rhs = Fab(torch.from_numpy(np.ascontiguousarray(r)[None]).cuda(), (0, 0, 0))
        phi = Fab(torch.zeros((1, nz + 2, ny + 2, nx + 2), dtype=torch.float64, device="cuda"), (-1, -1, -1))
        fc = PS.FaceCoeffs(*[Fab(torch.from_numpy(np.ascontiguousarray(c)[None]).cuda(), (-1, -1, -1))
                             for c in coeffs])
        self.solver.solve(rhs, phi, fc, bc)
        return phi.numpy()[0, 1:-1, 1:-1, 1:-1].copy()

    def __call__(self, rhs, coeffs, bc):
        m = self.tile
        J = m.detJ[1:-1, 1:-1, 1:-1]
        singular = all(b != orc.DIRICHLET for b in bc)
        b = rhs.copy()
        if singular:
            b -= np.sum(J * b) / np.sum(J)
        nb = float(np.linalg.norm(b))
        phi = np.zeros_like(b)
        res = nb
        for _ in range(self.max_refinements):
            r = b - PH.apply(m, phi, bc, coeffs)
            res = float(np.linalg.norm(r))
            if res <= self.rtol * nb:
                break
            phi = phi + self._device_solve(r, coeffs, bc)
            if singular:
                phi -= np.sum(J * phi) / np.sum(J)
        else:
            r = b - PH.apply(m, phi, bc, coeffs)
            res = float(np.linalg.norm(r))
        self.history.append(res / nb if nb > 0 else 0.0)
        assert res <= self.rtol * nb, f"refinement on the oracle operator stalled at {res / nb:.3e}"
        return phi
