"""The residual's physics on the GPU against the oracle: heights bit for
bit; the closure, the wall model, the ghost fills, the residual of all six
equations and the implicit rows of all five transported fields on a hill
tile with a random state, within 1e-12 of each field's scale (the device's
log and pow differ from the host's in the last bit; everything else is the
same operations in the same order)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import physics as P
from tests import physics_helpers as PHY

B1, B2 = (-1, -1, -1), (-2, -2, -2)


def _fab(a, begin):
    import torch
    from ablsteadyrans_amd.fab import Fab
    return Fab(torch.from_numpy(np.ascontiguousarray(a)).cuda(), begin)


def _inputs(phys_o: P.Physics, with_closure=True):
    from ablsteadyrans_amd import physics as GP
    f = phys_o.forcing
    gi = GP.ResidualInputs(
        closure=GP.ClosureOptions() if with_closure else None,
        surface=GP.SurfaceLayerOptions(z0=phys_o.surface.z0, wall_k_dirichlet=phys_o.surface.wall_k_dirichlet),
        forcing=GP.ForcingOptions(gravity=f.gravity, theta0=f.theta0, coriolis=f.coriolis, coriolis_3d=f.three_d,
                                  geostrophic=f.geostrophic, latitude=f.latitude, rotation_period=f.rotation_period,
                                  ug=f.ug, vg=f.vg, pressure_gradient=f.pressure_gradient),
        faces=(0, 0, phys_o.bc.zlo, 0, 0, phys_o.bc.zhi) if phys_o.bc is not None else None,
        zhi_theta_grad=phys_o.bc.zhi_theta_grad if phys_o.bc is not None else 0.0,
        theta_ref_z=phys_o.theta_ref_z, theta_ref=phys_o.theta_ref)
    return gi


def _setup(seed=3, wall=True, zhi=P.SLIP_WALL):
    tile = PHY.hill_mesh(n=(16, 12, 10), height=90.0, half=200.0)
    m = PHY.omesh(tile)
    q = PHY.random_state(tile, seed, U=6.0, amp=0.3)
    phys = P.Physics(bc=P.Boundary(zlo=P.SURFACE_LAYER if wall else P.SLIP_WALL, zhi=zhi, zhi_theta_grad=0.003),
                     surface=P.SurfaceOptions(z0=0.1),
                     forcing=P.Forcing(coriolis=True, latitude=45.0, geostrophic=True, ug=6.0, vg=1.0,
                                       pressure_gradient=(1e-4, 0.0, 0.0)),
                     theta_ref_z=(0.0, 300.0, 1000.0), theta_ref=(300.0, 300.0, 302.0))
    return tile, m, q, phys


def _dstate(tile, q):
    from ablsteadyrans_amd import physics as GP
    s = GP.DeviceState(tile.geom)
    for n in GP.FIELDS:
        s[n].data.copy_(__import__("torch").from_numpy(q[n]))
    return s


def _close(a, b, scale_rel=1e-12, what=""):
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max()
    assert err <= scale_rel * scale, f"{what}: max err {err:.3e} vs scale {scale:.3e}"


@pytest.mark.gpu
def test_heights_are_bitwise(cuda):
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd.fab import Fab
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile, m, _, _ = _setup()
    ref = P.height_above_ground(m)
    d = Fab.at_loc(tile.geom, 0, 1, 1)
    GP.height_above_ground(DeviceMesh(tile), d)
    assert np.array_equal(d.numpy(), ref)


@pytest.mark.gpu
def test_closure_and_surface_state_match(cuda):
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd.fab import Fab
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile, m, q, phys = _setup()
    dm = DeviceMesh(tile)
    dist = P.height_above_ground(m)
    ref = P.closure_evaluate(q, m, P.ClosureOptions(), dist, 0.1, phys.forcing)
    out = GP.ClosureFields(tile.geom)
    GP.closure_evaluate(dm, _dstate(tile, q), _fab(dist, B1), 0.1, GP.ClosureOptions(), out)
    for n in ("nu", "gamma_theta", "gamma_tke", "Ps", "Pb", "eps", "dSdk"):
        _close(out[n].numpy(), ref[n], 1e-13, n)
    ss_ref = np.zeros((P.SS_N, 1, 14, 18))
    orc.lib().oracle_surface_state(orc._v(q["u"], B2), orc._v(q["v"], B2), orc._v(m.z, B1), phys.surface.c(), m.geom,
                                   orc._v(ss_ref, (-1, -1, 0)))
    ss = GP.surface_slab(tile.geom)
    GP.surface_state(dm, _dstate(tile, q), GP.SurfaceLayerOptions(z0=0.1), ss)
    for c in range(P.SS_N):
        _close(ss.numpy()[c], ss_ref[c], 1e-14, f"ss[{c}]")


@pytest.mark.gpu
@pytest.mark.parametrize("zhi", [P.SLIP_WALL, P.GEOSTROPHIC, P.NO_SLIP_WALL])
def test_fill_ghosts_match(cuda, zhi):
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile, m, q, phys = _setup(zhi=zhi)
    ds = _dstate(tile, q)
    P.fill_ghosts(q, m, phys)
    GP.fill_ghosts(DeviceMesh(tile), ds, _inputs(phys))
    for n in ("u", "v", "w", "p", "theta", "tke"):
        _close(ds[n].numpy(), q[n], 1e-13, n)
    for n in ("w", "p", "theta", "tke"):
        assert np.array_equal(ds[n].numpy(), q[n]), n


@pytest.mark.gpu
@pytest.mark.parametrize("with_closure,wall", [(True, True), (True, False), (False, True)])
def test_residual_and_implicit_rows_match(cuda, with_closure, wall):
    import torch
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd.fab import Fab
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile, m, q, phys = _setup(seed=7, wall=wall)
    if not with_closure:
        rng = np.random.default_rng(1)
        phys.closure = None
        for name in ("nu", "gamma_theta", "gamma_tke"):
            setattr(phys, name, rng.uniform(0.5, 5.0, size=(1, 12, 14, 18)))
    P.fill_ghosts(q, m, phys)
    R_ref, cf_ref, ss_ref = P.residual(q, m, phys)
    gi = _inputs(phys, with_closure)
    if not with_closure:
        gi.nu, gi.gamma_theta, gi.gamma_tke = (_fab(getattr(phys, n), B1) for n in ("nu", "gamma_theta", "gamma_tke"))
    dm = DeviceMesh(tile)
    ds = _dstate(tile, q)
    R = {n: Fab.at_loc(tile.geom, loc, 1, 0) for n, loc in zip(GP.FIELDS, GP.LOCS)}
    cf = GP.ClosureFields(tile.geom) if with_closure else None
    ss = GP.surface_slab(tile.geom)
    GP.residual(dm, ds, gi, R, cf, ss)
    torch.cuda.synchronize()
    for n in GP.FIELDS:
        _close(R[n].numpy(), R_ref[n], 1e-11, n)
    # the implicit rows of the five transported fields
    Fx, Fy, Fz, Fb = orc.transport_fluxes(q["u"], q["v"], q["w"], B2, m.z, B1, m.ax, m.ay, B1, m.geom)
    F = [Fx[None], Fy[None], Fz[None]]
    Fd = [_fab(f, Fb) for f in F]
    for eq_o, eq_g, loc in ((P.EQ_U, GP.EQ_U, 1), (P.EQ_V, GP.EQ_V, 2), (P.EQ_W, GP.EQ_W, 4),
                            (P.EQ_THETA, GP.EQ_THETA, 0), (P.EQ_TKE, GP.EQ_TKE, 0)):
        A_ref = P.implicit_rows(eq_o, F, Fb, m, phys, cf_ref, ss_ref)
        A = Fab.at_loc(tile.geom, loc, 7, 0)
        GP.implicit_rows(eq_g, dm, Fd, gi, A, cf, ss if wall else None)
        _close(A.numpy(), A_ref, 1e-12, f"A[{eq_o}]")


# ---- non-periodic sides: inflow, outflow, slip and no-slip walls ----------------

def _gprofile(p):
    from ablsteadyrans_amd import initializer as GI
    return GI.InflowProfile(kind=p.kind, ustar=p.ustar, z0=p.z0, speed_ref=p.speed_ref, z_ref=p.z_ref,
                            alpha=p.alpha, direction=p.direction, Cmu0=p.Cmu0, kappa=p.kappa, unbounded=p.unbounded,
                            l_max=p.l_max, table=p.table, theta_ref_z=tuple(p.theta_ref_z),
                            theta_ref=tuple(p.theta_ref))


def _open_inputs(phys_o: P.Physics):
    gi = _inputs(phys_o)
    b = phys_o.bc
    gi.faces = b.faces()
    wv = [(0.0, 0.0, 0.0)] * 6
    for o, v in zip((0, 1, 3, 4), b.side_wall_velocity):
        wv[o] = tuple(v)
    for o, v in zip((2, 5), b.wall_velocity):
        wv[o] = tuple(v)
    gi.wall_velocity = tuple(wv)
    gi.inflow = _gprofile(b.profile) if b.profile is not None else None
    if b.damping is not None:
        from ablsteadyrans_amd import physics as GP
        d = b.damping
        gi.damping = GP.Damping(rayleigh_zdamp=d.rayleigh_zdamp, rayleigh_dampcoef=d.rayleigh_dampcoef,
                                fields=d.fields, above_ground=d.above_ground, sponge_width=d.sponge_width,
                                sponge_strength=d.sponge_strength, ztop=d.ztop, prob_lo=d.prob_lo, prob_hi=d.prob_hi)
    return gi


OPEN = [((0, 0), (P.INFLOW, P.SLIP_WALL, P.OUTFLOW, P.NO_SLIP_WALL)),
        ((0, 1), (P.INFLOW, P.PERIODIC, P.OUTFLOW, P.PERIODIC)),
        ((1, 0), (P.PERIODIC, P.NO_SLIP_WALL, P.PERIODIC, P.SLIP_WALL)),
        ((0, 0), (P.OUTFLOW, P.INFLOW, P.INFLOW, P.OUTFLOW))]


def _open_setup(case, seed=4):
    from ablsteadyrans_amd import mesh as M
    from oracle import initializer as OI
    per, (xlo, ylo, xhi, yhi) = OPEN[case]
    n = (14, 12, 10)
    lo, hi = (0.0, 0.0, 0.0), (700.0, 600.0, 500.0)
    geom = M.geometry(n, lo, hi, per)
    zl = M.make_zlevels(n[2], 500.0, dz0=10.0, stretch_ratio=1.15)
    tile = M.build_mesh(geom, lo, hi, zl, M.Terrain(kind="hills", hills=[(350.0, 300.0, 70.0, 130.0)]))
    m = PHY.omesh(tile)
    q = PHY.random_state(tile, seed, U=6.0, amp=0.3)
    prof = OI.InflowProfile(kind=OI.LOG_LAW, ustar=0.45, z0=0.1, direction=15.0, theta_ref_z=(0.0, 300.0, 1000.0),
                            theta_ref=(300.0, 300.0, 302.0))
    bc = P.Boundary(zlo=P.SURFACE_LAYER, zhi=P.SLIP_WALL, zhi_theta_grad=0.003, xlo=xlo, ylo=ylo, xhi=xhi, yhi=yhi,
                    side_wall_velocity=((0.0,) * 3, (1.2, 0.0, 0.0), (0.0,) * 3, (0.0, 0.0, 0.7)), profile=prof)
    phys = P.Physics(bc=bc, surface=P.SurfaceOptions(z0=0.1),
                     forcing=P.Forcing(coriolis=True, latitude=45.0, geostrophic=True, ug=6.0, vg=1.0),
                     theta_ref_z=(0.0, 300.0, 1000.0), theta_ref=(300.0, 300.0, 302.0))
    return tile, m, q, phys


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(OPEN)))
def test_open_side_fill_ghosts_match(cuda, case):
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile, m, q, phys = _open_setup(case)
    ds = _dstate(tile, q)
    ss_ref = P.fill_ghosts(q, m, phys)
    ss = GP.surface_slab(tile.geom)
    GP.fill_ghosts(DeviceMesh(tile), ds, _open_inputs(phys), ss)
    for n in ("u", "v", "w", "p", "theta", "tke"):
        _close(ds[n].numpy(), q[n], 1e-13, n)
    for c in range(P.SS_N):
        _close(ss.numpy()[c], ss_ref[c], 1e-13, f"ss[{c}]")


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(OPEN)))
def test_open_side_residual_and_implicit_rows_match(cuda, case):
    import torch
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd.fab import Fab
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile, m, q, phys = _open_setup(case, seed=9)
    P.fill_ghosts(q, m, phys)
    R_ref, cf_ref, ss_ref = P.residual(q, m, phys)
    gi = _open_inputs(phys)
    dm = DeviceMesh(tile)
    ds = _dstate(tile, q)
    R = {n: Fab.at_loc(tile.geom, loc, 1, 0) for n, loc in zip(GP.FIELDS, GP.LOCS)}
    cf = GP.ClosureFields(tile.geom)
    ss = GP.surface_slab(tile.geom)
    GP.residual(dm, ds, gi, R, cf, ss)
    torch.cuda.synchronize()
    for n in GP.FIELDS:
        _close(R[n].numpy(), R_ref[n], 1e-11, n)
    Fx, Fy, Fz, Fb = orc.transport_fluxes(q["u"], q["v"], q["w"], B2, m.z, B1, m.ax, m.ay, B1, m.geom)
    F = [Fx[None], Fy[None], Fz[None]]
    Fd = [_fab(f, Fb) for f in F]
    for eq_o, eq_g, loc in ((P.EQ_U, GP.EQ_U, 1), (P.EQ_V, GP.EQ_V, 2), (P.EQ_W, GP.EQ_W, 4),
                            (P.EQ_THETA, GP.EQ_THETA, 0), (P.EQ_TKE, GP.EQ_TKE, 0)):
        A_ref = P.implicit_rows(eq_o, F, Fb, m, phys, cf_ref, ss_ref)
        A = Fab.at_loc(tile.geom, loc, 7, 0)
        GP.implicit_rows(eq_g, dm, Fd, gi, A, cf, ss)
        _close(A.numpy(), A_ref, 1e-12, f"A[{eq_o}]")


@pytest.mark.gpu
@pytest.mark.parametrize("above_ground", [False, True])
def test_damping_layers_match(cuda, above_ground):
    """The Rayleigh layer (w and theta, absolute or above-ground height) and
    the sponge of the inflow/outflow faces in the residual and the rows."""
    import torch
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd.fab import Fab
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile, m, q, phys = _open_setup(0, seed=13)
    ztop = float(tile.zlevels[-1])
    phys.bc.damping = P.Damping(rayleigh_zdamp=0.45 * ztop, rayleigh_dampcoef=0.2, fields=("w", "theta", "u"),
                                above_ground=above_ground, ztop=ztop, sponge_width=140.0,
                                sponge_strength=0.05, prob_lo=(0.0, 0.0), prob_hi=(700.0, 600.0))
    P.fill_ghosts(q, m, phys)
    R_ref, cf_ref, ss_ref = P.residual(q, m, phys)
    gi = _open_inputs(phys)
    dm = DeviceMesh(tile)
    ds = _dstate(tile, q)
    R = {n: Fab.at_loc(tile.geom, loc, 1, 0) for n, loc in zip(GP.FIELDS, GP.LOCS)}
    cf = GP.ClosureFields(tile.geom)
    ss = GP.surface_slab(tile.geom)
    GP.residual(dm, ds, gi, R, cf, ss)
    torch.cuda.synchronize()
    for n in GP.FIELDS:
        _close(R[n].numpy(), R_ref[n], 1e-11, n)
    Fx, Fy, Fz, Fb = orc.transport_fluxes(q["u"], q["v"], q["w"], B2, m.z, B1, m.ax, m.ay, B1, m.geom)
    F = [Fx[None], Fy[None], Fz[None]]
    Fd = [_fab(f, Fb) for f in F]
    for eq_o, eq_g, loc in ((P.EQ_U, GP.EQ_U, 1), (P.EQ_W, GP.EQ_W, 4), (P.EQ_THETA, GP.EQ_THETA, 0)):
        A_ref = P.implicit_rows(eq_o, F, Fb, m, phys, cf_ref, ss_ref)
        A = Fab.at_loc(tile.geom, loc, 7, 0)
        GP.implicit_rows(eq_g, dm, Fd, gi, A, cf, ss)
        _close(A.numpy(), A_ref, 1e-12, f"A[{eq_o}]")
