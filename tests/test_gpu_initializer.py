"""The initializer on the GPU (gust_inflow_eval, gust_first_guess_state,
gust_obrien_omega, gust_initialize_state) against the oracle and the
reference's Askervein gold (Tests/Initialization/Mass_Consistent_Askervein:
divergence before to 1e-9, after at most 10 x reltol of before, max |w| and
the largest speed to 1e-6, theta 300 K and p' = 0, the state plotfile to
1e-8). The profile's log, pow and exp are the device's, so point values
agree with the oracle to a few units in the last place (1e-13 relative)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import initializer as OI
from oracle import oracle as orc
from oracle import physics as P
from tests import init_helpers as IH
from tests import physics_helpers as PHY

B1, B2 = (-1, -1, -1), (-2, -2, -2)
FIELDS = ("u", "v", "w", "p", "theta", "tke")


def _gp(p: OI.InflowProfile):
    from ablsteadyrans_amd import initializer as GI
    return GI.InflowProfile(kind=p.kind, ustar=p.ustar, z0=p.z0, speed_ref=p.speed_ref, z_ref=p.z_ref,
                            alpha=p.alpha, direction=p.direction, Cmu0=p.Cmu0, kappa=p.kappa, unbounded=p.unbounded,
                            l_max=p.l_max, table=p.table, theta_ref_z=tuple(p.theta_ref_z),
                            theta_ref=tuple(p.theta_ref))


def _go(o: OI.InitOptions):
    from ablsteadyrans_amd import initializer as GI
    return GI.InitOptions(first_guess=o.first_guess, projection=o.projection, z_agl_min=o.z_agl_min,
                          alpha_h=o.alpha_h, alpha_v=o.alpha_v, k_decay_height=o.k_decay_height, obrien=o.obrien,
                          column_theta=o.column_theta)


def _rel(a, b):
    s = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / s)


def _hill(periodic=(1, 1), n=(14, 12, 10)):
    from ablsteadyrans_amd import mesh as M
    lo, hi = (0.0, 0.0, 0.0), (700.0, 600.0, 500.0)
    geom = M.geometry(n, lo, hi, periodic)
    zl = M.make_zlevels(n[2], 500.0, dz0=10.0, stretch_ratio=1.15)
    return M.build_mesh(geom, lo, hi, zl, M.Terrain(kind="hills", hills=[(350.0, 300.0, 60.0, 120.0)]))


PROFILES = [
    OI.InflowProfile(kind=OI.LOG_LAW, ustar=0.45, z0=0.08, direction=25.0),
    OI.InflowProfile(kind=OI.LOG_LAW, ustar=0.45, z0=0.08, direction=-40.0, unbounded=True),
    OI.InflowProfile(kind=OI.POWER_LAW, speed_ref=7.0, z_ref=10.0, alpha=0.18, z0=0.05, direction=190.0,
                     theta_ref_z=(0.0, 200.0, 800.0), theta_ref=(295.0, 297.0, 305.0)),
    OI.InflowProfile(kind=OI.TABLE, z0=0.1, table=(np.array([0.0, 2.0, 10.0, 50.0, 200.0]),
                                                   np.array([0.0, 3.0, 5.0, 7.0, 9.0]),
                                                   np.array([0.0, 0.5, 1.0, 1.0, 0.5]),
                                                   np.array([300.0, 300.0, 301.0, 302.0, 305.0]), None)),
    OI.InflowProfile(kind=OI.TABLE, z0=0.1, table=(np.array([0.0, 2.0, 10.0, 50.0, 200.0]),
                                                   np.array([0.0, 3.0, 5.0, 7.0, 9.0]),
                                                   np.array([0.0, 0.5, 1.0, 1.0, 0.5]),
                                                   np.array([300.0, 300.0, 301.0, 302.0, 305.0]),
                                                   np.array([0.9, 0.8, 0.6, 0.3, 0.05]))),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(PROFILES)))
def test_inflow_eval_matches_the_oracle(cuda, case):
    import torch
    from ablsteadyrans_amd import initializer as GI
    p = PROFILES[case]
    z = np.concatenate([[-1.0, 0.0, 1e-3, 0.05], np.geomspace(0.1, 900.0, 60)])
    got = GI.inflow_eval(_gp(p), torch.from_numpy(z).cuda())
    ref = OI.inflow_eval(p, z)
    for n, a, b in zip(("u", "v", "theta", "k", "nut"), got, ref):
        b = np.broadcast_to(b, z.shape)
        assert _rel(a.cpu().numpy(), b) <= 1e-13, (case, n)


def _state(tile):
    from ablsteadyrans_amd import physics as GP
    return GP.DeviceState(tile.geom)


@pytest.mark.gpu
@pytest.mark.parametrize("case", [0, 2, 3, 4])
def test_first_guess_matches_the_oracle(cuda, case):
    from ablsteadyrans_amd import initializer as GI
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile = _hill()
    m = PHY.omesh(tile)
    p = PROFILES[case]
    opt = OI.InitOptions(first_guess=1, z_agl_min=1.5, k_decay_height=120.0 if case != 4 else 0.0,
                         obrien=case in (0, 3), column_theta=case == 3)
    q = PHY.zeros_state(tile.geom)
    ob = OI.first_guess_state(q, m, p, opt)
    dm = DeviceMesh(tile)
    ds = _state(tile)
    for n in FIELDS:
        ds[n].data.fill_(-7.0)
    gob = GI.first_guess_state(dm, _gp(p), _go(opt), ds)
    assert abs(gob - ob) <= 1e-12 * max(ob, 1e-30)
    for n in ("u", "v", "tke", "p"):   # every point, ghosts included
        assert _rel(ds[n].numpy()[0], q[n][0]) <= 1e-13, n
    # w: the valid faces (the ghosts are set to 0); theta: cells and one ghost
    w = ds["w"].numpy()[0]
    assert _rel(w[2:-2, 2:-2, 2:-2], q["w"][0, 2:-2, 2:-2, 2:-2]) <= 1e-12
    assert np.all(np.delete(w, np.s_[2:-2], axis=0) == 0.0)
    t = ds["theta"].numpy()[0]
    assert _rel(t[1:-1, 1:-1, 1:-1], q["theta"][0, 1:-1, 1:-1, 1:-1]) <= 1e-14


@pytest.mark.gpu
def test_obrien_omega_matches_the_oracle(cuda):
    import torch
    from ablsteadyrans_amd import initializer as GI
    from ablsteadyrans_amd.fab import Fab
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile = _hill(periodic=(0, 1))
    m = PHY.omesh(tile)
    q = PHY.zeros_state(tile.geom)
    OI.first_guess_state(q, m, PROFILES[2], OI.InitOptions(first_guess=1))
    om, worst = OI.obrien_omega(q["u"], q["v"], B2, m)
    dm = DeviceMesh(tile)
    u = Fab(torch.from_numpy(q["u"].copy()).cuda(), B2)
    v = Fab(torch.from_numpy(q["v"].copy()).cuda(), B2)
    nx, ny, nz = tile.geom.n
    omg = Fab(torch.full((1, nz + 1, ny, nx), 5.0, dtype=torch.float64).cuda(), (0, 0, 0))
    gw = GI.obrien_omega(dm, u, v, omg)
    assert worst > 1e-3 and abs(gw - worst) <= 1e-13 * worst
    assert _rel(omg.numpy()[0], om) <= 1e-12
    assert np.all(omg.numpy()[0, -1] == 0.0)


def _projection(tile, p, opt, faces, precision=0, reltol=1e-12):
    from ablsteadyrans_amd import initializer as GI
    from ablsteadyrans_amd import poisson as PS
    from ablsteadyrans_amd.fab import Fab
    dm = PS.DeviceMesh(tile)
    solver = PS.PoissonSolver(dm, PS.PoissonOptions(reltol=reltol, precision=precision))
    ds = _state(tile)
    nx, ny, nz = tile.geom.n
    lam = Fab.alloc((0, 0, 0), (nx - 1, ny - 1, nz - 1), ngrow=(1, 1, 1))
    res = GI.initialize_state(dm, _gp(p), _go(opt), faces, solver, ds, lam)
    return ds, lam, res


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_askervein_matches_the_reference_gold(cuda, precision):
    tile = IH.askervein_tile()
    ds, lam, res = _projection(tile, IH.askervein_profile(), IH.askervein_options(), IH.ASKERVEIN_FACES,
                               precision=precision, reltol=1e-10)
    gold = np.load(f"{IH.GOLD}/init_askervein.npz")
    print(f"askervein initializer (precision {precision}): {res.poisson.iterations} GMRES iterations, "
          f"divergence {res.divergence_before:.10g} -> {res.divergence_after:.3e}", flush=True)
    assert abs(res.divergence_before - float(gold["div_before"])) <= 1e-9 * float(gold["div_before"])
    assert res.divergence_after <= 10 * 1e-10 * res.divergence_before
    q = {n: ds[n].numpy() for n in FIELDS}
    cf = IH.cell_fields(q)
    for n in ("u", "v", "w", "tke"):
        assert IH.fcompare(cf[n], gold[n]) <= 1e-8, (n, IH.fcompare(cf[n], gold[n]))
    assert np.all(cf["theta"] == 300.0) and np.all(cf["p"] == 0.0) and cf["tke"].min() > 0
    speed = float(np.hypot(cf["u"], cf["v"]).max())
    wmax = float(np.abs(cf["w"]).max())
    assert abs(speed - float(gold["speed_max"])) <= 1e-6 * float(gold["speed_max"])
    assert abs(wmax - float(gold["wmax"])) <= 1e-6 * float(gold["wmax"])
    # and the oracle composite (exact solve) to the solve's tolerance
    qo = PHY.zeros_state(tile.geom)
    IH.initialize_state(tile, qo, IH.askervein_profile(), IH.askervein_options(), IH.ASKERVEIN_FACES)
    for n in ("u", "v", "w"):
        assert _rel(q[n][0, 2:-2, 2:-2, 2:-2], qo[n][0, 2:-2, 2:-2, 2:-2]) <= 1e-8, n


CASES = [
    # periodic sides, unit alphas (the plain projection), log law
    ((1, 1), (P.PERIODIC, P.PERIODIC, P.SURFACE_LAYER, P.PERIODIC, P.PERIODIC, P.SLIP_WALL), 0,
     dict(first_guess=1)),
    # periodic y, inflow/outflow x, anisotropic with O'Brien, power law
    ((0, 1), (P.INFLOW, P.PERIODIC, P.SURFACE_LAYER, P.OUTFLOW, P.PERIODIC, P.SLIP_WALL), 2,
     dict(first_guess=1, alpha_h=1.0, alpha_v=0.4, obrien=True)),
    # walls in y, alpha_v = alpha_h != 1 (no split), the table
    ((0, 0), (P.INFLOW, P.SLIP_WALL, P.SURFACE_LAYER, P.OUTFLOW, P.SLIP_WALL, P.SLIP_WALL), 4,
     dict(first_guess=1, alpha_h=0.7, alpha_v=0.7, k_decay_height=80.0)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES)))
def test_projection_matches_the_oracle(cuda, case):
    per, faces, pi, kw = CASES[case]
    tile = _hill(periodic=per)
    p = PROFILES[pi]
    opt = OI.InitOptions(**kw)
    qo = PHY.zeros_state(tile.geom)
    before, after, lam_o, ob = IH.initialize_state(tile, qo, p, opt, faces)
    ds, lam, res = _projection(tile, p, opt, faces)
    assert abs(res.divergence_before - before) <= 1e-12 * before
    assert res.divergence_after <= 1e-10 * before and after <= 1e-10 * before
    assert abs(res.obrien_top_residual - ob) <= 1e-12 * max(ob, 1e-30)
    q = {n: ds[n].numpy() for n in FIELDS}
    for n in ("u", "v", "w"):
        assert _rel(q[n][0, 2:-2, 2:-2, 2:-2], qo[n][0, 2:-2, 2:-2, 2:-2]) <= 1e-9, n
    # the one-ghost ring of u and v (periodic images and the first guess)
    for n in ("u", "v"):
        assert _rel(q[n][0, 1:-1, 1:-1, 1:-1], qo[n][0, 1:-1, 1:-1, 1:-1]) <= 1e-9, n
    if lam_o is not None:
        lg = lam.numpy()[0, 1:-1, 1:-1, 1:-1]
        off = 0.0
        if all(f not in (P.INFLOW, P.OUTFLOW) for f in faces):   # singular: lambda up to a constant
            J = tile.detJ[1:-1, 1:-1, 1:-1]
            off = np.sum(J * (lg - lam_o)) / np.sum(J)
        assert _rel(lg - off, lam_o) <= 1e-8


@pytest.mark.gpu
def test_no_projection_and_no_first_guess(cuda):
    tile = _hill()
    faces = CASES[0][1]
    ds, _, res = _projection(tile, PROFILES[0], OI.InitOptions(first_guess=1, projection=False), faces)
    qo = PHY.zeros_state(tile.geom)
    before, after, _, _ = IH.initialize_state(tile, qo, PROFILES[0], OI.InitOptions(first_guess=1, projection=False),
                                              faces)
    assert res.divergence_before == res.divergence_after
    assert abs(res.divergence_before - before) <= 1e-12 * before
    assert res.poisson.iterations == 0
    # first_guess = none leaves the state as it is
    from ablsteadyrans_amd import initializer as GI
    from ablsteadyrans_amd.poisson import DeviceMesh
    ds2 = _state(tile)
    ds2["u"].data.fill_(3.0)
    r = GI.initialize_state(DeviceMesh(tile), _gp(PROFILES[0]), GI.InitOptions(first_guess=0), faces, None, ds2)
    assert r.divergence_before == 0.0 and np.all(ds2["u"].numpy() == 3.0)
