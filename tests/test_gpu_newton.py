"""The Newton-Krylov finish on the GPU (Source/Newton.*, Continuation.*,
NewtonPreconditioner.*, NewtonStage.*): the homogeneous ghost fill bit for
bit against the oracle's; J v against the oracle's forward difference of
the same residual (the same eps, bit for bit) and exactly linear in powers of
two; continuation steps on the reference's Neutral_Column deck from a
SIMPLEC state converging ten orders at the Krylov tolerance's rate; the
steady.stage = simplec_newton run of that deck reaching the deck's criterion
in fewer iterations than SIMPLEC alone, at the state the oracle's composite
converges to (tests/golden/column_steady.npz) within 1e-6 relative L2; the
preconditioner with the stage's denominators equal to one SIMPLEC iteration
(NewtonPreconditioner.ApplyIsOneSimplecIteration); its GMRES iterations
bounded over CFL with the diagonal denominators
(NewtonPreconditioner.IterationsBoundedOverCfl); and the continuation with
either preconditioner reaching the same state
(NewtonPreconditioner.ContinuationIdentityVersusSimplec)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import physics as P
from tests import column_helpers as CO
from tests import newton_helpers as NH
from tests.test_gpu_physics import OPEN, _dstate, _open_inputs, _open_setup

pytestmark = pytest.mark.gpu

FIELDS = NH.FIELDS
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "column_steady.npz")


@pytest.mark.parametrize("case", range(len(OPEN)))
def test_homogeneous_fill_matches_the_oracle(cuda, case):
    from ablsteadyrans_amd import newton as NW
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile, m, q, phys = _open_setup(case, seed=12)
    ds = _dstate(tile, q)
    P.fill_ghosts_homogeneous(q, m, phys)
    NW.fill_ghosts_homogeneous(DeviceMesh(tile), ds, _open_inputs(phys))
    for n in FIELDS:
        assert np.array_equal(ds[n].numpy(), q[n]), n


def _jvp_setup(seed=5):
    from tests.test_oracle_newton import _jvp_case
    return _jvp_case(seed)


def _gpu_stage(tile, phys):
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd.poisson import DeviceMesh
    from tests.test_gpu_stage import _inputs
    dm = DeviceMesh(tile)
    stage = GP.SimplecStage(dm, _inputs(phys), GP.SimplecOptions(sweeps=2, tke_floor=1e-6, height=600.0))
    return dm, stage


def test_jvp_matches_the_oracle_difference(cuda):
    import torch
    from ablsteadyrans_amd import newton as NW
    from ablsteadyrans_amd import physics as GP
    tile, m, q, v, phys = _jvp_setup()
    dm, stage = _gpu_stage(tile, phys)
    gi = stage.inputs
    dq, dv = _dstate(tile, q), _dstate(tile, v)
    GP.fill_ghosts(dm, dq, gi)
    R0 = NW.residual_fields(tile.geom)
    GP.residual(dm, dq, gi, R0)
    Jv = NW.residual_fields(tile.geom)
    eps = NW.jvp(stage, dq, dv, R0, Jv)
    torch.cuda.synchronize()
    eps_ref, J_ref = NH.jvp(q, v, m, phys)
    print(f"eps device {eps!r} oracle {eps_ref!r}", flush=True)
    assert eps == eps_ref
    for n in FIELDS:
        a, b = Jv[n].numpy(), J_ref[n]
        scale = max(float(np.abs(b).max()), 1e-30)
        err = float(np.abs(a - b).max()) / scale
        print(f"  J v {n}: max |device - oracle| / max |oracle| = {err:.2e}", flush=True)
        assert err <= 1e-5, (n, err)
    # J(4 v) = 4 J v bit for bit (the same perturbed point)
    dv4 = _dstate(tile, {n: 4.0 * v[n] for n in FIELDS})
    J4 = NW.residual_fields(tile.geom)
    eps4 = NW.jvp(stage, dq, dv4, R0, J4)
    torch.cuda.synchronize()
    assert eps4 == eps / 4
    for n in FIELDS:
        assert torch.equal(J4[n].data, 4.0 * Jv[n].data), n


def _column_stage(pressure_reltol=1e-3):
    import torch
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile = CO.column_tile()
    dm = DeviceMesh(tile)
    inputs = GP.ResidualInputs(closure=GP.ClosureOptions(max_geom_lscale=1.0e12),
                               surface=GP.SurfaceLayerOptions(z0=0.1),
                               forcing=GP.ForcingOptions(pressure_gradient=(1.225e-4, 0.0, 0.0)),
                               log_face_mean=True)
    stage = GP.SimplecStage(dm, inputs, GP.SimplecOptions(sweeps=2, height=CO.HEIGHT,
                                                          pressure_reltol=pressure_reltol))
    q0 = CO.first_guess(tile)
    q = GP.DeviceState(tile.geom)
    for n in FIELDS:
        q[n].data.copy_(torch.from_numpy(q0[n]))
    return tile, dm, stage, q


def _rel_l2(q, gold):
    out = {}
    for n in FIELDS:
        a, b = q[n].numpy()[0, 2:-2, 2:-2, 2:-2], gold[f"q_{n}"][0, 2:-2, 2:-2, 2:-2]
        den = float(np.linalg.norm(b))
        out[n] = float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a))
    return out


def test_newton_steps_on_the_column_converge_to_round_off(cuda):
    """From the column's SIMPLEC state after 600 iterations (a 2-3 order
    drop), continuation steps (CFL 50 doubling to the cap, then infinite
    dtau: plain Newton) with the SIMPLEC preconditioner: every step accepted,
    |R| ten orders down to round-off, the last steps at the inexact Newton
    rate the Krylov tolerance sets (1e-2 per step, where SIMPLEC's iterations
    gain 1e-2 in about 90), at the composite's converged state; a plain
    Newton step from there keeps it."""
    from ablsteadyrans_amd import newton as NW
    from ablsteadyrans_amd import physics as GP
    tile, dm, stage, q = _column_stage()
    drv = GP.DriverOptions(max_iterations=600, cfl0=1.0, cfl_max=50.0, residual_tol=1.0e-8,
                           residual_drop_orders=6.0)
    r = stage.solve(q, drv)
    print(f"SIMPLEC: {r.iterations} iterations, relative residual {r.relative_residual:.2e}", flush=True)
    nk = NW.NewtonStage(stage, NW.NewtonOptions(gmres_restart=60, start_cfl=50.0))
    norms = [nk.residual_norm(q)]
    for it in range(30):
        ok, rec = nk.continuation_step(q)
        norms.append(nk.residual_norm(q))
        st = nk.continuation_state
        print(f"  PTC step {it}: |R| {rec['residual']:.3e} -> {norms[-1]:.3e}, CFL "
              f"{'inf' if rec['infinite'] else format(rec['cfl'], '.3g')}, lambda {rec['lambda']}, GMRES "
              f"{rec['gmres_iterations']}, {rec['outcome']}", flush=True)
        assert ok, rec
        if norms[-1] <= 1e-10 * norms[0]:
            break
    assert norms[-1] <= 1e-10 * norms[0], norms
    ratios = [b / a for a, b in zip(norms, norms[1:]) if a > 0]
    assert max(ratios[-3:]) <= 0.02, ratios
    assert st["infinite"] == 1
    info = nk.step(q)
    print(f"  plain Newton from there: |R| {info['residual_norm']:.3e}, applied {info['applied']}", flush=True)
    errs = _rel_l2(q, np.load(GOLD))
    print("  vs the composite's converged state:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= 1e-6, errs


def test_simplec_newton_run_of_the_column_deck(cuda):
    """steady.stage = simplec_newton on the Neutral_Column deck (its
    criterion: residual_tol 1e-8, a 6-order drop): SIMPLEC, the switch after
    20 iterations at r <= 1e-4 with the monitors settled, the Newton stage to
    the criterion; fewer iterations than the deck's 2231 with SIMPLEC alone,
    the composite's converged state within 1e-6."""
    from ablsteadyrans_amd import newton as NW
    from ablsteadyrans_amd import physics as GP
    tile, dm, stage, q = _column_stage()
    drv = GP.DriverOptions(max_iterations=20000, cfl0=1.0, cfl_max=50.0, residual_tol=1.0e-8,
                           residual_drop_orders=6.0)
    nk = NW.NewtonStage(stage, NW.NewtonOptions(gmres_restart=60, start_cfl=50.0))
    res = nk.solve(q, drv, NW.SwitchOptions())
    print(f"simplec_newton: {res}", flush=True)
    assert res["status"] == 1, res
    assert res["first_switch"] > 0 and res["newton_iterations"] > 0
    assert res["iterations"] < 2231, res
    errs = _rel_l2(q, np.load(GOLD))
    print("  vs the composite's converged state:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= 1e-6, errs


def _pc_case(case):
    """The two tiles of the stage's oracle tests: a periodic hill with
    stratification, rotation and the wall model; an inflow/outflow pair along
    x with slip and no-slip walls along y."""
    from ablsteadyrans_amd import physics as GP
    from tests import physics_helpers as PHY
    from tests import test_gpu_physics as TGP
    from tests.test_gpu_stage import _inputs
    if case == "hill":
        tile = PHY.hill_mesh(n=(16, 12, 10), height=60.0, half=220.0)
        q = PHY.random_state(tile, 11, U=6.0, amp=0.05)
        phys = P.Physics(closure=P.ClosureOptions(tke_floor=1e-6),
                         bc=P.Boundary(zlo=P.SURFACE_LAYER, zhi=P.SLIP_WALL, zhi_theta_grad=0.003),
                         surface=P.SurfaceOptions(z0=0.1),
                         forcing=P.Forcing(coriolis=True, latitude=45.0, geostrophic=True, ug=6.0, vg=1.0),
                         theta_ref_z=(0.0, 1000.0), theta_ref=(300.0, 303.0))
        return tile, q, phys, _inputs(phys), 600.0
    tile, _m, q, phys = TGP._open_setup(0, seed=11)
    phys.closure = P.ClosureOptions(tke_floor=1e-6)
    q["w"] *= 0.05
    q["tke"] += 0.5   # clear of the stage's k floor (apply has none)
    gi = TGP._open_inputs(phys)
    gi.closure = GP.ClosureOptions(tke_floor=1e-6)
    return tile, q, phys, gi, 500.0


def _seam_sync(a, n, geom):
    """OverrideSync of a valid-point array: the hi copies of a periodic
    direction's faces take the lo copies' values."""
    if n == "u" and geom.periodic[0]:
        a[..., -1] = a[..., 0]
    if n == "v" and geom.periodic[1]:
        a[..., -1, :] = a[..., 0, :]
    return a


@pytest.mark.parametrize("case", ["hill", "open"])
def test_stage_denominator_apply_is_one_simplec_iteration(cuda, case):
    """NewtonPreconditioner.ApplyIsOneSimplecIteration
    (Tests/Unit_Tests/test_NewtonPreconditioner.cpp:316-411): with the stage's
    denominators, apply(-R(q)) at a non-uniform step (local dtau at CFL 5)
    equals the increment of one iteration of the SIMPLEC stage from q at that
    step, within 1e-9 of each field's largest increment (the pressure modulo
    its J-weighted mean, the gauge); on the Dirichlet rows of u, v and w
    (identity rows of the Newton system) dq = r exactly."""
    import torch
    from ablsteadyrans_amd import newton as NW
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd.fab import Fab
    from ablsteadyrans_amd.poisson import DeviceMesh
    from oracle import oracle as orc
    from tests import physics_helpers as PHY
    tile, q, phys, gi, height = _pc_case(case)
    m = PHY.omesh(tile)
    g = m.geom
    dm = DeviceMesh(tile)
    stage = GP.SimplecStage(dm, gi, GP.SimplecOptions(pressure_reltol=1e-12, sweeps=2, tke_floor=1e-6,
                                                      height=height))
    ds = _dstate(tile, q)
    stage.init(ds)
    GP.fill_ghosts(dm, ds, gi)
    B1 = (-1, -1, -1)
    q1 = {n: ds[n].numpy()[:, 1:-1, 1:-1, 1:-1].copy() for n in FIELDS}
    dtau = orc.local_dtau(q1["u"], q1["v"], q1["w"], B1, m.z, B1, m.J, B1, g, 5.0)
    assert dtau.max() > 1.2 * dtau.min()
    dt = Fab(torch.from_numpy(dtau[None].copy()).cuda(), tuple(g.lo))
    # r = -R(q), its seam copies equal
    R = NW.residual_fields(tile.geom)
    GP.residual(dm, ds, gi, R)
    torch.cuda.synchronize()
    r = NW.residual_fields(tile.geom)
    rnp = {}
    for n in FIELDS:
        rnp[n] = _seam_sync(-R[n].numpy(), n, tile.geom)
        r[n].data.copy_(torch.from_numpy(rnp[n]))
    nk = NW.NewtonStage(stage, NW.NewtonOptions(pc_denominator=NW.STAGE_DENOMINATOR))
    nk.pc_setup(ds, dt)
    dq = NW.residual_fields(tile.geom)
    nk.pc_apply(r, dq)
    torch.cuda.synchronize()
    # one stage iteration from q at the same step
    qs = GP.DeviceState(tile.geom)
    for n in FIELDS:
        qs[n].data.copy_(ds[n].data)
    info = stage.iterate(qs, dt)
    torch.cuda.synchronize()
    assert info.tke.clipped == 0
    Jv = tile.detJ[1:-1, 1:-1, 1:-1]
    eqs = (P.EQ_U, P.EQ_V, P.EQ_W)
    for f, n in enumerate(FIELDS):
        d = dq[n].numpy()[0]
        inc = _seam_sync(qs[n].numpy()[0, 2:-2, 2:-2, 2:-2] - ds[n].numpy()[0, 2:-2, 2:-2, 2:-2], n, tile.geom)
        if n == "p":
            assert Jv.shape == d.shape
            d = d - (Jv * d).sum() / Jv.sum()
            inc = inc - (Jv * inc).sum() / Jv.sum()
        if f < 3:
            mask = P.dirichlet_rows(eqs[f], m, phys)[0]
            assert mask.shape == d.shape
            assert np.array_equal(d[mask > 0], rnp[n][0][mask > 0]), n
            if n == "w":
                assert mask.sum() > 0
            d, inc = d * (1 - mask), inc * (1 - mask)
        scale = float(np.abs(inc).max())
        err = float(np.abs(d - inc).max())
        print(f"{case} {n}: max |increment| {scale:.3e}, max |apply - increment| {err:.3e}", flush=True)
        assert scale > 0, n
        assert err <= 1e-9 * scale, (n, err, scale)
    # the diagonal denominators differ (the pressure increment no longer
    # scales with 1/dtau)
    nd = NW.NewtonStage(stage, NW.NewtonOptions())
    nd.pc_setup(ds, dt)
    dq2 = NW.residual_fields(tile.geom)
    nd.pc_apply(r, dq2)
    torch.cuda.synchronize()
    assert not np.allclose(dq2["p"].numpy(), dq["p"].numpy(), rtol=1e-3, atol=0)


def _base_value(f, x, y, z, Lx=900.0, Ly=600.0, H=400.0):
    """A smooth non-uniform state (Tests/Unit_Tests/test_NewtonPreconditioner.cpp:96-107)."""
    ax, ay, az = 2 * np.pi * x / Lx, 2 * np.pi * y / Ly, np.pi * z / H
    return (8 + 2 * np.sin(ax + 0.3) + z / H, 3 + np.cos(ay) + 0.5 * np.sin(az),
            0.5 * np.sin(az) * (1 + 0.2 * np.cos(ax)), 5 * np.sin(ax) * np.cos(ay), 300 + 0.01 * z + 0.5 * np.sin(ax),
            0.4 + 0.1 * np.cos(ay))[f]


def _open_hill(hill=30.0):
    """The open hill of NewtonPreconditioner.IterationsBoundedOverCfl
    (test_NewtonPreconditioner.cpp:163-211): 9 x 6 x 8 cells on 900 x 600 x
    400 m over a 30 m hill, inflow (log law, unbounded length) and outflow
    along x, periodic y, the wall law on the ground, a slip top, upwind
    convection, no buoyancy, fixed diffusivities of 2 m^2/s. The hill is
    round here (the reference's is a ridge along y)."""
    import torch
    from ablsteadyrans_amd import initializer as GI
    from ablsteadyrans_amd import mesh as M
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd.fab import Fab
    from tests import physics_helpers as PHY
    n, hi = (9, 6, 8), (900.0, 600.0, 400.0)
    geom = M.geometry(n, (0.0, 0.0, 0.0), hi, (0, 1))
    zl = M.make_zlevels(n[2], hi[2])
    tile = M.build_mesh(geom, (0.0, 0.0, 0.0), hi, zl, M.Terrain(kind="hills", hills=[(450.0, 300.0, hill, 150.0)]))
    q = GP.DeviceState(tile.geom)
    for f, name in enumerate(FIELDS):
        X, Y, Z = PHY.positions(tile, GP.LOCS[f])
        q[name].data.copy_(torch.from_numpy(_base_value(f, X, Y, Z)[None].copy()))
    diff = []
    for _ in range(3):
        d = Fab.at_loc(tile.geom, 0, 1, 1)
        d.data.fill_(2.0)
        diff.append(d)
    gi = GP.ResidualInputs(closure=None, nu=diff[0], gamma_theta=diff[1], gamma_tke=diff[2],
                           surface=GP.SurfaceLayerOptions(z0=0.1), forcing=GP.ForcingOptions(gravity=0.0),
                           faces=(GP.INFLOW, GP.PERIODIC, GP.SURFACE_LAYER, GP.OUTFLOW, GP.PERIODIC, GP.SLIP_WALL),
                           theta_ref_z=(-1000.0, 5000.0), theta_ref=(300.0, 300.0),
                           momentum_limiter=GP.UPWIND, theta_limiter=GP.UPWIND,
                           inflow=GI.InflowProfile(kind=GI.LOG_LAW, unbounded=True))
    return tile, q, gi, diff[0]


def _open_hill_steps(cfl, pc, denominator=1):
    """Four Newton steps at a fixed CFL (dtau from each step's state), GMRES(200)
    at Krylov tolerance 1e-3; the GMRES iterations of each step."""
    import torch
    from ablsteadyrans_amd import newton as NW
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd import pseudotime as PT
    from ablsteadyrans_amd.fab import Fab
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile, q, gi, nu = _open_hill()
    dm = DeviceMesh(tile)
    stage = GP.SimplecStage(dm, gi, GP.SimplecOptions(sweeps=2, height=400.0))
    GP.fill_ghosts(dm, q, gi)
    nk = NW.NewtonStage(stage, NW.NewtonOptions(gmres_restart=200, krylov_rtol=1e-3, preconditioner=pc,
                                                pc_denominator=denominator, cfl_max=1e5, start_cfl=10.0))
    dtau = Fab.at_loc(tile.geom, 0, 1, 0)
    its = []
    for _ in range(4):
        PT.local_dtau(dm, q["u"], q["v"], q["w"], cfl, dtau, nu=nu)
        info = nk.step(q, dtau)
        assert info["applied"], (cfl, pc, denominator, info)
        its.append(info["gmres_iterations"])
    torch.cuda.synchronize()
    return its


def test_preconditioner_iterations_bounded_over_cfl(cuda):
    """NewtonPreconditioner.IterationsBoundedOverCfl
    (test_NewtonPreconditioner.cpp:558-575): with the diagonal denominators
    the mean GMRES iterations per Newton step stay within 1.5 x their CFL-10
    value up to CFL 1e4; with the stage's denominators they grow at least
    1.5-fold to CFL 1e4 and exceed the diagonal form's five-fold there; the
    identity needs at least five times the diagonal form's at CFL 10 and
    1e4."""
    from ablsteadyrans_amd import newton as NW
    cfls = (1e1, 1e2, 1e3, 1e4)
    sc, st, ids = [], [], []
    for cfl in cfls:
        ends = cfl in (cfls[0], cfls[-1])
        sc.append(float(np.mean(_open_hill_steps(cfl, NW.SIMPLEC))))
        st.append(float(np.mean(_open_hill_steps(cfl, NW.SIMPLEC, NW.STAGE_DENOMINATOR))) if ends else 0.0)
        ids.append(float(np.mean(_open_hill_steps(cfl, NW.IDENTITY))) if ends else 0.0)
        print(f"CFL {cfl:g}: mean GMRES iterations simplec {sc[-1]}, stage's d {st[-1]}, identity {ids[-1]}",
              flush=True)
    for n in range(1, len(cfls)):
        assert sc[n] < 1.5 * sc[0], (cfls[n], sc)
    assert st[3] >= 1.5 * st[0], st
    assert st[3] >= 5 * sc[3], (st, sc)
    assert ids[0] >= 5 * sc[0] and ids[3] >= 5 * sc[3], (ids, sc)


def _continuation_run(pc, max_steps=30, tol=1e-8):
    from ablsteadyrans_amd import newton as NW
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile, q, gi, _nu = _open_hill()
    dm = DeviceMesh(tile)
    stage = GP.SimplecStage(dm, gi, GP.SimplecOptions(sweeps=2, height=400.0))
    GP.fill_ghosts(dm, q, gi)
    nk = NW.NewtonStage(stage, NW.NewtonOptions(gmres_restart=200, preconditioner=pc))
    r0 = nk.residual_norm(q)
    steps = gmres = 0
    r = r0
    for _ in range(max_steps):
        ok, rec = nk.continuation_step(q)
        assert ok, rec
        steps += 1
        gmres += rec["gmres_iterations"]
        r = nk.residual_norm(q)
        if r <= tol * r0:
            break
    return q, r0, r, steps, gmres


def test_continuation_identity_versus_simplec(cuda):
    """NewtonPreconditioner.ContinuationIdentityVersusSimplec
    (test_NewtonPreconditioner.cpp:586-620): the pseudo-transient
    continuation on the open hill with GMRES(200) converges to 1e-8 of its
    first residual with either preconditioner, to the same state (1e-6 of
    the speed scale), the SIMPLEC preconditioner in at most as many steps
    with at most a fifth of the GMRES iterations (those of each step's
    accepted attempt)."""
    from ablsteadyrans_amd import newton as NW
    qa, r0a, ra, sa, ga = _continuation_run(NW.IDENTITY)
    qb, r0b, rb, sb, gb = _continuation_run(NW.SIMPLEC)
    print(f"identity: |R0| {r0a:.4e}, {sa} steps, {ga} GMRES iterations, |R| {ra:.3e}", flush=True)
    print(f"simplec:  |R0| {r0b:.4e}, {sb} steps, {gb} GMRES iterations, |R| {rb:.3e}", flush=True)
    assert r0a == r0b
    assert ra <= 1e-8 * r0a and rb <= 1e-8 * r0b
    assert sb <= sa and 5 * gb <= ga, (sa, sb, ga, gb)
    speed = float(max(np.abs(qa[n].numpy()).max() for n in ("u", "v", "w")))
    for n in FIELDS:
        a, b = qa[n].numpy()[0, 2:-2, 2:-2, 2:-2], qb[n].numpy()[0, 2:-2, 2:-2, 2:-2]
        err = float(np.abs(a - b).max())
        print(f"  {n}: max |identity - simplec| {err:.2e}", flush=True)
        if n in ("u", "v", "w"):
            assert err <= 1e-6 * speed, (n, err, speed)
