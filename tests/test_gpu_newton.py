"""The Newton-Krylov finish on the GPU (Source/Newton.*, Continuation.*,
NewtonPreconditioner.*, NewtonStage.*): the homogeneous ghost fill bit for
bit against the oracle's; J v against the oracle's forward difference of
the same residual (the same eps, bit for bit) and exactly linear in powers of
two; continuation steps on the reference's Neutral_Column deck from a
SIMPLEC state converging ten orders at the Krylov tolerance's rate; and the steady.stage = simplec_newton run of
that deck reaching the deck's criterion in fewer iterations than SIMPLEC
alone, at the state the oracle's composite converges to
(tests/golden/column_steady.npz) within 1e-6 relative L2."""
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
