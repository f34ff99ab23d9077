"""The HIP momentum predictor and scalar sweep against the oracle composites
(tests/simplec_helpers.py): bit for bit without the rotation (a_P, d, the
sweeps, the updated fields; plain and slope-aware w caps; z-line Jacobi and
LU-SGS), and to 1e-12 with the point-implicit rotation and buoyancy (the
per-level mean of the other residual is summed in another order; the
reference itself sums it with atomics)."""
import numpy as np
import pytest

from ablsteadyrans_amd.fab import LineSweepOptions
from oracle import oracle as orc
from tests import massflux_helpers as MH
from tests import simplec_helpers as S
from tests.test_oracle_implicit import _random_fields
from tests.test_oracle_predictor import _setup

pytestmark = pytest.mark.gpu
B = (-1, -1, -1)


def _fab(a, begin):
    import torch
    from ablsteadyrans_amd.fab import Fab
    return Fab(torch.from_numpy(np.ascontiguousarray(a)).cuda(), begin)


def _gpu_predictor(m, vel, R, A, dt, opt, mask=None, theta=None, R_theta=None, A_theta=None):
    from ablsteadyrans_amd import simplec as SP
    from ablsteadyrans_amd.fab import Fab
    from ablsteadyrans_amd.poisson import DeviceMesh
    lo = tuple(m.geom.lo)
    fv = [_fab(v, B) for v in vel]
    ap = [Fab.at_loc(m.geom, 1 << c, 1, 0) for c in range(3)]
    d = [Fab.at_loc(m.geom, 1 << c, 1, 0) for c in range(3)]
    res = SP.momentum_predictor(DeviceMesh(m), fv, [_fab(r[None], lo) for r in R], [_fab(a, lo) for a in A],
                                [_fab(x[None], lo) for x in dt], opt, ap, d,
                                theta=_fab(theta, B) if theta is not None else None,
                                R_theta=_fab(R_theta[None], lo) if R_theta is not None else None,
                                A_theta=_fab(A_theta, lo) if A_theta is not None else None,
                                mask=[_fab(x[None], lo) for x in mask] if mask is not None else None)
    return [f.numpy() for f in fv], [x.numpy()[0] for x in ap], [x.numpy()[0] for x in d], res


@pytest.mark.parametrize("kind,cap,margin", [(0, 0.0, 0.0), (0, 16.0, 0.0), (0, 16.0, 4.0), (1, 16.0, 4.0)])
def test_predictor_matches_the_composite_bitwise(cuda, kind, cap, margin):
    from ablsteadyrans_amd.simplec import PredictorOptions
    m = MH.hill_mesh(0, smooth_ghosts=False)
    F, g, A, R, dt, vel = _setup(m, 5)
    sweep = LineSweepOptions(kind=kind)
    opt = PredictorOptions(sweeps=3, w_vnn_cap=cap, w_vnn_slope_margin=margin, sweep=sweep)
    print(f"predictor kind={kind} cap={cap} margin={margin}", flush=True)
    gv, gap, gd, res = _gpu_predictor(m, vel, R, A, dt, opt)
    if kind == 0:
        rv, rap, rd, dmax, last = S.momentum_predictor(m, vel, R, A, dt, 3, w_vnn_cap=cap, margin=margin)
    else:
        # the composite with the oracle's line LU-SGS (lagged ghosts,
        # sequential order, no alternation: LineSweepOptions' defaults)
        with S.lusgs_sweeps():
            rv, rap, rd, dmax, last = S.momentum_predictor(m, vel, R, A, dt, 3, w_vnn_cap=cap, margin=margin)
    for c in range(3):
        assert np.array_equal(gap[c], rap[c]) and np.array_equal(gd[c], rd[c]), c
        assert np.array_equal(gv[c][0, 1:-1, 1:-1, 1:-1], rv[c][0, 1:-1, 1:-1, 1:-1]), c
    assert [res.max_du, res.max_dv, res.max_dw] == dmax
    assert [res.last_sweep_du, res.last_sweep_dv, res.last_sweep_dw] == last


def test_predictor_with_rotation_and_buoyancy(cuda):
    from ablsteadyrans_amd.simplec import PredictorOptions
    m = MH.hill_mesh(0, smooth_ghosts=False)
    F, g, A, R, dt, vel = _setup(m, 7)
    nx, ny, nz = m.geom.n
    At = orc.implicit_coefficients(0, F, B, g, B, m.z_nd, B, m.detJ, B, MH.ogeom(m))
    rng = np.random.default_rng(8)
    theta = rng.uniform(300, 303, size=(1, nz + 2, ny + 2, nx + 2))
    Rt = rng.uniform(-1e-3, 1e-3, size=(nz, ny, nx))
    mask = [np.zeros_like(x) for x in dt]
    mask[2][0] = mask[2][-1] = 1.0
    opt = PredictorOptions(sweeps=2, coriolis_f=1e-4, rotation_relax=0.3, gravity=9.81, theta0=300.0)
    gv, gap, gd, res = _gpu_predictor(m, vel, R, A, dt, opt, mask, theta, Rt, At)
    rv, rap, rd, dmax, last = S.momentum_predictor(m, vel, R, A, dt, 2, mask=mask, coriolis_f=1e-4, relax=0.3,
                                                   gravity=9.81, theta0=300.0, theta=theta, R_theta=Rt,
                                                   A_theta=At)
    for c in range(3):
        assert np.max(np.abs(gap[c] - rap[c]) / np.abs(rap[c])) <= 1e-12, c
        assert np.max(np.abs(gd[c] - rd[c])) <= 1e-12 * np.max(np.abs(rd[c])), c
        dv = gv[c][0, 1:-1, 1:-1, 1:-1] - rv[c][0, 1:-1, 1:-1, 1:-1]
        assert np.max(np.abs(dv)) <= 1e-12 * np.max(np.abs(rv[c])), c


def test_scalar_sweep_matches_the_composite_bitwise(cuda):
    from ablsteadyrans_amd import simplec as SP
    m = MH.hill_mesh(0, smooth_ghosts=False)
    F, g = _random_fields(m, 5.0, 20.0, 3)
    A = orc.implicit_coefficients(0, F, B, g, B, m.z_nd, B, m.detJ, B, MH.ogeom(m))
    nx, ny, nz = m.geom.n
    rng = np.random.default_rng(4)
    k = np.zeros((1, nz + 2, ny + 2, nx + 2))
    k[0, 1:-1, 1:-1, 1:-1] = rng.uniform(0.005, 0.02, size=(nz, ny, nx))
    R = rng.uniform(-1, 1, size=(nz, ny, nx))
    dt = rng.uniform(20, 60, size=(nz, ny, nx))
    lo = tuple(m.geom.lo)
    for eq, clip in ((SP.THETA, False), (SP.TKE, True)):
        fk = _fab(k, B)
        res = SP.scalar_sweep(m.geom, eq, fk, _fab(R[None], lo), _fab(A, lo), _fab(dt[None], lo), 2, floor=1e-3,
                              clip=clip)
        f, dq, ch, clipped = S.scalar_sweep(m, k, R, A, dt, 2, floor=1e-3, clip=clip)
        assert np.array_equal(fk.numpy()[0, 1:-1, 1:-1, 1:-1], f[0, 1:-1, 1:-1, 1:-1])
        assert res.max_dq == dq and res.last_sweep == ch and res.clipped == clipped
