"""The predictor and scalar-sweep composites (tests/simplec_helpers.py) after
Tests/Unit_Tests/test_Predictor.cpp and test_Simplec.cpp:
  * a_P = 1/dtau + diag(A1) and d = a_P - sum a_nb (test_Predictor.cpp:256-367);
    the Dirichlet rows (mask 1) get a_P = A_P and d = 0;
  * the plain w cap adds (a_ZM + a_ZP)/cap to both; the slope-aware cap is
    that on flat ground (test_Predictor.cpp:756-1081) and bounded by
    max(0, 1/(s r_c) - 1)/dtau on a slope;
  * with many sweeps the corrections solve (a_P - A_nb) delta = -R: the
    residual of the full linear system vanishes (the sweeps converge);
  * the rotation with a level-uniform other residual adds exactly
    -f tau_c <R_v> to u's right-hand side (the mean term alone);
  * the k sweep clips at the floor and counts.
"""
import numpy as np

from oracle import oracle as orc
from tests import massflux_helpers as MH
from tests import simplec_helpers as S
from tests.test_oracle_implicit import _random_fields

B = (-1, -1, -1)


def _setup(m, seed=1):
    F, g = _random_fields(m, 5.0, 20.0, seed)
    A = [orc.implicit_coefficients(1 << c, F, B, g, B, m.z_nd, B, m.detJ, B, MH.ogeom(m)) for c in range(3)]
    rng = np.random.default_rng(seed + 10)
    R = [rng.uniform(-1, 1, size=a.shape[1:]) for a in A]
    dt = [rng.uniform(20, 60, size=a.shape[1:]) for a in A]
    vel = [np.zeros((1, a.shape[1] + 2, a.shape[2] + 2, a.shape[3] + 2)) for a in A]
    return F, g, A, R, dt, vel


def test_predictor_diagonal_and_denominator():
    m = MH.hill_mesh(0, smooth_ghosts=False)
    F, g, A, R, dt, vel = _setup(m)
    ap, d = S.predictor_fields(A[0], dt[0])
    assert np.array_equal(ap, 1.0 / dt[0] + A[0][0])
    assert np.allclose(d, 1.0 / dt[0], rtol=1e-10)   # a_P = sum a_nb: d = 1/dtau up to round-off
    mask = np.zeros_like(dt[2])
    mask[0] = mask[-1] = 1.0
    ap, d = S.predictor_fields(A[2], dt[2], mask)
    assert np.array_equal(ap[0], A[2][0][0]) and np.all(d[0] == 0) and np.all(d[-1] == 0)


def test_w_caps():
    m = MH.hill_mesh(0, smooth_ghosts=False)
    F, g, A, R, dt, vel = _setup(m)
    _, aps0, ds0, _, _ = S.momentum_predictor(m, vel, R, A, dt, 1)
    _, aps, ds, _, _ = S.momentum_predictor(m, vel, R, A, dt, 1, w_vnn_cap=16.0)
    # (a + t) - a recovers t to an ulp of a
    assert np.max(np.abs((aps[2] - aps0[2]) - (A[2][5] + A[2][6]) / 16.0) / aps[2]) <= 1e-15
    flat = MH.hill_mesh(0, smooth_ghosts=False, flat=True)
    Ff, gf, Af, Rf, dtf, velf = _setup(flat)
    _, ap_plain, _, _, _ = S.momentum_predictor(flat, velf, Rf, Af, dtf, 1, w_vnn_cap=16.0)
    _, ap_slope, _, _, _ = S.momentum_predictor(flat, velf, Rf, Af, dtf, 1, w_vnn_cap=16.0, margin=4.0)
    assert np.array_equal(ap_plain[2], ap_slope[2])   # flat ground: K/V exactly
    # long steps: the slope bound max(0, 1/(s r_c) - 1)/dtau binds on the hill's flanks
    dtl = [x * 1e4 for x in dt]
    _, ap0l, _, _, _ = S.momentum_predictor(m, vel, R, A, dtl, 1)
    _, ap_s, _, _, _ = S.momentum_predictor(m, vel, R, A, dtl, 1, w_vnn_cap=16.0, margin=4.0)
    added = ap_s[2] - ap0l[2]
    assert np.all(added <= (A[2][5] + A[2][6]) / 16.0 + 1e-15 * ap_s[2])
    assert np.any(added < 0.5 * (A[2][5] + A[2][6]) / 16.0)


def test_many_sweeps_solve_the_linear_system():
    m = MH.hill_mesh(0, smooth_ghosts=False)
    F, g, A, R, dt, vel = _setup(m)
    # a periodic seam face is one unknown: its two copies carry the same row
    A[0][..., -1] = A[0][..., 0]
    R[0][..., -1] = R[0][..., 0]
    dt[0][..., -1] = dt[0][..., 0]
    A[1][..., -1, :] = A[1][..., 0, :]
    R[1][..., -1, :] = R[1][..., 0, :]
    dt[1][..., -1, :] = dt[1][..., 0, :]
    new, aps, ds, dmax, last = S.momentum_predictor(m, vel, R, A, dt, 400)
    for c in range(3):
        delta = new[c][0, 1:-1, 1:-1, 1:-1]
        q = np.zeros_like(new[c])
        q[0, 1:-1, 1:-1, 1:-1] = delta
        # neighbours across the periodic sides are images; above and below the column, nothing
        q = S.periodic_images(q, m.geom.periodic, (c == 0, c == 1))
        A1d = orc.apply_implicit(1 << c, A[c], q, B, MH.ogeom(m))   # A_P d - sum a_nb d_nb
        res = (aps[c] - A[c][0]) * delta + A1d + R[c]
        own = (slice(None), slice(0, res.shape[1] - (c == 1)), slice(0, res.shape[2] - (c == 0)))   # seam copy once
        assert np.max(np.abs(res[own])) <= 1e-9 * np.max(np.abs(R[c])), c


def test_rotation_mean_term_with_a_level_uniform_residual():
    m = MH.hill_mesh(0, smooth_ghosts=False)
    F, g, A, R, dt, vel = _setup(m)
    f = 1e-4
    Rv = np.broadcast_to(np.linspace(-1, 1, R[1].shape[0])[:, None, None], R[1].shape).copy()
    R2 = [R[0], Rv, R[2]]
    _, aps0, _, _, _ = S.momentum_predictor(m, vel, R2, A, dt, 1)
    _, aps, _, _, _ = S.momentum_predictor(m, vel, R2, A, dt, 1, coriolis_f=f, relax=0.3)
    nb = sum(A[0][n] for n in range(1, 7))
    net = np.maximum(A[0][0] - nb, 0.0)
    tau_c = 1.0 / np.maximum(f / 0.3, net)
    tau_own = 1.0 / (1.0 / dt[0] + net)
    assert np.max(np.abs((aps[0] - aps0[0]) - f * f * np.maximum(tau_c, tau_own)) / aps[0]) <= 1e-15


def test_k_sweep_clips_at_the_floor():
    m = MH.hill_mesh(0, smooth_ghosts=False)
    F, g = _random_fields(m, 5.0, 20.0, 3)
    A = orc.implicit_coefficients(0, F, B, g, B, m.z_nd, B, m.detJ, B, MH.ogeom(m))
    nx, ny, nz = m.geom.n
    rng = np.random.default_rng(4)
    k = np.zeros((1, nz + 2, ny + 2, nx + 2))
    k[0, 1:-1, 1:-1, 1:-1] = 0.01
    R = rng.uniform(-1, 1, size=(nz, ny, nx))
    dt = np.full((nz, ny, nx), 50.0)
    f, dq, ch, clipped = S.scalar_sweep(m, k, R, A, dt, 2, floor=1e-3, clip=True)
    assert clipped > 0 and np.min(f[0, 1:-1, 1:-1, 1:-1]) == 1e-3
