"""The oracle's pieces of the Newton-Krylov finish, pinned by what the
mathematics fixes: the homogeneous ghost fill is the linear part of the fill
(fill(q + v) - fill(q) without the wall law, exact scaling, zero at the
inflow profile's points), and the forward-difference J v is the residual's
directional derivative (a Taylor remainder of second order, exact scaling
by powers of two)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import physics as P
from tests import newton_helpers as NH
from tests import physics_helpers as PHY
from tests.test_gpu_physics import OPEN, _open_setup

FIELDS = NH.FIELDS


def _copy(q):
    return {n: a.copy() for n, a in q.items()}


def _slip_ground(phys):
    b = phys.bc
    b.zlo = P.SLIP_WALL
    return phys


@pytest.mark.parametrize("case", range(len(OPEN)))
def test_homogeneous_fill_is_the_fill_s_linear_part(case):
    tile, m, q, phys = _open_setup(case)
    phys = _slip_ground(phys)
    v = PHY.random_state(tile, 17, U=1.0, amp=0.4)
    a = _copy(q)
    P.fill_ghosts(a, m, phys)
    b = {n: q[n] + v[n] for n in FIELDS}
    P.fill_ghosts(b, m, phys)
    h = _copy(v)
    P.fill_ghosts_homogeneous(h, m, phys)
    for n in FIELDS:
        d = b[n] - a[n]
        scale = max(np.abs(q[n]).max(), np.abs(v[n]).max(), 1.0)
        assert np.abs(d - h[n]).max() <= 1e-12 * scale, n
    # exact scaling by two
    h2 = {n: 2.0 * v[n] for n in FIELDS}
    P.fill_ghosts_homogeneous(h2, m, phys)
    for n in FIELDS:
        assert np.array_equal(h2[n], 2.0 * h[n]), n


def test_homogeneous_fill_zeroes_the_inflow_points_and_the_wall_gradient():
    tile, m, q, phys = _open_setup(0)
    h = _copy(q)
    P.fill_ghosts_homogeneous(h, m, phys)
    # xlo is an inflow face: the face point of u and the two ghost layers are 0
    assert np.all(h["u"][0, 2:-2, 2:-2, 0:3] == 0.0)
    assert np.all(h["theta"][0, 2:-2, 2:-2, 0:2] == 0.0)
    # the surface layer's gradient is 0: the centred difference about the
    # first cell (BoundaryConditions.cpp:512-513) makes the ghosts k = -1 and
    # k = -2 the value of cell k = 1
    assert np.array_equal(h["u"][0, 1, 2:-2, 3:-3], h["u"][0, 3, 2:-2, 3:-3])
    assert np.array_equal(h["u"][0, 0, 2:-2, 3:-3], h["u"][0, 3, 2:-2, 3:-3])


def _jvp_case(seed=5):
    tile = PHY.hill_mesh(n=(10, 8, 8), height=60.0, half=180.0)
    m = PHY.omesh(tile)
    q = PHY.random_state(tile, seed, U=6.0, amp=0.1)
    phys = P.Physics(closure=P.ClosureOptions(tke_floor=1e-6),
                     bc=P.Boundary(zlo=P.SURFACE_LAYER, zhi=P.SLIP_WALL, zhi_theta_grad=0.003),
                     surface=P.SurfaceOptions(z0=0.1),
                     forcing=P.Forcing(coriolis=True, latitude=45.0, geostrophic=True, ug=6.0, vg=1.0),
                     theta_ref_z=(0.0, 1000.0), theta_ref=(300.0, 303.0))
    P.fill_ghosts(q, m, phys)
    v = PHY.random_state(tile, seed + 1, U=0.5, amp=0.2)
    v["theta"] -= 300.0
    v["tke"] *= 0.01
    P.fill_ghosts_homogeneous(v, m, phys)
    return tile, m, q, v, phys


def test_jvp_is_the_directional_derivative():
    """|R(q + h v) - R(q) - h J v| falls as h^2: halving h about quarters
    it (a J v that misses a term leaves a remainder of first order, which
    halves)."""
    tile, m, q, v, phys = _jvp_case()
    R0 = P.residual(q, m, phys)[0]
    eps, Jv = NH.jvp(q, v, m, phys, R0=R0)
    assert eps > 0
    rem = []
    for h in (4e-3, 2e-3, 1e-3):
        hs = h / eps * eps
        qh = {n: q[n] + hs * v[n] for n in FIELDS}
        P.fill_ghosts(qh, m, phys)
        Rh = P.residual(qh, m, phys)[0]
        rem.append(sum(float(np.linalg.norm(Rh[n] - R0[n] - hs * Jv[n]) / NH.SCALES[n]) for n in FIELDS))
    print("Taylor remainders", rem)
    assert 3.0 <= rem[0] / rem[1] <= 7.0 and 3.0 <= rem[1] / rem[2] <= 7.0


def test_jvp_scales_exactly_by_powers_of_two():
    """eps of 4 v is eps / 4: the perturbed point and the difference are the
    same, so J(4 v) = 4 J v bit for bit."""
    tile, m, q, v, phys = _jvp_case(seed=8)
    R0 = P.residual(q, m, phys)[0]
    e1, J1 = NH.jvp(q, v, m, phys, R0=R0)
    v4 = {n: 4.0 * v[n] for n in FIELDS}
    e4, J4 = NH.jvp(q, v4, m, phys, R0=R0)
    assert e4 == e1 / 4
    for n in FIELDS:
        assert np.array_equal(J4[n], 4.0 * J1[n]), n
    z = {n: 0.0 * v[n] for n in FIELDS}
    e0, J0 = NH.jvp(q, z, m, phys, R0=R0)
    assert e0 == 0 and all(not np.any(J0[n]) for n in FIELDS)
