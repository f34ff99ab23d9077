"""Convection and diffusion of the oracle (oracle/transport.c) pinned after
Tests/Unit_Tests/test_Convection.cpp and test_Diffusion.cpp, and against the
independently written implicit operator (test_ImplicitOperator.cpp:132-217):
  * unlimited MUSCL (van Albada, eps >> differences) is second order over the
    hill for the scalar (both forms) and the three momentum control volumes;
    first-order upwind is first order;
  * scalar and momentum diffusion are second order (closed-form operators of
    the manufactured fields with constant diffusivity; >= 1.9 on the finest
    pair, >= 1.85 on the coarsest);
  * bounded upwind convection equals A1 q with diffusivity 0 (every location);
    on a flat mesh -D(q) = A1 q with transport 0, and -D_c of a field with
    only component c equals A1 u_c;
  * the bounded form of a constant is zero; the limiters' closed forms.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import massflux_helpers as H

B1, B2 = (-1, -1, -1), (-2, -2, -2)
LOCS = (0, 1, 2, 4)


def _scalar(X, Y, Z, fl):
    return 2.0 + np.cos(fl.a * X) * np.cos(fl.b * Y) * np.cos(fl.c * Z)


def _field(fl, loc, X, Y, Z):
    return _scalar(X, Y, Z, fl) if loc == 0 else fl.vel({1: 0, 2: 1, 4: 2}[loc], X, Y, Z)


def _D1(f, X, Y, Z, d, h=1e-2):
    e = [0.0, 0.0, 0.0]
    e[d] = 1.0
    P = lambda s: f(X + s * e[0], Y + s * e[1], Z + s * e[2])
    return (8 * (P(h) - P(-h)) - (P(2 * h) - P(-2 * h))) / (12 * h)


def _transport(m, fl):
    vel = H.velocity(m, fl, grow=2)
    Fx, Fy, Fz, fb = orc.transport_fluxes(*vel, B2, m.z_nd, B1, m.ax, m.ay, B1, H.ogeom(m))
    return [Fx[None].copy(), Fy[None].copy(), Fz[None].copy()], fb


def _conv_error(mm, loc, limiter, eps, bounded):
    m = H.hill_mesh(mm)
    fl = H.Flow()
    F, fb = _transport(m, fl)
    q = _field(fl, loc, *H.positions(m, loc, 2))[None].copy()
    out = orc.convection(loc, q, B2, F, fb, m.detJ, B1, H.ogeom(m), limiter, eps, bounded)
    X, Y, Z = H.positions(m, loc, 0)
    qf = lambda a, b, c: _field(fl, loc, a, b, c)
    ex = np.zeros_like(X)
    for d in range(3):
        if bounded:
            ex += fl.vel(d, X, Y, Z) * _D1(qf, X, Y, Z, d)
        else:
            ex += _D1(lambda a, b, c: fl.vel(d, a, b, c) * qf(a, b, c), X, Y, Z, d)
    err = (out - ex)[1:-1]
    return np.sqrt(np.mean(err ** 2)), np.max(np.abs(err))


@pytest.mark.parametrize("loc,bounded", [(0, False), (0, True), (1, False), (2, False), (4, False)])
def test_unlimited_muscl_is_second_order(loc, bounded):
    e = [_conv_error(m, loc, orc.VAN_ALBADA, 1e6, bounded) for m in range(3)]
    assert np.log2(e[0][0] / e[1][0]) >= 1.9 and np.log2(e[1][0] / e[2][0]) >= 1.9, e
    assert np.log2(e[1][1] / e[2][1]) >= 1.9, e


def test_upwind_is_first_order():
    e = [_conv_error(m, 0, orc.UPWIND, 1e-5, True) for m in range(3)]
    assert abs(np.log2(e[1][0] / e[2][0]) - 1.0) <= 0.1, e


def test_diffusion_is_second_order():
    fl = H.Flow()
    K2 = fl.a ** 2 + fl.b ** 2 + fl.c ** 2
    Kd = 0.3 * fl.a + 0.2 * fl.b + 0.1 * fl.c
    G = 2.5
    es, em = [], []
    for mm in range(3):
        m = H.hill_mesh(mm)
        nx, ny, nz = m.geom.n
        g = np.full((1, nz + 2, ny + 2, nx + 2), G)
        q = _scalar(*H.positions(m, 0, 1), fl)[None].copy()
        d = orc.scalar_diffusion(q, B1, g, B1, m.z_nd, B1, m.detJ, B1, H.ogeom(m))
        X, Y, Z = H.positions(m, 0, 0)
        ex = -G * K2 * (_scalar(X, Y, Z, fl) - 2.0)
        es.append(np.sqrt(np.mean(((d - ex)[1:-1]) ** 2)))
        vel = H.velocity(m, fl, grow=2)
        D = orc.momentum_diffusion(vel, B2, g, B1, m.z_nd, B1, m.detJ, B1, H.ogeom(m))
        e = 0.0
        for c, (amp, kc) in enumerate(((0.3, fl.a), (0.2, fl.b), (0.1, fl.c))):
            X, Y, Z = H.positions(m, 1 << c, 0)
            base = fl.vel(c, X, Y, Z) - (1.0 if c == 0 else 0.0)   # the oscillating part
            # d_c (div u) = -Kd kc times the same trigonometric product as u_c's
            ex = G * (-K2 * base - Kd * kc * base / amp)
            e = max(e, np.sqrt(np.mean(((D[c] - ex)[1:-1]) ** 2)))
        em.append(e)
    for e in (es, em):   # the coarsest pair (16 x 12 x 8) is slightly pre-asymptotic
        assert np.log2(e[0] / e[1]) >= 1.85 and np.log2(e[1] / e[2]) >= 1.9, e


def _random(m, loc, seed, lo=-1.0, hi=2.0, grow=2):
    nx, ny, nz = m.geom.n
    nd = [(loc >> d) & 1 for d in range(3)]
    return np.random.default_rng(seed).uniform(lo, hi, size=(1, nz + nd[2] + 2 * grow, ny + nd[1] + 2 * grow,
                                                             nx + nd[0] + 2 * grow))


def test_bounded_upwind_convection_is_the_implicit_operator():
    m = H.hill_mesh(1, smooth_ghosts=False)
    nx, ny, nz = m.geom.n
    F = [_random(m, 1 << d, 41 + d, -20, 20, 1) for d in range(3)]
    g0 = np.zeros((1, nz + 2, ny + 2, nx + 2))
    for loc in LOCS:
        q = _random(m, loc, 51, grow=2)
        Cq = orc.convection(loc, q, B2, F, B1, m.detJ, B1, H.ogeom(m), orc.UPWIND, 1e-5, True)
        A = orc.implicit_coefficients(loc, F, B1, g0, B1, m.z_nd, B1, m.detJ, B1, H.ogeom(m))
        Aq = orc.apply_implicit(loc, A, q[:, 1:-1, 1:-1, 1:-1].copy(), B1, H.ogeom(m))
        assert np.max(np.abs(Cq - Aq)) <= 1e-12 * np.max(np.abs(Cq)), loc
        # the bounded form of a constant is zero
        c1 = orc.convection(loc, np.full_like(q, 3.0), B2, F, B1, m.detJ, B1, H.ogeom(m), orc.VAN_LEER, 1e-5, True)
        assert np.max(np.abs(c1)) <= 1e-12 * np.max(np.abs(Cq)), loc


def test_flat_mesh_diffusion_is_the_implicit_operator():
    m = H.hill_mesh(0, smooth_ghosts=False, flat=True)
    nx, ny, nz = m.geom.n
    F0 = [np.zeros((1, nz + 2 + (d == 2), ny + 2 + (d == 1), nx + 2 + (d == 0))) for d in range(3)]
    nu = _random(m, 0, 61, 0.0, 50.0, 1)
    g = H.ogeom(m)
    q = _random(m, 0, 71, grow=1)
    Dq = orc.scalar_diffusion(q, B1, nu, B1, m.z_nd, B1, m.detJ, B1, g)
    A = orc.implicit_coefficients(0, F0, B1, nu, B1, m.z_nd, B1, m.detJ, B1, g)
    Aq = orc.apply_implicit(0, A, q, B1, g)
    assert np.max(np.abs(Dq + Aq)) <= 1e-12 * np.max(np.abs(Dq))
    for c in range(3):
        vel = [np.zeros_like(_random(m, 1 << d, 0, grow=2)) for d in range(3)]
        vel[c] = _random(m, 1 << c, 81 + c, grow=2)
        D = orc.momentum_diffusion(vel, B2, nu, B1, m.z_nd, B1, m.detJ, B1, g)
        A = orc.implicit_coefficients(1 << c, F0, B1, nu, B1, m.z_nd, B1, m.detJ, B1, g)
        Au = orc.apply_implicit(1 << c, A, vel[c][:, 1:-1, 1:-1, 1:-1].copy(), B1, g)
        rows = slice(1, -1) if c == 2 else slice(None)   # the ground row of w is a half volume
        assert np.max(np.abs((D[c] + Au)[rows])) <= 1e-12 * np.max(np.abs(D[c])), c


def test_limiter_closed_forms_and_eps_refusal():
    m = H.hill_mesh(0)
    nx, ny, nz = m.geom.n
    F = [np.ones((1, nz + 2 + (d == 2), ny + 2 + (d == 1), nx + 2 + (d == 0))) for d in range(3)]
    # q linear in i: both limiters give the central slope, the face value is the midpoint
    I = np.arange(-2, nx + 2)[None, None, None, :].astype(float)
    q = np.broadcast_to(I, (1, nz + 4, ny + 4, nx + 4)).copy()
    for lim in (orc.VAN_LEER, orc.VAN_ALBADA):
        G, b = orc.convective_fluxes(0, q, B2, F, B1, H.ogeom(m), lim, 1e-5)
        assert np.allclose(G[0][0], (np.arange(0, nx + 1) - 0.5)[None, None, :], rtol=0, atol=1e-12)
    for eps in (0.0, -1.0, float("nan"), 1e300):
        with pytest.raises(orc.OracleError):
            orc.convective_fluxes(0, q, B2, F, B1, H.ogeom(m), orc.VAN_ALBADA, eps)
