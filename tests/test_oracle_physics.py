"""The oracle's residual physics pinned to closed forms and to the
reference's own tests: heights above the ground on flat and hilly meshes,
the k-L closure in its neutral, stable and unstable branches against ERF's
relations (extern_erf/ERF_RANSClosure.H, written with branches), the neutral
surface layer's log law, the ghost-fill rules per face type, the rest state
over steep terrain (test_Residual.cpp:274-302), the closed-form sources
(test_Residual.cpp:566-608), and the wall and boundary rows."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import physics as P
from tests import physics_helpers as PHY

B1, B2 = (-1, -1, -1), (-2, -2, -2)


def test_height_above_ground_flat_and_hill():
    flat = PHY.flat_mesh()
    d = P.height_above_ground(PHY.omesh(flat))
    nz = flat.geom.n[2]
    dz = 800.0 / nz
    assert np.array_equal(d[0, 1:-1, 3, 3], (np.arange(nz) + 0.5) * dz)
    assert d[0, 0, 3, 3] == d[0, 1, 3, 3] and d[0, -1, 3, 3] == d[0, -2, 3, 3]   # vertical ghosts copied
    hill = PHY.hill_mesh()
    d = P.height_above_ground(PHY.omesh(hill))
    Z = hill.z_nd
    zc = (Z[1:-2, :-1, :-1] + Z[1:-2, 1:, :-1] + Z[1:-2, :-1, 1:] + Z[1:-2, 1:, 1:] + Z[2:-1, :-1, :-1] +
          Z[2:-1, 1:, :-1] + Z[2:-1, :-1, 1:] + Z[2:-1, 1:, 1:]) / 8
    zg = (Z[1, :-1, :-1] + Z[1, 1:, :-1] + Z[1, :-1, 1:] + Z[1, 1:, 1:]) / 4
    np.testing.assert_allclose(d[0, 1:-1], zc - zg[None], rtol=1e-13, atol=1e-12)


def _erf_point(k, N2, S2, l_g, o: P.ClosureOptions):
    """ERF's AL01 relations as ComputeTurbulentViscosityRANS composes them
    (extern_erf/ERF_RANSClosure.H), with branches."""
    c3 = o.Cmu0 ** 3
    inv_cb2 = 1.0 / (o.Cb * o.Cb)
    eps_n2 = np.finfo(float).eps

    def smooth(rt):
        if rt >= o.Rt_crit:
            return rt
        x, a = rt - o.Rt_crit, o.Rt_min - o.Rt_crit
        return o.Rt_crit + a * x / (x + a)

    if abs(N2) <= eps_n2:
        L = l_g
    elif N2 > eps_n2:
        L = math.sqrt(1.0 / (1.0 / (l_g * l_g) + inv_cb2 * N2 / k))
    else:
        rt = smooth(l_g * l_g * N2 / (k * c3 * c3))
        L = l_g * math.sqrt(1.0 - c3 * c3 * inv_cb2 * rt)
    eps = c3 * k ** 1.5 / L
    Rt = smooth(L * L * N2 / (k * c3 * c3))
    cmu = (o.Cmu0 + 0.108 * Rt) / (1.0 + 0.308 * Rt + 0.00837 * Rt * Rt)
    cmup = o.Cmu0 / (1 + 0.277 * Rt)
    nut = cmu * math.sqrt(k) * L
    return dict(L=L, eps=eps, nut=nut, nutp=cmup / cmu * nut, Ps=nut * S2, Pb=-(cmup / cmu * nut) * N2)


@pytest.mark.parametrize("dthdz,shear", [(0.0, 0.02), (0.004, 0.0), (-0.004, 0.0), (-0.05, 0.01)])
def test_closure_fields_match_erf_relations(dthdz, shear):
    """Flat mesh, u = a z, theta = 300 + b z: S^2 = a^2 and N^2 = g b / theta0
    in the interior; nu_t, gamma's, P_s, P_b, eps and the Patankar diagonal
    equal ERF's relations at the point (neutral, stable, unstable and the
    smoothed strongly unstable branch)."""
    tile = PHY.flat_mesh(n=(4, 4, 12), top=240.0)
    m = PHY.omesh(tile)
    q = PHY.zeros_state(m.geom)
    _, _, zu = PHY.positions(tile, 1)
    _, _, zc = PHY.positions(tile, 0)
    q["u"][0] = shear * zu
    q["theta"][0] = 300.0 + dthdz * zc
    q["tke"][:] = 0.3
    o = P.ClosureOptions()
    dist = P.height_above_ground(m)
    z0 = 0.1
    f = P.closure_evaluate(q, m, o, dist, z0, P.Forcing())
    N2 = 9.81 / 300.0 * dthdz
    for k in (2, 5, 9):
        lg = o.max_geom_lscale * (o.kappa * (dist[0, k + 1, 2, 2] + z0)) / (o.max_geom_lscale +
                                                                           o.kappa * (dist[0, k + 1, 2, 2] + z0))
        e = _erf_point(0.3, N2, shear * shear, lg, o)
        cell = (0, k, 1, 1)
        assert f["nu"][0, k + 1, 2, 2] == pytest.approx(e["nut"], rel=1e-12)
        assert f["gamma_theta"][0, k + 1, 2, 2] == pytest.approx(e["nutp"], rel=1e-12)
        assert f["gamma_tke"][0, k + 1, 2, 2] == pytest.approx(e["nut"], rel=1e-12)
        assert f["eps"][cell] == pytest.approx(e["eps"], rel=1e-12)
        assert f["Ps"][cell] == pytest.approx(e["Ps"], rel=1e-12, abs=1e-300)
        assert f["Pb"][cell] == pytest.approx(e["Pb"], rel=1e-9, abs=1e-18)
        assert f["dSdk"][cell] == pytest.approx((e["eps"] + max(-e["Pb"], 0.0)) / 0.3, rel=1e-9)
    # neutral closed form: L = l_g, nu_t = Cmu0 sqrt(k) l_g
    if dthdz == 0.0:
        lg = o.max_geom_lscale * (o.kappa * (dist[0, 4, 2, 2] + z0)) / (o.max_geom_lscale + o.kappa *
                                                                         (dist[0, 4, 2, 2] + z0))
        assert f["nu"][0, 4, 2, 2] == pytest.approx(o.Cmu0 * math.sqrt(0.3) * lg, rel=1e-14)
    # vertical ghosts copy the first and last cells
    assert np.array_equal(f["nu"][0, 0], f["nu"][0, 1]) and np.array_equal(f["nu"][0, -1], f["nu"][0, -2])


def test_neutral_surface_layer_is_the_log_law():
    """Flat ground, uniform U along x: u* = kappa U / ln((d0 + z0)/z0),
    d0 = dz/2, the stress u*^2 along x, k_w = u*^2 / Cmu0^2
    (test_SurfaceLayer.cpp:186-232, 396-451)."""
    tile = PHY.flat_mesh(n=(6, 4, 10), top=200.0)
    m = PHY.omesh(tile)
    q = PHY.zeros_state(m.geom)
    q["u"][:] = 6.0
    s = P.SurfaceOptions(z0=0.05)
    ss = np.zeros((P.SS_N, 1, 6, 8))
    assert orc.lib().oracle_surface_state(orc._v(q["u"], B2), orc._v(q["v"], B2), orc._v(m.z, B1), s.c(), m.geom,
                                          orc._v(ss, (-1, -1, 0))) == 0
    d0 = 10.0
    us = 0.41 * 6.0 / math.log((d0 + 0.05) / 0.05)
    assert ss[P.SS_D0, 0, 2, 3] == d0
    assert ss[P.SS_USTAR, 0, 2, 3] == pytest.approx(us, rel=1e-15)
    assert ss[P.SS_TAU0, 0, 2, 3] == pytest.approx(us * us, rel=1e-15)
    assert ss[P.SS_TAU1, 0, 2, 3] == 0.0
    assert ss[P.SS_KWALL, 0, 2, 3] == pytest.approx(us * us / 0.5562 ** 2, rel=1e-15)
    assert ss[P.SS_DRAG, 0, 2, 3] == pytest.approx(us * us / 6.0, rel=1e-15)
    # below the 0.1 m/s floor the speed is the floor
    q["u"][:] = 0.01
    orc.lib().oracle_surface_state(orc._v(q["u"], B2), orc._v(q["v"], B2), orc._v(m.z, B1), s.c(), m.geom,
                                   orc._v(ss, (-1, -1, 0)))
    assert ss[P.SS_SPEED, 0, 2, 3] == 0.1


def test_fill_rules_per_face_type():
    """Slip top: u, v, k even, w odd about its face value, theta by its
    gradient; surface-layer ground: w = 0 on flat ground, u by the wall
    law's gradient (the centred difference at the first centre equals
    u*/(kappa (d0 + z0))), k even (BoundaryConditions.cpp:38-65, 540-576)."""
    tile = PHY.flat_mesh(n=(6, 4, 10), top=200.0)
    m = PHY.omesh(tile)
    rng = np.random.default_rng(2)
    q = PHY.log_law_state(tile, ustar=0.4, z0=0.1)
    for n in ("v", "w", "tke", "p"):
        q[n] += rng.uniform(0.1, 0.2, size=q[n].shape)
    phys = P.Physics(bc=P.Boundary(zlo=P.SURFACE_LAYER, zhi=P.SLIP_WALL, zhi_theta_grad=0.003),
                     surface=P.SurfaceOptions(z0=0.1))
    ss = P.fill_ghosts(q, m, phys)
    nz = 10
    u, v, w, th, k = q["u"][0], q["v"][0], q["w"][0], q["theta"][0], q["tke"][0]
    K = lambda kk: kk + 2   # noqa: E731
    # top: even
    assert np.array_equal(u[K(nz)], u[K(nz - 1)]) and np.array_equal(u[K(nz + 1)], u[K(nz - 2)])
    assert np.array_equal(k[K(nz)], k[K(nz - 1)]) and np.array_equal(k[K(-1)], k[K(0)])
    # w: the face value 0 (flat), ghosts mirrored with the sign flipped
    assert np.all(w[K(0)] == 0) and np.all(w[K(nz)] == 0)
    assert np.array_equal(w[K(-1)], -w[K(1)]) and np.array_equal(w[K(nz + 2)], -w[K(nz - 2)])
    # theta at the top: (theta(nz) - theta(nz-2)) / (2 dz) = 0.003
    dz = 20.0
    np.testing.assert_allclose((th[K(nz)] - th[K(nz - 2)]) / (2 * dz), 0.003, rtol=1e-12)
    # u at the ground: the centred difference at the first centre is the wall
    # law's u*/(kappa (d0 + z0)) along the stress, the mean of the two columns
    us = ss[P.SS_USTAR, 0, 1:-1, 1:-1]
    t0, t1 = ss[P.SS_TAU0, 0, 1:-1, 1:-1], ss[P.SS_TAU1, 0, 1:-1, 1:-1]
    Gc = us / (0.41 * (10.0 + 0.1)) * t0 / np.hypot(t0, t1)
    G = 0.5 * (np.roll(Gc, 1, axis=1) + Gc)
    np.testing.assert_allclose((u[K(1), 2:-2, 2:-3] - u[K(-1), 2:-2, 2:-3]) / (2 * dz), G, rtol=1e-12)
    # periodic sides, ghosts included
    assert np.array_equal(u[:, :, 0], u[:, :, 6]) and np.array_equal(u[:, :, -1], u[:, :, 4])
    assert np.array_equal(th[:, 0, :], th[:, 4, :])


def _mesh_and_state_at_rest(theta_ref):
    tile = PHY.hill_mesh(n=(32, 24, 16), height=300.0, half=250.0)
    m = PHY.omesh(tile)
    q = PHY.zeros_state(m.geom)
    q["tke"][:] = 0.2
    phys = P.Physics(closure=None, surface=None, bc=None, theta_ref_z=theta_ref[0], theta_ref=theta_ref[1])
    phys.forcing.coriolis = True
    phys.forcing.latitude = 45.0
    t = P.theta_ref_field(m, phys)
    q["theta"][:] = 290.0
    q["theta"][0, 1:-1, 1:-1, 1:-1] = t[0]
    for name, val in (("nu", 10.0), ("gamma_theta", 3.0), ("gamma_tke", 3.0)):
        setattr(phys, name, np.full((1, 18, 26, 34), val))
    return tile, m, q, phys


def test_rest_over_steep_terrain_is_balanced():
    """test_Residual.cpp:274-302: at rest with theta = theta_ref(z) (a linear
    profile, 290 K + 3 K/km) momentum, continuity and k balance exactly."""
    zs = tuple(float(z) for z in np.linspace(-200.0, 2000.0, 12))
    ts = tuple(290.0 + 0.003 * z for z in zs)
    tile, m, q, phys = _mesh_and_state_at_rest((zs, ts))
    R, _, _ = P.residual(q, m, phys)
    for n in ("u", "v", "w", "p", "tke"):
        assert np.abs(R[n]).max() <= 1e-14, n


def test_sources_have_their_closed_forms():
    """test_Residual.cpp:566-608: uniform u, v, w, theta on a flat mesh:
    R_u = -(f (v - v_g) - f_cos w), R_v = f (u - u_g), R_w = -g (theta -
    theta_ref)/theta0 - f_cos u; with three_d off and the geostrophic wind
    itself, R = 0."""
    tile = PHY.flat_mesh(n=(16, 12, 8), top=400.0)
    m = PHY.omesh(tile)
    q = PHY.zeros_state(m.geom)
    q["u"][:], q["v"][:], q["w"][:], q["tke"][:], q["theta"][:] = 7.0, -2.0, 0.5, 0.3, 302.0
    phys = P.Physics(closure=None, surface=None, bc=None)
    for name in ("nu", "gamma_theta", "gamma_tke"):
        setattr(phys, name, np.zeros((1, 10, 14, 18)))
    phys.forcing = P.Forcing(coriolis=True, latitude=30.0, geostrophic=True, ug=5.0, vg=1.0)
    assert phys.forcing.coriolis_parameter() == pytest.approx(7.27221e-5, abs=1e-9)
    fc = 4 * math.pi / 86400.0
    f, fcos = fc * math.sin(math.pi / 6), fc * math.cos(math.pi / 6)
    R, _, _ = P.residual(q, m, phys)
    np.testing.assert_allclose(R["u"], -(f * (-2.0 - 1.0) - fcos * 0.5), rtol=0, atol=1e-15)
    np.testing.assert_allclose(R["v"], f * (7.0 - 5.0), rtol=0, atol=1e-15)
    np.testing.assert_allclose(R["w"], -9.81 * 2.0 / 300.0 - fcos * 7.0, rtol=0, atol=1e-13)
    phys.forcing.three_d = False
    q["u"][:], q["v"][:], q["w"][:], q["theta"][:] = 5.0, 1.0, 0.0, 300.0
    R, _, _ = P.residual(q, m, phys)
    assert max(np.abs(R[n]).max() for n in R) <= 1e-15


def test_wall_and_boundary_rows():
    """The surface layer's rows: R_w = w - 0 on the flat ground and top
    faces; R_k = a (k - k_w) on the first cells (dirichlet); the implicit
    rows: identity for w's face points, the ground neighbour removed and the
    Picard drag added for u, the wall-k row a alone, even folds at the top."""
    tile = PHY.flat_mesh(n=(6, 4, 10), top=200.0)
    m = PHY.omesh(tile)
    q = PHY.log_law_state(tile, ustar=0.4, z0=0.1)
    q["w"] += 0.01
    phys = P.Physics(bc=P.Boundary(), surface=P.SurfaceOptions(z0=0.1))
    ss = P.fill_ghosts(q, m, phys)
    R, cf, ss2 = P.residual(q, m, phys)
    assert np.array_equal(ss, ss2)
    assert np.all(R["w"][0, 0] == q["w"][0, 2, 2:-2, 2:-2]) and np.all(R["w"][0, -1] == q["w"][0, 12, 2:-2, 2:-2])
    a = np.maximum(ss[P.SS_USTAR, 0, 1:-1, 1:-1], 0.01) / ss[P.SS_D0, 0, 1:-1, 1:-1]
    np.testing.assert_array_equal(R["tke"][0, 0], a * (q["tke"][0, 2, 2:-2, 2:-2] - ss[P.SS_KWALL, 0, 1:-1, 1:-1]))
    Fx, Fy, Fz, Fb = orc.transport_fluxes(q["u"], q["v"], q["w"], B2, m.z, B1, m.ax, m.ay, B1, m.geom)
    F = [Fx[None], Fy[None], Fz[None]]
    A1 = orc.implicit_coefficients(1, F, Fb, cf["nu"], B1, m.z, B1, m.J, B1, m.geom)
    Au = P.implicit_rows(P.EQ_U, F, Fb, m, phys, cf, ss)
    drag = 0.5 * (ss[P.SS_DRAG, 0, 1:-1, :-1] + ss[P.SS_DRAG, 0, 1:-1, 1:])
    np.testing.assert_allclose(Au[0, 0], A1[0, 0] - A1[5, 0] + drag / 20.0 / 1.0, rtol=1e-14)
    assert np.all(Au[5, 0] == 0)
    np.testing.assert_array_equal(Au[0, -1], A1[0, -1] - A1[6, -1])   # even top: P -= ZP
    assert np.all(Au[6, -1] == 0)
    Aw = P.implicit_rows(P.EQ_W, F, Fb, m, phys, cf, ss)
    assert np.all(Aw[0, 0] == 1) and np.all(Aw[1:, 0] == 0) and np.all(Aw[0, -1] == 1)
    Ak = P.implicit_rows(P.EQ_TKE, F, Fb, m, phys, cf, ss)
    np.testing.assert_array_equal(Ak[0, 0], a)
    assert np.all(Ak[1:, 0] == 0)
