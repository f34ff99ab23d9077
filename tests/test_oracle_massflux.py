"""The mass-flux oracle (oracle/massflux.c) pinned without the reference
binary, after Tests/Unit_Tests/test_MassFlux.cpp:
  * second-order convergence of D(U) and of the three components of G p
    over a hill (test_MassFlux.cpp:86-108);
  * G is minus the adjoint of D under the control-volume inner product
    (197-225);
  * the horizontal components of G ignore the ghost cells below the ground
    (110-134);
  * D(U) is, bit for bit, the Poisson layer's face_divergence of
    (u, v, Omega) (227-241) -- the Poisson layer's face_divergence is itself
    pinned by the Askervein gold;
  * Omega and w round trip (273-294); Omega = w exactly on a flat mesh;
    transport fluxes are a_x u, a_y v and Omega on the grown faces.
"""
import numpy as np

from oracle import oracle as orc
from tests import massflux_helpers as H

B = H.B1


def _div(mesh, u, v, w):
    return orc.velocity_divergence(u, v, w, B, mesh.z_nd, B, mesh.ax, mesh.ay, mesh.az, B, mesh.detJ, B,
                                   H.ogeom(mesh))


def _grad(mesh, p):
    return orc.pressure_gradient(p, B, mesh.z_nd, B, mesh.detJ, B, mesh.ax, mesh.ay, B, H.ogeom(mesh))


def test_divergence_and_gradient_are_second_order_over_a_hill():
    fl = H.Flow()
    errs = []
    for m in range(3):
        mesh = H.hill_mesh(m)
        u, v, w = H.velocity(mesh, fl)
        div = _div(mesh, u, v, w)
        dex = fl.div(*H.positions(mesh, 0, 0))
        G = _grad(mesh, H.pressure(mesh, fl))
        e = {"div": np.sqrt(np.mean((div - dex) ** 2))}
        for d in range(3):
            gex = fl.grad(d, *H.positions(mesh, 1 << d, 0))
            err = G[d] - gex
            if d == 2:   # the ground and top rows of G_z are boundary rows
                err = err[1:-1]
            e[f"g{d}"] = np.sqrt(np.mean(err ** 2))
        errs.append(e)
    for key in errs[0]:
        o1 = np.log2(errs[0][key] / errs[1][key])
        o2 = np.log2(errs[1][key] / errs[2][key])
        assert o1 >= 1.9 and o2 >= 1.9, (key, [e[key] for e in errs], o1, o2)


def test_gradient_is_minus_the_adjoint_of_the_divergence():
    mesh = H.hill_mesh(1)
    vel = [H.random_field(mesh, 1 << d, 10 + d, 2) for d in range(3)]
    p = H.random_field(mesh, 0, 20, 0)
    div = _div(mesh, *vel)
    G = _grad(mesh, p)
    gu, scale = 0.0, 0.0
    for d in range(3):
        vv = vel[d][0, 1:-1, 1:-1, 1:-1]
        t = H.weighted_dot(mesh, G[d], vv, 1 << d)
        gu += t
        scale += abs(t)
    pd = H.weighted_dot(mesh, p[0, 1:-1, 1:-1, 1:-1], div, 0)
    assert scale > 1.0
    assert abs(gu + pd) <= 1e-12 * scale, (gu, pd)


def test_gradient_ignores_the_ghost_below_the_ground():
    mesh = H.hill_mesh(1)
    p = H.pressure(mesh, H.Flow())
    G1 = _grad(mesh, p)
    p2 = p.copy()
    p2[0, 0] = 1.0e3   # k = -1
    G2 = _grad(mesh, p2)
    assert np.array_equal(G1[0], G2[0]) and np.array_equal(G1[1], G2[1])
    assert not np.array_equal(G1[2][0], G2[2][0])   # the ground row of G_z does read it


def test_divergence_is_the_poisson_layers_face_divergence_of_omega():
    mesh = H.hill_mesh(1)
    u, v, w = H.velocity(mesh, H.Flow())
    om, ob = orc.omega_from_w(u, v, w, B, mesh.z_nd, B, H.ogeom(mesh), 0)
    # face_divergence with per-array begins: u, v at (-1,-1,-1), Omega at lo
    n = mesh.geom.n
    F = [u, v, np.zeros_like(w)]
    F[2][0, 1:n[2] + 2, 1:-1, 1:-1] = om
    d1 = orc.face_divergence(F, B, (mesh.ax, mesh.ay, mesh.az), B, mesh.detJ, B, H.ogeom(mesh))
    d2 = _div(mesh, u, v, w)
    assert np.array_equal(d1, d2)
    assert np.max(np.abs(d1)) > 1e-6


def test_omega_and_w_round_trip_and_omega_is_w_on_a_flat_mesh():
    mesh = H.hill_mesh(1)
    u, v, w = H.velocity(mesh, H.Flow())
    om, ob = orc.omega_from_w(u, v, w, B, mesh.z_nd, B, H.ogeom(mesh), 0)
    w2 = orc.w_from_omega(u, v, om[None].copy(), B, ob, mesh.z_nd, B, H.ogeom(mesh))
    wv = w[0, 1:-1, 1:-1, 1:-1]
    assert np.max(np.abs(w2 - wv)) <= 1e-14 * np.max(np.abs(wv))
    assert np.max(np.abs(om - wv)) > 1e-3   # the flux through the sloped faces differs from w
    flat = H.hill_mesh(1, flat=True)
    u, v, w = H.velocity(flat, H.Flow())
    om, _ = orc.omega_from_w(u, v, w, B, flat.z_nd, B, H.ogeom(flat), 0)
    assert np.array_equal(om, w[0, 1:-1, 1:-1, 1:-1])


def test_transport_fluxes_are_the_area_weighted_velocities_and_omega():
    mesh = H.hill_mesh(0)
    nx, ny, nz = mesh.geom.n
    rng = np.random.default_rng(5)
    vel = [rng.uniform(-1, 1, size=(1, nz + 4 + (d == 2), ny + 4 + (d == 1), nx + 4 + (d == 0))) for d in range(3)]
    B2 = (-2, -2, -2)
    # the mesh areas have one ghost layer: pad to two
    ax = np.pad(mesh.ax, 1)[None].copy()
    ay = np.pad(mesh.ay, 1)[None].copy()
    z = np.pad(mesh.z_nd, 1, mode="edge")[None].copy()
    Fx, Fy, Fz, b = orc.transport_fluxes(vel[0], vel[1], vel[2], B2, z, B2, ax, ay, B2, H.ogeom(mesh))
    # valid faces + one ghost face: offsets of 1 into the 2-ghost arrays
    assert np.array_equal(Fx, ax[0, 1:-1, 1:-1, 1:-1] * vel[0][0, 1:-1, 1:-1, 1:-1])
    assert np.array_equal(Fy, ay[0, 1:-1, 1:-1, 1:-1] * vel[1][0, 1:-1, 1:-1, 1:-1])
    om, _ = orc.omega_from_w(vel[0], vel[1], vel[2], B2, z, B2, H.ogeom(mesh), 1)
    assert np.array_equal(Fz, om)
