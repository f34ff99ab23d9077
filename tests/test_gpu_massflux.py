"""The HIP mass-flux operators equal the oracle bit for bit (FP64, same
expression order) on the hill mesh of test_MassFlux.cpp with the manufactured
flow and with random fields: omega_from_w (0 and 1 ghost), w_from_omega,
transport_fluxes, velocity_divergence and pressure_gradient."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import massflux_helpers as H

pytestmark = pytest.mark.gpu
B = H.B1


def _fab(a, begin):
    import torch
    from ablsteadyrans_amd.fab import Fab
    return Fab(torch.from_numpy(np.ascontiguousarray(a)).cuda(), begin)


def _case(kind):
    from ablsteadyrans_amd.poisson import DeviceMesh
    mesh = H.hill_mesh(1)
    if kind == "flow":
        vel = H.velocity(mesh, H.Flow(), grow=1)
        p = H.pressure(mesh, H.Flow())
    else:
        vel = [H.random_field(mesh, 1 << d, 30 + d, 0) for d in range(3)]
        p = H.random_field(mesh, 0, 40, 0)
        p[0, 0] = np.random.default_rng(41).uniform(-1, 1, size=p[0, 0].shape)   # a non-zero ground ghost
    return mesh, DeviceMesh(mesh), vel, p


@pytest.mark.parametrize("kind", ["flow", "random"])
def test_divergence_and_gradient_match_the_oracle_bitwise(cuda, kind):
    import torch
    from ablsteadyrans_amd import massflux as MF
    from ablsteadyrans_amd.fab import Fab
    mesh, dm, vel, p = _case(kind)
    print(f"massflux {kind} {mesh.geom.n}", flush=True)
    g = H.ogeom(mesh)
    u, v, w = (_fab(a, B) for a in vel)
    div = Fab.at_loc(mesh.geom, 0, 1, 0)
    MF.velocity_divergence(dm, u, v, w, div)
    ref = orc.velocity_divergence(*vel, B, mesh.z_nd, B, mesh.ax, mesh.ay, mesh.az, B, mesh.detJ, B, g)
    torch.cuda.synchronize()
    assert np.array_equal(div.numpy()[0], ref)
    Gp = [Fab.at_loc(mesh.geom, 1 << d, 1, 0) for d in range(3)]
    MF.pressure_gradient(dm, _fab(p, B), Gp)
    gref = orc.pressure_gradient(p, B, mesh.z_nd, B, mesh.detJ, B, mesh.ax, mesh.ay, B, g)
    for d in range(3):
        assert np.array_equal(Gp[d].numpy()[0], gref[d]), d


@pytest.mark.parametrize("ngrow", [0, 1])
def test_omega_and_w_match_the_oracle_bitwise(cuda, ngrow):
    from ablsteadyrans_amd import massflux as MF
    from ablsteadyrans_amd.fab import Fab
    mesh, dm, vel, _ = _case("flow")
    g = H.ogeom(mesh)
    if ngrow == 1:   # u, v need two ghosts: pad the one-ghost fields with their edge values
        vel2 = [np.pad(a, ((0, 0), (1, 1), (1, 1), (1, 1)), mode="edge") for a in vel]
        vb = (-2, -2, -2)
    else:
        vel2, vb = vel, B
    u, v, w = (_fab(a, vb) for a in vel2)
    om = Fab.at_loc(mesh.geom, 4, 1, ngrow)
    MF.omega_from_w(dm, u, v, w, om, ngrow=ngrow)
    ref, rb = orc.omega_from_w(*vel2, vb, mesh.z_nd, B, g, ngrow)
    assert np.array_equal(om.numpy()[0], ref)
    if ngrow == 0:
        w2 = Fab.at_loc(mesh.geom, 4, 1, 0)
        MF.w_from_omega(dm, u, v, om, w2)
        wref = orc.w_from_omega(vel[0], vel[1], ref[None].copy(), B, rb, mesh.z_nd, B, g)
        assert np.array_equal(w2.numpy()[0], wref)


def test_transport_fluxes_match_the_oracle_bitwise(cuda):
    from ablsteadyrans_amd import massflux as MF
    from ablsteadyrans_amd.fab import Fab
    mesh, dm, vel, _ = _case("flow")
    g = H.ogeom(mesh)
    vel2 = [np.pad(a, ((0, 0), (1, 1), (1, 1), (1, 1)), mode="edge") for a in vel]
    vb = (-2, -2, -2)
    F = [Fab.at_loc(mesh.geom, 1 << d, 1, 1) for d in range(3)]
    MF.transport_fluxes(dm, *(_fab(a, vb) for a in vel2), F)
    ref = orc.transport_fluxes(*vel2, vb, mesh.z_nd, B, mesh.ax, mesh.ay, B, g)
    for d in range(3):
        assert np.array_equal(F[d].numpy()[0], ref[d]), d
