"""The pseudo-time oracle (oracle/pseudotime.c) against the analytic limits of
Tests/Unit_Tests/test_Driver.cpp:167-255 (local_dtau), the SER law and its
caps (257-358), and the face means of face_dtau (Source/Predictor.cpp:30-77)."""
import math

import numpy as np
import pytest

from ablsteadyrans_amd import mesh as M
from oracle import oracle as orc
from tests import massflux_helpers as H

B = (-1, -1, -1)


def _flat(n=(16, 8, 8)):
    lo, hi = (0.0, 0.0, 0.0), (800.0, 400.0, 200.0)
    geom = M.geometry(n, lo, hi, (1, 1))
    return M.build_mesh(geom, lo, hi, M.make_zlevels(n[2], hi[2]), M.Terrain(kind="flat"))


def _const_vel(mesh, U, V, W):
    nx, ny, nz = mesh.geom.n
    return [np.full((1, nz + 2 + (d == 2), ny + 2 + (d == 1), nx + 2 + (d == 0)), val)
            for d, val in enumerate((U, V, W))]


def _dtau(mesh, vel, cfl, nu=None):
    nub = tuple(mesh.geom.lo) if nu is not None else None
    return orc.local_dtau(*vel, B, mesh.z_nd, B, mesh.detJ, B, H.ogeom(mesh), cfl, nu=nu, nu_begin=nub)


def test_local_dtau_matches_the_analytic_limits():
    m = _flat()
    dx = m.geom.dx
    U, V, W, NU, CFL = 7.0, -3.0, 0.5, 2.0, 1.5
    vel = _const_vel(m, U, V, W)
    nu = np.full((1,) + tuple(reversed(m.geom.n)), NU)
    want = CFL / (abs(U) / dx[0] + abs(V) / dx[1] + abs(W) / dx[2] + 2 * NU * (1 / dx[0] ** 2 + 1 / dx[1] ** 2))
    d = _dtau(m, vel, CFL, nu)
    assert np.max(np.abs(d - want)) <= 1e-14 * want
    want0 = CFL / (abs(U) / dx[0] + abs(V) / dx[1] + abs(W) / dx[2])
    assert np.max(np.abs(_dtau(m, vel, CFL) - want0)) <= 1e-14 * want0
    # at rest without diffusion: the floor, exactly
    assert np.all(_dtau(m, _const_vel(m, 0.0, 0.0, 0.0), CFL) == CFL / orc.K_MIN_RATE)
    # u varying in x: the mean of the cell's two faces
    vel = _const_vel(m, 0.0, 0.0, 0.0)
    vel[0][0] = np.arange(-1, m.geom.n[0] + 2)[None, None, :]
    d = _dtau(m, vel, CFL)
    uc = np.arange(m.geom.n[0]) + 0.5
    want = np.broadcast_to(CFL / (uc / dx[0]), d.shape)
    assert np.max(np.abs(d - want) / want) <= 1e-14


def test_local_dtau_over_a_hill_sees_only_the_flux_through_the_layers():
    h = H.hill_mesh(0, smooth_ghosts=False)
    dx = h.geom.dx
    U, V, W, CFL = 7.0, -3.0, 0.5, 1.5
    vel = _const_vel(h, U, V, 0.0)
    nx, ny, nz = h.geom.n
    om0 = np.zeros((1, nz + 1, ny, nx))
    w = orc.w_from_omega(vel[0], vel[1], om0, B, tuple(h.geom.lo), h.z_nd, B, H.ogeom(h))
    assert np.max(np.abs(w)) > 0.1
    vel[2][0, 1:-1, 1:-1, 1:-1] = w
    # the ghost faces of w are read only through the slopes' u/v means; fill them from their own columns
    vel[2] = H.periodic_fill(vel[2], 1, (0, 0, 1))
    d = _dtau(h, vel, CFL)
    want = CFL / (abs(U) / dx[0] + abs(V) / dx[1])
    assert np.max(np.abs(d - want)) <= 1e-12 * want
    # a plain vertical velocity: dtau = CFL J dzeta / |W| cell by cell
    d = _dtau(h, _const_vel(h, 0.0, 0.0, W), CFL)
    J = h.detJ[1:-1, 1:-1, 1:-1]
    want = CFL * J * dx[2] / W
    assert np.max(np.abs(d - want) / want) <= 1e-12


def test_local_dtau_refuses_a_bad_cfl():
    m = _flat((4, 4, 4))
    vel = _const_vel(m, 1.0, 1.0, 1.0)
    for cfl in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(orc.OracleError):
            _dtau(m, vel, cfl)


def test_ser_law_and_its_caps():
    # the ramp of test_Driver.cpp:273-287 (cfl0 1, cfl_max 64, p 1, growth 4)
    assert orc.ser_cfl(1, 1 / 16, 1, 64, 1, 4) == 4
    assert orc.ser_cfl(4, 1 / 64, 1, 64, 1, 4) == 16
    assert orc.ser_cfl(16, 1 / 256, 1, 64, 1, 4) == 64
    assert orc.ser_cfl(64, 1 / 1024, 1, 64, 1, 4) == 64
    # p = 2, loose caps: the law itself (256 = 1 / (1/16)^2)
    assert orc.ser_cfl(1, 1 / 16, 1, 1e6, 2, 1000) == 256
    assert orc.ser_cfl(256, 1 / 64, 1, 1e6, 2, 1000) == 4096
    # growth cap 1.5 binding
    assert orc.ser_cfl(1, 1 / 16, 1, 64, 1, 1.5) == 1.5
    assert orc.ser_cfl(1.5, 1 / 64, 1, 64, 1, 1.5) == 2.25
    # p = 0: cfl0
    assert orc.ser_cfl(1, 1e-6, 1, 64, 0, 4) == 1
    # a vanished residual and a tiny one with p = 2: the cap through the growth, no overflow
    assert orc.ser_cfl(4, 0.0, 1, 64, 1, 4) == 16
    assert orc.ser_cfl(4, 2.2250738585072014e-308, 1, 64, 2, 4) == 16
    # r above r_min: cfl0 r^-p below the caps
    r = 0.5
    assert orc.ser_cfl(100, r, 1, 64, 1.5, 4) == 1 * math.pow(r, -1.5)


def test_face_dtau_means_and_boundary_faces():
    rng = np.random.default_rng(3)
    for per in ((0, 0), (1, 0), (0, 1), (1, 1)):
        g = orc.geom_c((0, 0, 0), (4, 3, 2), per + (0,))
        d = rng.uniform(0.5, 2.0, size=(3, 4, 5))
        fx, fy, fz = orc.face_dtau(d, g)
        # interior means
        assert np.array_equal(fx[:, :, 1:-1], 0.5 * (d[:, :, :-1] + d[:, :, 1:]))
        assert np.array_equal(fy[:, 1:-1, :], 0.5 * (d[:, :-1, :] + d[:, 1:, :]))
        assert np.array_equal(fz[1:-1], 0.5 * (d[:-1] + d[1:]))
        assert np.array_equal(fz[0], d[0]) and np.array_equal(fz[-1], d[-1])
        if per[0]:
            wrap = 0.5 * (d[:, :, -1] + d[:, :, 0])
            assert np.array_equal(fx[:, :, 0], wrap) and np.array_equal(fx[:, :, -1], wrap)
        else:
            assert np.array_equal(fx[:, :, 0], d[:, :, 0]) and np.array_equal(fx[:, :, -1], d[:, :, -1])
        if per[1]:
            wrap = 0.5 * (d[:, -1, :] + d[:, 0, :])
            assert np.array_equal(fy[:, 0, :], wrap) and np.array_equal(fy[:, -1, :], wrap)
        else:
            assert np.array_equal(fy[:, 0, :], d[:, 0, :]) and np.array_equal(fy[:, -1, :], d[:, -1, :])
    g = orc.geom_c((0, 0, 0), (4, 3, 2), (0, 0, 0))
    for bad in (0.0, -1.0, float("nan"), float("inf")):
        d = np.ones((3, 4, 5))
        d[1, 2, 3] = bad
        with pytest.raises(orc.OracleError):
            orc.face_dtau(d, g)
