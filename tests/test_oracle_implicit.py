"""The implicit operator A1 of the oracle (oracle/implicit.c) pinned after
Tests/Unit_Tests/test_ImplicitOperator.cpp: diagonal dominance for random
wind and diffusivity over the hill (90-130); on a flat uniform mesh the
diffusion part is the closed-form second difference with the own-direction
factor (166-217), the convection part the upwind inflow weight F / (J dxi)
(143-164); A1 annihilates constants; the logarithmic face mean is the
closed form (a - b) / ln(a / b)."""
import numpy as np

from ablsteadyrans_amd import mesh as M
from oracle import oracle as orc
from tests import massflux_helpers as MH

B = (-1, -1, -1)
LOCS = (0, 1, 2, 4)


def _random_fields(mesh, fmax, numax, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = mesh.geom.n
    F = [rng.uniform(-fmax, fmax, size=(1, nz + 2 + (d == 2), ny + 2 + (d == 1), nx + 2 + (d == 0))) for d in range(3)]
    g = rng.uniform(0, numax, size=(1, nz + 2, ny + 2, nx + 2))
    g[rng.random(g.shape) < 0.2] = 0.0
    return F, g


def _A(mesh, loc, F, g, logz=False):
    return orc.implicit_coefficients(loc, F, B, g, B, mesh.z_nd, B, mesh.detJ, B, MH.ogeom(mesh), logz)


def test_diagonally_dominant_for_random_wind_and_nu():
    m = MH.hill_mesh(0, smooth_ghosts=False)
    F, g = _random_fields(m, 20.0, 50.0, 31)
    for loc in LOCS:
        for logz in (False, True):
            A = _A(m, loc, F, g, logz)
            nb = A[1:]
            assert np.all(nb >= 0), loc
            s = nb.sum(axis=0)
            assert np.all(A[0] > 0)
            assert np.max(np.abs(A[0] - s) / A[0]) <= 1e-14, loc


def _flat(n=(12, 10, 8), L=(600.0, 400.0, 160.0)):
    lo, hi = (0.0, 0.0, 0.0), L
    geom = M.geometry(n, lo, hi, (1, 1))
    return M.build_mesh(geom, lo, hi, M.make_zlevels(n[2], hi[2]), M.Terrain(kind="flat"))


def test_diffusion_part_is_the_flat_second_difference():
    m = _flat()
    nx, ny, nz = m.geom.n
    dx, dy, dz = m.geom.dx[0], m.geom.dx[1], m.zlevels[1]
    G = 3.0
    F = [np.zeros((1, nz + 2 + (d == 2), ny + 2 + (d == 1), nx + 2 + (d == 0))) for d in range(3)]
    g = np.full((1, nz + 2, ny + 2, nx + 2), G)
    for loc in LOCS:
        A = _A(m, loc, F, g)
        C = {0: -1, 1: 0, 2: 1, 4: 2}[loc]
        nd = [(loc >> d) & 1 for d in range(3)]
        shape = (1, nz + nd[2] + 2, ny + nd[1] + 2, nx + nd[0] + 2)
        K, Jj, I = np.meshgrid(*(np.arange(s) - 1.0 for s in shape[1:]), indexing="ij")
        for d, X, h in ((0, I * dx, dx), (1, Jj * dy, dy), (2, K * dz, dz)):
            q = (X ** 2)[None].copy()
            y = orc.apply_implicit(loc, A, q, B, MH.ogeom(m))
            own = 2.0 if C == d else 1.0
            want = -2.0 * G * own
            if loc == 4:   # the ground row of w: a half control volume (J = 0 below), replaced by the BCs
                y = y[1:]
            assert np.max(np.abs(y - want)) <= 1e-9 * abs(want), (loc, d)
        # and A1 of a constant is zero
        y = orc.apply_implicit(loc, A, np.ones(shape), B, MH.ogeom(m))
        assert np.max(np.abs(y)) <= 1e-12 * np.max(A[0])


def test_convection_part_is_the_upwind_inflow_weight():
    m = _flat()
    nx, ny, nz = m.geom.n
    dx, dy, dz = m.geom.dx
    J = m.detJ[1:-1, 1:-1, 1:-1]
    rng = np.random.default_rng(4)
    F = [rng.uniform(-5, 5, size=(1, nz + 2 + (d == 2), ny + 2 + (d == 1), nx + 2 + (d == 0))) for d in range(3)]
    g = np.zeros((1, nz + 2, ny + 2, nx + 2))
    A = _A(m, 0, F, g)
    fx = F[0][0, 1:-1, 1:-1, 1:-1]
    fz = F[2][0, 1:-1, 1:-1, 1:-1]
    assert np.allclose(A[1], np.maximum(fx[:, :, :-1], 0) / (J * dx), rtol=1e-15, atol=0)
    assert np.allclose(A[2], np.maximum(-fx[:, :, 1:], 0) / (J * dx), rtol=1e-15, atol=0)
    assert np.allclose(A[5], np.maximum(fz[:-1], 0) / (J * m.geom.dx[2]), rtol=1e-15, atol=0)
    assert np.allclose(A[6], np.maximum(-fz[1:], 0) / (J * m.geom.dx[2]), rtol=1e-15, atol=0)


def test_logarithmic_face_mean():
    m = _flat((6, 5, 4))
    nx, ny, nz = m.geom.n
    F = [np.zeros((1, nz + 2 + (d == 2), ny + 2 + (d == 1), nx + 2 + (d == 0))) for d in range(3)]
    g = np.ones((1, nz + 2, ny + 2, nx + 2))
    g[0, 2] = 4.0          # level k = 1: the z-face k = 1 sees (1, 4), k = 2 sees (4, 1)
    g[0, 3] = 1.0005       # inside the series window with level 2
    A = _A(m, 0, F, g, logz=True)
    Aa = _A(m, 0, F, g, logz=False)
    dz = m.geom.dx[2]
    J = m.detJ[1:-1, 1:-1, 1:-1]
    lm = 3.0 / np.log(4.0)
    # z-face k=1 is cell 0's high face and cell 1's low face: dif / (zz J dz^2)
    zz = J[0]
    assert np.allclose(A[6][0], lm / (zz * dz) / dz / J[0], rtol=1e-15)
    a, b = 4.0, 1.0005
    lm2 = (a - b) / np.log(a / b)
    assert np.allclose(A[6][1], lm2 / (J[1] * dz) / dz / J[1], rtol=1e-14)
    # horizontal faces keep the arithmetic mean
    assert np.array_equal(A[1], Aa[1]) and np.array_equal(A[3], Aa[3])
