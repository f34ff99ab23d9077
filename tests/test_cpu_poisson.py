"""The CPU Poisson solve the oracle composite uses on tiles too large to
factor (tests/cpu_poisson.py): the z-line coefficients are the assembled
oracle operator's tridiagonal entries exactly, the batched Thomas solve
inverts them, and GMRES reaches its tolerance and a sparse direct solve's
answer (periodic and open sides, with face coefficients, a Dirichlet face
as the pressure correction of the open-channel tiles has)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from ablsteadyrans_amd import mesh as M
from oracle import oracle as orc
from tests import cpu_poisson as CP
from tests import poisson_helpers as PH


def _tile(periodic, n=(11, 8, 7)):
    lo, hi = (0.0, 0.0, 0.0), (1100.0, 800.0, 700.0)
    geom = M.geometry(n, lo, hi, periodic)
    zl = M.make_zlevels(n[2], 700.0, dz0=20.0, stretch_ratio=1.2)
    return M.build_mesh(geom, lo, hi, zl, M.Terrain(kind="hills", hills=[(550.0, 400.0, 120.0, 250.0)]))


@pytest.mark.parametrize("periodic,bc", [((1, 0), (orc.PERIODIC, orc.NEUMANN, orc.NEUMANN, orc.PERIODIC,
                                                     orc.NEUMANN, orc.DIRICHLET)),
                                          ((0, 0), (orc.NEUMANN, orc.NEUMANN, orc.NEUMANN, orc.DIRICHLET,
                                                    orc.NEUMANN, orc.NEUMANN))])
def test_line_coefficients_and_gmres(periodic, bc):
    t = _tile(periodic)
    nx, ny, nz = t.geom.n
    rng = np.random.default_rng(4)
    coeffs = tuple(rng.uniform(0.5, 2.0, size=a.shape) for a in (t.ax, t.ay, t.az))
    A = PH.sparse_operator(t, bc, coeffs=coeffs).tocsr()
    lo, di, up = CP.zline_coefficients(t, bc, coeffs)
    idx = np.arange(nx * ny * nz).reshape(nz, ny, nx)
    assert np.array_equal(di.ravel(), A.diagonal())
    assert np.array_equal(lo[1:].ravel(), np.asarray(A[idx[1:].ravel(), idx[:-1].ravel()]).ravel())
    assert np.array_equal(up[:-1].ravel(), np.asarray(A[idx[:-1].ravel(), idx[1:].ravel()]).ravel())
    r = rng.uniform(-1, 1, size=(nz, ny, nx))
    x = CP.thomas(lo, di, up, r)
    T = np.zeros_like(r)
    T += di * x
    T[1:] += lo[1:] * x[:-1]
    T[:-1] += up[:-1] * x[1:]
    assert np.max(np.abs(T - r)) <= 1e-12 * np.max(np.abs(r))
    b = rng.uniform(-1, 1, size=(nz, ny, nx))
    s = CP.LineGMRES(t, rtol=1e-10, maxiter=50)
    phi = s(b, coeffs, bc)
    res = PH.apply(t, phi, bc, coeffs) - b
    assert np.linalg.norm(res) <= 1e-9 * np.linalg.norm(b)
    ex = spla.spsolve(A.tocsc(), b.ravel()).reshape(b.shape)
    assert np.max(np.abs(phi - ex)) <= 1e-8 * np.max(np.abs(ex))
