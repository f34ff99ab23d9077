"""The pressure-correction composite of the oracle (tests/simplec_helpers.py)
follows Source/Simplec.cpp:134-220: one correction with an exact solve
removes the divergence to round-off (test_Simplec.cpp:168-216 with reltol ->
0), leaves the Dirichlet rows of w on a flat mesh untouched, changes p by
alpha_p phi, and on a singular problem the removed J-weighted mean of the
divergence is round-off."""
import numpy as np

from ablsteadyrans_amd import mesh as M
from oracle import oracle as orc
from tests import massflux_helpers as MH
from tests import poisson_helpers as PH
from tests import simplec_helpers as S

N, D, P = orc.NEUMANN, orc.DIRICHLET, orc.PERIODIC


def flat_case():
    n = (8, 4, 6)
    lo, hi = (0.0, 0.0, 0.0), (1000.0, 800.0, 500.0)
    geom = M.geometry(n, lo, hi, (1, 1))
    return M.build_mesh(geom, lo, hi, M.make_zlevels(n[2], hi[2]), M.Terrain(kind="flat"))


def test_one_exact_correction_removes_the_divergence_on_a_flat_mesh():
    m = flat_case()
    u, v, w, p, d = S.random_state(m, 5)
    bc = (P, P, N, P, P, N)
    for alpha_p in (1.0, 0.7):
        u1, v1, w1, p1, phi, F, div0, div1 = S.pressure_correction(m, u, v, w, p, d, bc, alpha_p)
        assert np.max(np.abs(div0)) > 1e-3
        assert np.sqrt(np.sum(div1 ** 2)) <= 1e-11 * np.sqrt(np.sum(div0 ** 2))
        # the Dirichlet rows of w (d = 0) are untouched on a flat mesh
        assert np.array_equal(w1[0, 1], w[0, 1]) and np.array_equal(w1[0, -2], w[0, -2])
        assert np.array_equal(p1[0, 1:-1, 1:-1, 1:-1] - p[0, 1:-1, 1:-1, 1:-1], alpha_p * phi) or \
            np.max(np.abs(p1[0, 1:-1, 1:-1, 1:-1] - p[0, 1:-1, 1:-1, 1:-1] - alpha_p * phi)) <= 1e-14 * np.max(np.abs(phi))


def test_one_exact_correction_over_a_hill_with_an_open_top():
    m = MH.hill_mesh(0, smooth_ghosts=False)
    u, v, w, p, d = S.random_state(m, 7, open_top=True)
    bc = (P, P, N, P, P, D)
    u1, v1, w1, p1, phi, F, div0, div1 = S.pressure_correction(m, u, v, w, p, d, bc, 1.0)
    assert np.sqrt(np.sum(div1 ** 2)) <= 1e-10 * np.sqrt(np.sum(div0 ** 2))


def test_the_singular_problem_has_no_net_source():
    m = flat_case()
    u, v, w, p, d = S.random_state(m, 9)
    div0 = orc.velocity_divergence(u, v, w, S.B, m.z_nd, S.B, m.ax, m.ay, m.az, S.B, m.detJ, S.B, PH.ogeom(m))
    J = m.detJ[1:-1, 1:-1, 1:-1]
    mean = np.sum(J * div0) / np.sum(J)
    assert abs(mean) <= 1e-12 * np.max(np.abs(div0))
