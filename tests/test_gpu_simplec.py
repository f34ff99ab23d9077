"""The HIP pressure correction (gust_pressure_correction) against the oracle
composite of Source/Simplec.cpp:134-220 (tests/simplec_helpers.py, an exact
solve): the corrected u, v, w, p agree to the Poisson tolerance, the
reported divergence before is the GPU divergence's max bit for bit, one
correction at reltol r leaves ||div|| <= r ||div*|| (test_Simplec.cpp:204),
the Dirichlet rows of w stay untouched on a flat mesh, and the refusals and
the net-source abort behave as the reference's."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import massflux_helpers as MH
from tests import simplec_helpers as S
from tests.test_oracle_simplec import flat_case

pytestmark = pytest.mark.gpu
N, D, P = orc.NEUMANN, orc.DIRICHLET, orc.PERIODIC
B = (-1, -1, -1)


def _fab(a, begin=B):
    import torch
    from ablsteadyrans_amd.fab import Fab
    return Fab(torch.from_numpy(np.ascontiguousarray(a)).cuda(), begin)


def _run(m, state, bc, alpha_p, reltol, source=None, project=False):
    from ablsteadyrans_amd import simplec as SP
    from ablsteadyrans_amd.poisson import DeviceMesh, PoissonOptions, PoissonSolver
    u, v, w, p, d = state
    dm = DeviceMesh(m)
    solver = PoissonSolver(dm, PoissonOptions(reltol=reltol))
    fu, fv, fw, fp = (_fab(a) for a in (u, v, w, p))
    lo = tuple(m.geom.lo)
    fd = [_fab(x[None], lo) for x in d]
    src = _fab(source[None], lo) if source is not None else None
    res = SP.pressure_correction(solver, dm, fu, fv, fw, fp, *fd, bc, alpha_p, source=src, project_net_source=project)
    return res, fu.numpy(), fv.numpy(), fw.numpy(), fp.numpy(), dm


@pytest.mark.parametrize("case", ["flat_singular", "hill_open_top"])
def test_pressure_correction_matches_the_exact_oracle_correction(cuda, case):
    if case == "flat_singular":
        m, bc, open_top = flat_case(), (P, P, N, P, P, N), False
    else:
        m, bc, open_top = MH.hill_mesh(0, smooth_ghosts=False), (P, P, N, P, P, D), True
    print(f"pressure_correction {case} {m.geom.n}", flush=True)
    state = S.random_state(m, 21, open_top=open_top)
    alpha_p = 0.7
    res, u, v, w, p, _ = _run(m, state, bc, alpha_p, 1e-12)
    ru, rv, rw, rp, phi, F, div0, div1 = S.pressure_correction(m, *state[:4], state[4], bc, alpha_p)
    for name, a, b in (("u", u, ru), ("v", v, rv), ("w", w, rw), ("p", p, rp)):
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        assert err <= 1e-9, (name, err)
    assert res.divergence_before == np.max(np.abs(div0))   # the GPU divergence is the oracle's, bitwise
    assert res.divergence_after <= 1e-9 * res.divergence_before
    assert abs(res.max_dp - alpha_p * np.max(np.abs(phi))) <= 1e-9 * res.max_dp
    assert res.poisson.singular == (case == "flat_singular")
    if case == "flat_singular":
        assert np.array_equal(w[0, 1], state[2][0, 1]) and np.array_equal(w[0, -2], state[2][0, -2])


@pytest.mark.parametrize("reltol", [0.1, 1e-10])
def test_one_correction_meets_its_tolerance(cuda, reltol):
    m = flat_case()
    state = S.random_state(m, 5)
    bc = (P, P, N, P, P, N)
    res, u, v, w, p, _ = _run(m, state, bc, 1.0, reltol)
    g = S.PH.ogeom(m)
    div0 = orc.velocity_divergence(*state[:3], B, m.z_nd, B, m.ax, m.ay, m.az, B, m.detJ, B, g)
    div1 = orc.velocity_divergence(u, v, w, B, m.z_nd, B, m.ax, m.ay, m.az, B, m.detJ, B, g)
    assert np.sqrt(np.sum(div1 ** 2)) <= reltol * np.sqrt(np.sum(div0 ** 2)) * (1 + 1e-6)
    assert res.divergence_after == np.max(np.abs(div1))
    assert res.max_dp > 0 and res.max_du > 0


def test_refusals_and_the_net_source_abort(cuda):
    import ablsteadyrans_amd as G
    m = flat_case()
    state = S.random_state(m, 3)
    bc = (P, P, N, P, P, N)
    with pytest.raises(G.GustError, match="alpha_p must be > 0 and <= 1"):
        _run(m, state, bc, 1.5, 1e-8)
    bad = [x.copy() for x in state[4]]
    bad[0][1, 1, 1] = -1.0
    with pytest.raises(G.GustError, match="the SIMPLEC denominator d_u must be finite and >= 0 everywhere"):
        _run(m, state[:4] + (bad,), bc, 1.0, 1e-8)
    src = np.full(tuple(reversed(m.geom.n)), 0.05)   # a net source the periodic / wall tile cannot carry
    with pytest.raises(G.GustError, match="net source the boundaries cannot carry"):
        _run(m, state, bc, 1.0, 1e-8, source=src)
    res, *_ = _run(m, state, bc, 1.0, 1e-8, source=src, project=True)
    assert res.poisson.singular and abs(res.poisson.removed_mean) > 0.01
