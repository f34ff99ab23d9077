"""A1's coefficients and apply on the GPU against the oracle: bit for bit
with the arithmetic face mean; with the logarithmic one to 1e-13 relative:
(a - b) / (ln a - ln b) just outside the series window (|a - b| / (a + b) =
0.01) amplifies an ulp of either log by about |ln a| / (2 * 0.01), and the
device log need not round as the host's does. Random wind and diffusivity
over the hill, all four locations."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import massflux_helpers as MH
from tests.test_oracle_implicit import _random_fields

pytestmark = pytest.mark.gpu
B = (-1, -1, -1)


def _fab(a, begin=B):
    import torch
    from ablsteadyrans_amd.fab import Fab
    return Fab(torch.from_numpy(np.ascontiguousarray(a)).cuda(), begin)


@pytest.mark.parametrize("loc", [0, 1, 2, 4])
@pytest.mark.parametrize("logz", [False, True])
def test_implicit_coefficients_and_apply_match_the_oracle(cuda, loc, logz):
    from ablsteadyrans_amd import implicit as IM
    from ablsteadyrans_amd.fab import Fab
    from ablsteadyrans_amd.poisson import DeviceMesh
    m = MH.hill_mesh(1, smooth_ghosts=False)
    F, g = _random_fields(m, 20.0, 50.0, 31 + loc)
    print(f"implicit loc={loc} logz={logz} {m.geom.n}", flush=True)
    ref = orc.implicit_coefficients(loc, F, B, g, B, m.z_nd, B, m.detJ, B, MH.ogeom(m), logz)
    A = Fab.at_loc(m.geom, loc, 7, 0)
    dm = DeviceMesh(m)
    IM.implicit_coefficients(loc, dm, [_fab(f) for f in F], _fab(g), A, face_mean=int(logz))
    got = A.numpy()
    if logz:
        assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)) <= 1e-13
    else:
        assert np.array_equal(got, ref)
    nd = [(loc >> d) & 1 for d in range(3)]
    nx, ny, nz = m.geom.n
    q = np.random.default_rng(7).uniform(-1, 2, size=(1, nz + nd[2] + 2, ny + nd[1] + 2, nx + nd[0] + 2))
    out = Fab.at_loc(m.geom, loc, 1, 0)
    IM.apply_implicit(loc, m.geom, _fab(ref, tuple(m.geom.lo)), _fab(q), out)
    assert np.array_equal(out.numpy()[0], orc.apply_implicit(loc, ref, q, B, MH.ogeom(m)))
