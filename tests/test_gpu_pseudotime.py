"""local_dtau and face_dtau on the GPU equal the oracle bit for bit (FP64):
the manufactured flow over the hill with and without a diffusivity, and a
random positive dtau on periodic and walled tiles; ser_cfl equals the
oracle's law on a grid of arguments (host code behind the C-ABI)."""
import itertools

import numpy as np
import pytest

from oracle import oracle as orc
from tests import massflux_helpers as H

B = H.B1


def _fab(a, begin):
    import torch
    from ablsteadyrans_amd.fab import Fab
    return Fab(torch.from_numpy(np.ascontiguousarray(a)).cuda(), begin)


@pytest.mark.gpu
@pytest.mark.parametrize("with_nu", [False, True])
def test_local_dtau_matches_the_oracle_bitwise(cuda, with_nu):
    from ablsteadyrans_amd import pseudotime as PT
    from ablsteadyrans_amd.fab import Fab
    from ablsteadyrans_amd.poisson import DeviceMesh
    mesh = H.hill_mesh(1)
    vel = H.velocity(mesh, H.Flow())
    nx, ny, nz = mesh.geom.n
    nu = np.random.default_rng(9).uniform(0.1, 5.0, size=(1, nz, ny, nx)) if with_nu else None
    lo = tuple(mesh.geom.lo)
    ref = orc.local_dtau(*vel, B, mesh.z_nd, B, mesh.detJ, B, H.ogeom(mesh), 1.7, nu=nu, nu_begin=lo)
    d = Fab.at_loc(mesh.geom, 0, 1, 0)
    PT.local_dtau(DeviceMesh(mesh), *(_fab(a, B) for a in vel), 1.7, d, nu=_fab(nu, lo) if with_nu else None)
    assert np.array_equal(d.numpy()[0], ref)


@pytest.mark.gpu
@pytest.mark.parametrize("per", [(0, 0), (1, 1), (1, 0)])
def test_face_dtau_matches_the_oracle_bitwise(cuda, per):
    from ablsteadyrans_amd import pseudotime as PT
    from ablsteadyrans_amd.fab import Fab, Geometry
    geom = Geometry.of_size(13, 9, 7, periodic=per)
    d = np.random.default_rng(11).uniform(0.01, 3.0, size=(7, 9, 13))
    ref = orc.face_dtau(d, orc.geom_c(geom.lo, geom.hi, geom.periodic, geom.dx))
    F = [Fab.at_loc(geom, 1 << q, 1, 0) for q in range(3)]
    PT.face_dtau(geom, _fab(d[None], (0, 0, 0)), F)
    for q in range(3):
        assert np.array_equal(F[q].numpy()[0], ref[q]), q
    d[3, 4, 5] = 0.0
    with pytest.raises(Exception, match="gust: face_dtau: dtau_cc must be finite and > 0 everywhere"):
        PT.face_dtau(geom, _fab(d[None], (0, 0, 0)), F)


def test_ser_cfl_equals_the_oracle_law():
    from ablsteadyrans_amd import pseudotime as PT
    for cfl, r, p, g in itertools.product((0.5, 1.0, 7.0, 64.0), (0.0, 1e-300, 1e-4, 1 / 16, 0.5, 1.0, 3.0),
                                          (0.0, 0.5, 1.0, 2.0), (1.5, 4.0)):
        assert PT.ser_cfl(cfl, r, 1.0, 64.0, p, g) == orc.ser_cfl(cfl, r, 1.0, 64.0, p, g), (cfl, r, p, g)
