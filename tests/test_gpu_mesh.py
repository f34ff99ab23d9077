"""The device mesh builder against the host builder (mesh.py, itself pinned
to the reference's mesh gold files): every array bitwise, over periodic and
non-periodic sides and each terrain kind; and its refusals."""
import os

import numpy as np
import pytest

from ablsteadyrans_amd import mesh as M
from tests.conftest import ROOT

pytestmark = pytest.mark.gpu


def _askervein_grid():
    d = np.load(os.path.join(ROOT, "tests", "golden", "askervein_grid.npz"))
    return M.ErfGrid(d["x"], d["y"], d["z"])


CASES = [
    ("flat", (6, 5, 7), (1, 1), M.Terrain(kind="flat"), 0.0, 1.08),
    ("agnesi", (16, 12, 10), (1, 1), M.Terrain(kind="agnesi", hill_height=100.0, hill_half_width=300.0, hill_shape="round"), 0.0, 0.0),
    ("hills", (14, 11, 9), (0, 1), M.Terrain(kind="hills", hills=[(300.0, 200.0, 60.0, 150.0),
                                                                (100.0, 500.0, 30.0, 90.0)]), 20.0, 1.1),
    ("hills_open", (9, 13, 8), (0, 0), M.Terrain(kind="hills", hills=[(200.0, 300.0, 80.0, 200.0)]), 5.0, 0.0),
    ("hills_xper", (12, 7, 6), (1, 0), M.Terrain(kind="hills", hills=[(550.0, 100.0, 50.0, 120.0)]), 0.0, 0.0),
]


def _frame_and_tile(n, periodic, terrain, dz0, ratio, lo=(0.0, 0.0, 0.0), hi=None):
    hi = hi or (n[0] * 50.0, n[1] * 50.0, 800.0)
    geom = M.geometry(n, lo, hi, periodic)
    zl = M.make_zlevels(n[2], hi[2], dz0=dz0, stretch_ratio=ratio) if (dz0 or ratio) else M.make_zlevels(n[2], hi[2])
    return M.tile_frame(geom, lo, hi, zl, terrain), M.build_mesh(geom, lo, hi, zl, terrain)


def _check(frame, tile):
    import torch
    from ablsteadyrans_amd.poisson import DeviceMesh
    dev = DeviceMesh(frame)
    torch.cuda.synchronize()
    for name, ref in (("z_nd", tile.z_nd), ("detJ", tile.detJ), ("ax", tile.ax), ("ay", tile.ay),
                      ("az", tile.az), ("z_cc", tile.z_cc)):
        got = getattr(dev, name).numpy()[0]
        assert got.shape == ref.shape, name
        assert np.array_equal(got, ref), (name, np.max(np.abs(got - ref)))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_device_mesh_is_bitwise_the_host_mesh(cuda, case):
    _, n, periodic, terrain, dz0, ratio = case
    frame, tile = _frame_and_tile(n, periodic, terrain, dz0, ratio)
    _check(frame, tile)


def test_device_mesh_of_the_askervein_grid(cuda):
    from ablsteadyrans_amd import decks as D
    tile = D.askervein_tile(_askervein_grid())
    frame = M.tile_frame(tile.geom, tile.prob_lo, tile.prob_hi, tile.zlevels, M.Terrain(kind="erf_grid",
                                                                                          grid=_askervein_grid()))
    _check(frame, tile)


def test_device_mesh_refuses_terrain_at_the_top(cuda):
    from ablsteadyrans_amd import GustError
    from ablsteadyrans_amd.poisson import DeviceMesh
    terrain = M.Terrain(kind="hills", hills=[(150.0, 150.0, 900.0, 100.0)])
    frame, _ = None, None
    geom = M.geometry((6, 6, 4), (0, 0, 0), (300.0, 300.0, 800.0), (1, 1))
    frame = M.tile_frame(geom, (0, 0, 0), (300.0, 300.0, 800.0), M.make_zlevels(4, 800.0), terrain)
    with pytest.raises(GustError, match="the terrain reaches .* m, at or above the top level 800.0 m"):
        DeviceMesh(frame)
    with pytest.raises(ValueError, match="the terrain reaches .* m, at or above the top level 800.0 m"):
        M.build_mesh(geom, (0, 0, 0), (300.0, 300.0, 800.0), M.make_zlevels(4, 800.0), terrain)
