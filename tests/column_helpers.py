"""The reference's Neutral_Column deck as the oracle holds it (TEST
INFRASTRUCTURE), shared by the fixture script tests/golden/make_column_steady.py
and the GPU test that compares against its output.

Tests/Canonical_RANS/Neutral_Column/inputs_steady_column: 4 x 4 x 64 over
2560 x 2560 x 1000 m, periodic sides, surface layer z0 = 0.1 m, slip top,
uniform theta, the force F = 1.225e-4 m/s^2, the closure uncapped
(max_geom_lscale 1e12) with the logarithmic face mean, two sweeps; the first
guess the log law with u* = 0.25 m/s and its equilibrium k. The pseudo-time
step is local dtau with the previous iteration's nu at CFL min(50, 1.05^n)
(a fixed schedule in place of the SER law, so both sides take the same steps).
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as orc
from oracle import physics as P
from tests import physics_helpers as PHY

N_ITER = 2600
CFL_GROWTH = 1.05
CFL_MAX = 50.0
HEIGHT = 1000.0


def cfl_at(n: int) -> float:
    return min(CFL_MAX, CFL_GROWTH ** n)


def column_tile():
    from ablsteadyrans_amd import mesh as M
    geom = M.geometry((4, 4, 64), (0.0, 0.0, 0.0), (2560.0, 2560.0, 1000.0), (1, 1))
    return M.build_mesh(geom, (0.0, 0.0, 0.0), (2560.0, 2560.0, 1000.0), M.make_zlevels(64, 1000.0),
                        M.Terrain(kind="flat"))


def first_guess(tile) -> dict:
    q = PHY.zeros_state(PHY.P_geom(tile))
    _, _, zu = PHY.positions(tile, 1)
    q["u"][0] = 0.25 / 0.41 * np.log((np.maximum(zu, 0.0) + 0.1) / 0.1)
    q["theta"][:] = 300.0
    q["tke"][:] = 0.25 * 0.25 / (0.5562 * 0.5562)
    return q


def column_case():
    tile = column_tile()
    m = PHY.omesh(tile)
    phys = P.Physics(closure=P.ClosureOptions(max_geom_lscale=1.0e12),
                     bc=P.Boundary(zlo=P.SURFACE_LAYER, zhi=P.SLIP_WALL),
                     surface=P.SurfaceOptions(z0=0.1),
                     forcing=P.Forcing(pressure_gradient=(1.225e-4, 0.0, 0.0)),
                     log_face_mean=True)
    q = first_guess(tile)
    P.fill_ghosts(q, m, phys)
    return tile, m, q, phys


def norms(R, q, m) -> list:
    """The stage norms (Source/Simplec.cpp:385-409) with the default scales
    and the column's scale height."""
    res, _ = orc.stage_norms([R[n] for n in ("u", "v", "w", "p", "theta", "tke")], (0, 0, 0), q["u"], q["v"],
                             q["w"], (-2, -2, -2), q["tke"], (-2, -2, -2), 10.0, 10.0, 0.1, HEIGHT, m.geom)
    return res
