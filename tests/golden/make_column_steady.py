"""Writes tests/golden/column_steady.npz: the oracle's SIMPLEC composite of
the reference's Neutral_Column deck (Tests/Canonical_RANS/Neutral_Column/
inputs_steady_column: 4 x 4 x 64 over 2560 x 2560 x 1000 m, periodic sides,
surface layer z0 = 0.1 m, slip top, uniform theta, F = 1.225e-4 m/s^2, the
uncapped closure with the logarithmic face mean, two sweeps) iterated to
steady from the deck's first guess (the log law with u* = 0.25 m/s and its
equilibrium k), the pressure correction solved exactly, local dtau with the
previous iteration's nu at the fixed CFL schedule min(50, 1.05^n).

Stored: the fields after N_ITER iterations (two-ghost arrays, as the oracle
holds them), the stage norms of every 10th residual and of the last, and the
schedule's parameters. The GPU test runs the device stage on the same
schedule and compares the converged fields (tests/test_gpu_column_steady.py).

Run from the repository root: python tests/golden/make_column_steady.py
(about two minutes on one core).
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from oracle import physics as P  # noqa: E402
from tests import column_helpers as CO  # noqa: E402
from tests import stage_helpers as STH  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "column_steady.npz")


def main():
    tile, m, q, phys = CO.column_case()
    g = m.geom
    B1 = (-1, -1, -1)
    nu = None
    norms, its = [], []
    t0 = time.time()
    for n in range(CO.N_ITER):
        cfl = CO.cfl_at(n)
        q1 = {k: a[:, 1:-1, 1:-1, 1:-1].copy() for k, a in q.items()}
        if nu is None:
            dtau = orc.local_dtau(q1["u"], q1["v"], q1["w"], B1, m.z, B1, m.J, B1, g, cfl)
        else:
            dtau = orc.local_dtau(q1["u"], q1["v"], q1["w"], B1, m.z, B1, m.J, B1, g, cfl, nu=nu, nu_begin=B1)
        R, cf = STH.simplec_iteration(tile, m, q, phys, dtau, sweeps=2, tke_floor=0.0)
        nu = np.ascontiguousarray(cf["nu"])
        if n % 10 == 0 or n == CO.N_ITER - 1:
            norms.append(CO.norms(R, q, m))
            its.append(n)
        if n % 200 == 0:
            print(f"iteration {n}: cfl {cfl:.3g} norms {norms[-1]} ({time.time() - t0:.0f} s)", flush=True)
    np.savez_compressed(OUT, **{f"q_{k}": v for k, v in q.items()}, norms=np.array(norms),
                        norm_iterations=np.array(its), n_iter=CO.N_ITER, cfl_growth=CO.CFL_GROWTH,
                        cfl_max=CO.CFL_MAX)
    print(f"wrote {OUT}: {CO.N_ITER} iterations in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
