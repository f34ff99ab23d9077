"""The reference's Neutral_Column deck converged on the GPU against the
oracle's composite converged the same way (tests/golden/column_steady.npz,
written by tests/golden/make_column_steady.py): the same first guess, the
same pseudo-time schedule (local dtau with the previous iteration's nu at
CFL min(50, 1.05^n)), 2600 iterations each, the device's pressure
correction at reltol 1e-12 and the composite's exact.

Checks: the composite's residuals fell six orders (the deck's criterion);
every converged field on the device within 1e-8 relative L2 of the
composite's; and the deck's expected results
(Tests/Canonical_RANS/Neutral_Column/README.md:20-31) on the device's
converged state: u* = sqrt(F H) to 1e-4, the first cell the wall law's to
1e-6, the second cell within 1 % of the log law and the lowest eight within
2 %, the wall k = u*^2 / Cmu0^2 to 1e-6, nu_t of the second cell within 6 %
of kappa u* (z + z0) (the oracle's closure of the device's state), every
column identical to 1e-10.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

from oracle import physics as P
from tests import column_helpers as CO
from tests import physics_helpers as PHY

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "column_steady.npz")
FIELDS = ("u", "v", "w", "p", "theta", "tke")


def _device_run():
    import torch
    from ablsteadyrans_amd import _lib
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd.fab import Fab
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile = CO.column_tile()
    dm = DeviceMesh(tile)
    inputs = GP.ResidualInputs(closure=GP.ClosureOptions(max_geom_lscale=1.0e12),
                               surface=GP.SurfaceLayerOptions(z0=0.1),
                               forcing=GP.ForcingOptions(pressure_gradient=(1.225e-4, 0.0, 0.0)),
                               log_face_mean=True)
    stage = GP.SimplecStage(dm, inputs, GP.SimplecOptions(sweeps=2, height=CO.HEIGHT, pressure_reltol=1e-12))
    q0 = CO.first_guess(tile)
    q = GP.DeviceState(tile.geom)
    for n in FIELDS:
        q[n].data.copy_(torch.from_numpy(q0[n]))
    stage.init(q)
    g = tile.geom.c()
    dt = Fab.at_loc(tile.geom, 0, 1, 0)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for n in range(CO.N_ITER):
        nu = stage.nu_view()
        _lib.check(_lib.lib.gust_local_dtau(C.byref(g), C.byref(q["u"].array4()), C.byref(q["v"].array4()),
                                            C.byref(q["w"].array4()), C.byref(dm.z_nd.array4()),
                                            C.byref(dm.detJ.array4()), C.byref(nu), CO.cfl_at(n),
                                            C.byref(dt.array4()), st))
        stage.iterate(q, dt)
    return tile, {n: q[n].numpy().copy() for n in FIELDS}


def test_neutral_column_converges_to_the_oracle_composite(cuda):
    gold = np.load(GOLD)
    assert int(gold["n_iter"]) == CO.N_ITER
    norms = gold["norms"]
    drop = norms[-1][0] / norms[0][0]
    print(f"composite: u residual {norms[0][0]:.3e} -> {norms[-1][0]:.3e} ({drop:.2e})", flush=True)
    assert drop <= 1e-6 and norms[-1][0] <= 1e-8   # the deck's 6 orders and residual_tol
    tile, f = _device_run()
    for n in FIELDS:
        a, b = f[n][0, 2:-2, 2:-2, 2:-2], gold[f"q_{n}"][0, 2:-2, 2:-2, 2:-2]
        den = float(np.linalg.norm(b))
        e = float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a))
        print(f"  {n}: device vs composite relative L2 {e:.2e}", flush=True)
        assert e <= 1e-8 if den > 0 else e <= 1e-12, (n, e)
    check_expected_results(tile, f)


def check_expected_results(tile, f):
    """Tests/Canonical_RANS/Neutral_Column/README.md:20-31 on a converged state."""
    m = PHY.omesh(tile)
    _, _, _, phys = CO.column_case()
    qd = {n: f[n].copy() for n in FIELDS}
    P.fill_ghosts(qd, m, phys)
    R, cf, ss = P.residual(qd, m, phys)
    ustar = float(np.mean(ss[P.SS_USTAR]))
    assert abs(ustar - np.sqrt(1.225e-4 * 1000.0)) <= 1e-4 * 0.35, ustar
    u = f["u"][0, 2:-2, 2:-2, 2:-2]
    k = f["tke"][0, 2:-2, 2:-2, 2:-2]
    for a in (u, k):
        assert np.abs(a - a[:, :1, :1]).max() <= 1e-10 * np.abs(a).max()
    zl = tile.zlevels
    zc = 0.5 * (zl[:-1] + zl[1:])
    uc = 0.5 * (u[:, 0, 0] + u[:, 0, 1])           # cell-centred speed of the first column
    law = ustar / 0.41 * np.log((zc + 0.1) / 0.1)
    assert abs(uc[0] - law[0]) <= 1e-6 * uc[0]
    rel = uc / law - 1.0
    print(f"  log-law deviation of cells 1-8: {np.round(100 * rel[:8], 2)} %", flush=True)
    assert abs(rel[1]) <= 0.01
    assert np.all(np.abs(rel[:8]) <= 0.02)
    assert abs(k[0, 0, 0] - ustar * ustar / 0.5562 ** 2) <= 1e-6 * k[0, 0, 0]
    nu = cf["nu"][0, 1:-1, 1:-1, 1:-1]
    nut = nu[1, 0, 0]
    ref_nut = 0.41 * ustar * (zc[1] + 0.1)
    print(f"  nu_t of the second cell {nut:.4f} vs kappa u* (z + z0) {ref_nut:.4f}", flush=True)
    assert abs(nut - ref_nut) <= 0.06 * ref_nut
