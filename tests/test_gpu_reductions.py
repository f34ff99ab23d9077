"""The control loop's reductions on the GPU against the oracle: the
column-ordered sum bit for bit (per box of the 3 x 3 layout it reproduces
the bits test_Reductions.cpp pins), the exact maximum with NaN and +0, the
scaled sums bit for bit, the stage norms to 1e-13 relative (a fixed device
tree against the oracle's loop order; the same bits on every call) and the
confidence share and height exactly."""
from __future__ import annotations

import math
import struct

import numpy as np
import pytest

from oracle import oracle as orc
from tests import reduction_helpers as RH

B1 = (-1, -1, -1)
Z3 = (0, 0, 0)


def _fab(a, begin):
    import torch
    from ablsteadyrans_amd.fab import Fab
    return Fab(torch.from_numpy(np.ascontiguousarray(a)).cuda(), begin)


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


@pytest.mark.gpu
def test_deterministic_sum_per_box_reproduces_the_pinned_bits(cuda):
    from ablsteadyrans_amd import reductions as RD
    from ablsteadyrans_amd.fab import Geometry
    a = RH.pattern_field(3)
    fab = _fab(a, B1)
    total = 0.0
    for x in RH.XS:
        for y in RH.YS:
            geom = Geometry((x[0], y[0], 0), (x[1], y[1], 15))
            s = RD.deterministic_sum(geom, fab)
            assert bits(s) == bits(orc.box_sum(a, B1, geom.lo, geom.hi)), (x, y)
            total += s
    assert bits(total) == RH.SUM_BITS


@pytest.mark.gpu
def test_deterministic_sum_random_tile_and_component(cuda):
    from ablsteadyrans_amd import reductions as RD
    from ablsteadyrans_amd.fab import Geometry
    rng = np.random.default_rng(41)
    a = rng.normal(size=(2, 50, 66, 68))   # two components, one ghost around 64 x 64 x 48
    geom = Geometry((0, 0, 0), (63, 63, 47))
    got = RD.deterministic_sum(geom, _fab(a, B1), comp=1)
    assert bits(got) == bits(orc.box_sum(np.ascontiguousarray(a[1:2]), B1, geom.lo, geom.hi))


@pytest.mark.gpu
def test_global_max_is_exact_with_nan_and_positive_zero(cuda):
    from ablsteadyrans_amd import reductions as RD
    from ablsteadyrans_amd.fab import Geometry
    geom = Geometry((0, 0, 0), (23, 19, 15))
    a = RH.pattern_field(5)
    a[0, 0] = 1e30
    assert RD.global_max(geom, _fab(a, B1)) == a[0, 1:-1, 1:-1, 1:-1].max()
    m = RD.global_max(geom, _fab(np.full_like(a, -0.0), B1))
    assert m == 0 and math.copysign(1.0, m) > 0
    n = np.full_like(a, -1.0)
    n[0, 1 + 9, 1 + 3, 1 + 17] = np.nan
    assert math.isnan(RD.global_max(geom, _fab(n, B1)))
    # x-faces: the hi face counts
    u = np.full((1, 16, 20, 25), -2.0)
    u[0, 4, 5, 24] = 3.5
    assert RD.global_max(geom, _fab(u, Z3), loc=1) == 3.5


@pytest.mark.gpu
@pytest.mark.parametrize("loc", [0, 1, 2, 4])
@pytest.mark.parametrize("per", [(0, 0), (1, 1)])
def test_scaled_sums_match_the_oracle_bitwise(cuda, loc, per):
    from ablsteadyrans_amd import reductions as RD
    from ablsteadyrans_amd.fab import Geometry
    geom = Geometry((0, 0, 0), (30, 17, 11), periodic=per)
    g = orc.geom_c(geom.lo, geom.hi, geom.periodic)
    n = [31 + (loc & 1), 18 + ((loc >> 1) & 1), 12 + ((loc >> 2) & 1)]
    rng = np.random.default_rng(loc + 10 * per[0])
    q, v, off = (rng.normal(size=(1, n[2], n[1], n[0])) for _ in range(3))
    for offset in (None, off):
        ref = orc.scaled_sums(q, v, Z3, 3.0, loc, g, offset=offset)
        got = RD.scaled_sums(geom, loc, _fab(q, Z3), _fab(v, Z3), 3.0,
                             offset=_fab(off, Z3) if offset is not None else None)
        assert [bits(x) for x in got] == [bits(x) for x in ref]


def _random_state(geom, seed):
    n = geom.n
    rng = np.random.default_rng(seed)
    R = []
    for loc in (1, 2, 4, 0, 0, 0):
        sh = (1, n[2] + ((loc >> 2) & 1), n[1] + ((loc >> 1) & 1), n[0] + (loc & 1))
        R.append(rng.normal(scale=1e-3, size=sh))
    vel = [rng.normal(size=(1, n[2] + 2 + (d == 2), n[1] + 2 + (d == 1), n[0] + 2 + (d == 0))) for d in range(3)]
    tke = rng.uniform(0.01, 1.0, size=(1, n[2] + 2, n[1] + 2, n[0] + 2))
    ustar = rng.uniform(0.1, 0.5, size=(1, 1, n[1] + 2, n[0] + 2))
    return R, vel, tke, ustar


@pytest.mark.gpu
@pytest.mark.parametrize("with_ustar", [False, True])
def test_stage_norms_match_the_oracle(cuda, with_ustar):
    from ablsteadyrans_amd import reductions as RD
    from ablsteadyrans_amd.fab import Geometry
    geom = Geometry.of_size(64, 48, 40, periodic=(1, 0))
    g = orc.geom_c(geom.lo, geom.hi, geom.periodic)
    R, vel, tke, ustar = _random_state(geom, 5)
    sc = RD.JvpScales()
    H = 1700.0
    ref_res, ref_mon = orc.stage_norms(R, Z3, *vel, B1, tke, B1, sc.velocity, sc.theta, sc.tke, H, g,
                                       ustar if with_ustar else None, (-1, -1, 0))
    dR = [_fab(r, Z3) for r in R]
    dv = [_fab(x, B1) for x in vel]
    dus = _fab(ustar, (-1, -1, 0)) if with_ustar else None
    res, mon = RD.stage_norms(geom, dR, *dv, _fab(tke, B1), sc, H, ustar=dus)
    np.testing.assert_allclose(res, ref_res, rtol=1e-13, atol=0)
    assert mon[0] == ref_mon[0]
    np.testing.assert_allclose(mon[1:], ref_mon[1:], rtol=1e-13, atol=0)
    assert len(mon) == (3 if with_ustar else 2)
    res2, mon2 = RD.stage_norms(geom, dR, *dv, _fab(tke, B1), sc, H, ustar=dus)
    assert [bits(x) for x in res2 + mon2] == [bits(x) for x in res + mon]


@pytest.mark.gpu
def test_confidence_matches_the_oracle_exactly(cuda):
    from ablsteadyrans_amd import reductions as RD
    from ablsteadyrans_amd.poisson import DeviceMesh
    from tests import massflux_helpers as MH
    mesh = MH.hill_mesh(1)
    geom = mesh.geom
    g = MH.ogeom(mesh)
    R, _, _, _ = _random_state(geom, 8)
    sc = RD.JvpScales()
    H = 600.0
    for th in ([1e-2] * 6, [4e-2, 4e-2, 4e-2, 1.0, 1.0, 1.0], [1.0] * 6):
        ref = orc.confidence(R, Z3, mesh.z_nd, B1, sc.velocity, sc.theta, sc.tke, H, th, g)
        got = RD.confidence(geom, [_fab(r, Z3) for r in R], DeviceMesh(mesh).z_nd, sc, H, th)
        assert got == ref, th
    assert 0.0 < ref[0] <= 1.0
