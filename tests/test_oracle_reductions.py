"""The oracle's reductions pinned to Tests/Unit_Tests/test_Reductions.cpp
(the sum's bits on the 3 x 3-box layout, its distance to the exact sum, the
exact maximum, NaN and +0) and to closed forms of the scaled sums, the stage
norms, the monitors and the confidence share (constant fields, a flat mesh
whose heights are known, a single failing cell or face)."""
from __future__ import annotations

import math
import struct

import numpy as np
import pytest

from oracle import oracle as orc
from tests import reduction_helpers as RH

B1 = (-1, -1, -1)


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _box_sums(a, order):
    t = 0.0
    for x, y in order:
        t += orc.box_sum(a, B1, (x[0], y[0], 0), (x[1], y[1], RH.N_CELL[2] - 1))
    return t


def test_box_sums_reproduce_the_pinned_bits():
    """test_Reductions.cpp:44-79: the layout's sum has the pinned bits for
    either box order, lies within 16 eps sum|x| of the exact sum, and a
    plane-by-plane order (a plausible mistake) gives other bits."""
    a = RH.pattern_field(3)
    xmajor = [(x, y) for x in RH.XS for y in RH.YS]
    ymajor = [(x, y) for y in RH.YS for x in RH.XS]
    s = _box_sums(a, xmajor)
    assert bits(s) == RH.SUM_BITS
    assert bits(_box_sums(a, ymajor)) == RH.SUM_BITS
    valid = a[0, 1:-1, 1:-1, 1:-1]
    assert math.fsum(valid.ravel()) == RH.EXACT_SUM
    assert abs(s - RH.EXACT_SUM) <= 16 * np.finfo(float).eps * np.abs(valid).sum()
    wrong = 0.0
    for x, y in xmajor:
        box = 0.0
        for k in range(RH.N_CELL[2]):
            plane = 0.0
            for j in range(y[0], y[1] + 1):
                for i in range(x[0], x[1] + 1):
                    plane += valid[k, j, i]
            box += plane
        wrong += box
    assert bits(wrong) != RH.SUM_BITS


def test_box_sum_ignores_ghost_cells():
    a = RH.pattern_field(3)
    b = a.copy()
    b[0, 0], b[0, -1] = 1e30, 1e30
    b[0, :, 0], b[0, :, -1], b[0, :, :, 0], b[0, :, :, -1] = 1e30, 1e30, 1e30, 1e30
    lo, hi = (0, 0, 0), (23, 19, 15)
    assert orc.box_sum(a, B1, lo, hi) == orc.box_sum(b, B1, lo, hi)


def test_max_is_exact_nan_wins_and_zero_is_positive():
    a = RH.pattern_field(5)
    a[0, 0] = 1e30
    lo, hi = (0, 0, 0), (23, 19, 15)
    assert orc.global_max(a, B1, lo, hi) == a[0, 1:-1, 1:-1, 1:-1].max()
    z = np.full_like(a, -0.0)
    m = orc.global_max(z, B1, lo, hi)
    assert m == 0 and not math.copysign(1.0, m) < 0
    n = np.full_like(a, -1.0)
    n[0, 1 + 9, 1 + 3, 1 + 17] = np.nan
    assert math.isnan(orc.global_max(n, B1, lo, hi))


@pytest.mark.parametrize("loc", [0, 1, 2, 4])
@pytest.mark.parametrize("per", [(0, 0), (1, 1)])
def test_scaled_sums_closed_forms(loc, per):
    """Constant dyadic q, v, offset: the sums are exact products of the
    owned-point count; a periodic nodal axis owns one copy of its seam."""
    g = orc.geom_c((0, 0, 0), (5, 4, 3), per)
    n = [6 + (loc & 1), 5 + ((loc >> 1) & 1), 4 + ((loc >> 2) & 1)]
    q = np.full((1, n[2], n[1], n[0]), -3.0)
    v = np.full_like(q, 0.75)
    off = np.full_like(q, 1.0)
    own = [n[0] - ((loc & 1) and per[0]), n[1] - (((loc >> 1) & 1) and per[1]), n[2]]
    N = own[0] * own[1] * own[2]
    s0, s1, s2 = orc.scaled_sums(q, v, (0, 0, 0), 0.5, loc, g)
    assert (s0, s1, s2) == (N * 6.0, N * 2.25, float(N))
    t0, _, _ = orc.scaled_sums(q, v, (0, 0, 0), 0.5, loc, g, offset=off)
    assert t0 == N * 8.0


def test_scaled_sums_random_against_a_brute_force_sum():
    g = orc.geom_c((0, 0, 0), (9, 7, 5), (1, 0))
    rng = np.random.default_rng(3)
    q = rng.normal(size=(1, 6, 8, 11))
    v = rng.normal(size=(1, 6, 8, 11))
    s0, s1, s2 = orc.scaled_sums(q, v, (0, 0, 0), 2.5, 1, g)
    owned = (slice(None), slice(None), slice(0, 10))   # x-face 10 is face 0's copy
    assert s2 == 6 * 8 * 10
    assert s0 == pytest.approx(math.fsum(np.abs(q[0][owned]).ravel() / 2.5), rel=1e-14)
    assert s1 == pytest.approx(math.fsum(((v[0][owned] / 2.5) ** 2).ravel()), rel=1e-14)


def test_residual_scales_closed_form():
    # Source/Simplec.cpp:374-382 with U = 10, theta_s = 10, k_s = 0.1, H = 1000
    s = orc.residual_scales(10.0, 10.0, 0.1, 1000.0)
    assert s == [0.1, 0.1, 0.1, 0.01, 0.1, 10.0 * 0.1 / 1000.0]


def _flat_z(nx, ny, nz, dz):
    """Node heights k dz with one ghost layer, begin (-1,)*3."""
    z = np.zeros((1, nz + 3, ny + 3, nx + 3))
    z[0] = (np.arange(-1, nz + 2) * dz)[:, None, None]
    return z


def _fields(g, vals):
    n = [g.hi[d] - g.lo[d] + 1 for d in range(3)]
    out = []
    for q, loc in enumerate((1, 2, 4, 0, 0, 0)):
        sh = (1, n[2] + ((loc >> 2) & 1), n[1] + ((loc >> 1) & 1), n[0] + (loc & 1))
        out.append(np.full(sh, vals[q]))
    return out


def test_stage_norms_closed_forms():
    """Constant residuals give |c| / scale; the monitors are the largest
    speed, the mean k and the mean u* over the ground slab."""
    g = orc.geom_c((0, 0, 0), (7, 5, 3), (1, 1), (10.0, 10.0, 1.0))
    c = [0.5, -0.25, 2.0, 1e-3, -4.0, 0.125]
    R = _fields(g, c)
    u, v, w = _fields(g, [1.0, -2.0, 0.5, 0, 0, 0])[:3]
    v[0, 2, 3, 4] = -7.5
    tke = np.full((1, 4, 6, 8), 0.5)
    tke[0, 1, 2, 3] = 0.5 + 192.0   # mean 0.5 + 1
    ustar = np.full((1, 1, 6, 8), 0.25)
    H, U = 800.0, 10.0
    res, mon = orc.stage_norms(R, (0, 0, 0), u, v, w, (0, 0, 0), tke, (0, 0, 0), U, 10.0, 0.1, H, g, ustar,
                               (0, 0, 0))
    scale = orc.residual_scales(U, 10.0, 0.1, H)
    for q in range(6):
        assert res[q] == pytest.approx(abs(c[q]) / scale[q], rel=1e-15)
    assert mon == [7.5, 1.5, 0.25]
    _, mon2 = orc.stage_norms(R, (0, 0, 0), u, v, w, (0, 0, 0), tke, (0, 0, 0), U, 10.0, 0.1, H, g)
    assert mon2 == [7.5, 1.5]


def test_confidence_counts_cells_and_reports_their_height():
    """All residuals zero: share 1, height 0. One failing p cell at level k:
    share 1 - 1/N and height (k + 1/2) dz on the flat mesh. One failing u
    face fails both cells beside it."""
    nx, ny, nz, dz = 8, 6, 5, 4.0
    g = orc.geom_c((0, 0, 0), (nx - 1, ny - 1, nz - 1), (1, 1), (10.0, 10.0, dz))
    z = _flat_z(nx, ny, nz, dz)
    N = nx * ny * nz
    th = [1e-3] * 6
    R = _fields(g, [0.0] * 6)
    assert orc.confidence(R, (0, 0, 0), z, B1, 10.0, 10.0, 0.1, 1000.0, th, g) == (1.0, 0.0)
    R[3][0, 3, 2, 1] = 1.0
    share, top = orc.confidence(R, (0, 0, 0), z, B1, 10.0, 10.0, 0.1, 1000.0, th, g)
    assert share == 1.0 - 1.0 / N and top == 3.5 * dz
    R = _fields(g, [0.0] * 6)
    R[0][0, 1, 2, 4] = 1.0   # x-face 4: cells 3 and 4
    share, top = orc.confidence(R, (0, 0, 0), z, B1, 10.0, 10.0, 0.1, 1000.0, th, g)
    assert share == 1.0 - 2.0 / N and top == 1.5 * dz
    # below the threshold times the scale: converged
    scale = orc.residual_scales(10.0, 10.0, 0.1, 1000.0)
    R[0][0, 1, 2, 4] = th[0] * scale[0]
    assert orc.confidence(R, (0, 0, 0), z, B1, 10.0, 10.0, 0.1, 1000.0, th, g)[0] == 1.0
