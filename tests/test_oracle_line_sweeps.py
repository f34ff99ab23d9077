"""Pins the oracle's line solvers (oracle/line_sweeps.c) without the reference binary.

Each check is independent of the oracle's code: a banded LAPACK solve for the
Thomas column (Source/LineSolve_K.H:22-45); dense block-Jacobi steps and the
dense solution (Tests/Unit_Tests/test_Predictor.cpp:369-437); dense
block-triangular line Gauss-Seidel passes with the lagged seam ghosts
(test_Predictor.cpp:1172-1387); a replay of the box-local rule written from
the header text (test_LineSweeps.cpp:191-286); the alternating ordering's exact
solve (test_LineSweeps.cpp:529-557); the zero/lagged identity on one
non-periodic box (test_LineSweeps.cpp:354-362); and the refusals
(test_LineSweeps.cpp:618-657).
"""
import itertools

import numpy as np
import pytest
import scipy.linalg

from oracle import oracle as orc
from tests import linesys as L

LOCS = (0, 1, 2, 4)
CASES = [((6, 5, 4), (0, 0)), ((6, 5, 4), (1, 1)), ((7, 4, 5), (1, 0)), ((5, 6, 3), (0, 1))]


def test_thomas_matches_banded_lapack_solve():
    rng = np.random.default_rng(3)
    for nz in (1, 2, 7, 48):
        n = (1, 1, nz)
        A, ap, rhs = L.random_system(n, 0, (0, 0), seed=nz)
        x, ch, _ = L.oracle_jacobi(n, 0, (0, 0), A, ap, rhs, 1)
        ab = np.zeros((3, nz))
        ab[0, 1:] = -A[L.ZP, :-1, 0, 0]
        ab[1, :] = ap[0, :, 0, 0]
        ab[2, :-1] = -A[L.ZM, 1:, 0, 0]
        ref = scipy.linalg.solve_banded((1, 1), ab, rhs[0, :, 0, 0])
        assert np.max(np.abs(x[:, 0, 0] - ref)) <= 1e-13 * np.max(np.abs(ref))
        assert ch == np.max(np.abs(x))
    del rng


def test_zero_rhs_gives_exact_zero():
    n = (5, 4, 6)
    for loc in LOCS:
        A, ap, rhs = L.random_system(n, loc, (1, 0), seed=5)
        rhs[:] = 0.0
        x, ch, d = L.oracle_jacobi(n, loc, (1, 0), A, ap, rhs, 3)
        assert np.all(d == 0.0) and ch == 0.0
        x, ch, d = L.oracle_lusgs(n, loc, (1, 0), A, ap, rhs, 2, ghosts=1, overlap=2)
        assert np.all(d == 0.0) and ch == 0.0


@pytest.mark.parametrize("n,periodic", CASES)
@pytest.mark.parametrize("loc", LOCS)
def test_jacobi_sweeps_are_dense_block_jacobi_steps(n, periodic, loc):
    A, ap, rhs = L.random_system(n, loc, periodic, seed=11 + loc)
    N, M, _ = L.dense_matrices(n, loc, periodic, A, ap)
    b = rhs.reshape(-1)
    x = np.zeros_like(b)
    for sweeps in (1, 2, 3):
        prev = x
        x = np.linalg.solve(M, b - (N - M) @ prev)
        got, ch, d = L.oracle_jacobi(n, loc, periodic, A, ap, rhs, sweeps)
        scale = np.max(np.abs(x))
        assert np.max(np.abs(got.reshape(-1) - x)) <= 1e-12 * scale
        assert abs(ch - np.max(np.abs(x - prev))) <= 1e-12 * scale
        # ghosts outside a non-periodic face stay zero
        if not periodic[0]:
            assert np.all(d[0, :, :, 0] == 0) and np.all(d[0, :, :, -1] == 0)


@pytest.mark.parametrize("loc", LOCS)
def test_jacobi_converges_to_the_dense_solution(loc):
    n, periodic = (6, 5, 4), (1, 0)
    A, ap, rhs = L.random_system(n, loc, periodic, seed=21, extra=0.5, xscale=1.0)
    N, _, _ = L.dense_matrices(n, loc, periodic, A, ap)
    ref = np.linalg.solve(N, rhs.reshape(-1))
    got, ch, _ = L.oracle_jacobi(n, loc, periodic, A, ap, rhs, 400)
    assert np.max(np.abs(got.reshape(-1) - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert ch <= 1e-13


def _dense_lusgs(n, loc, periodic, A, ap, rhs, sweeps, alternate):
    lo, hi = L.valid_box(n, loc)
    nz, ny, nx = L.shape_of(lo, hi)
    x = np.zeros(nz * ny * nx)
    change = 0.0
    for s in range(sweeps):
        for forward in (True, False):
            snap = x.copy()
            before = x.copy()
            cols = L.pass_orders(nx, ny, forward, bool(alternate and s % 2 == 1))
            x = L.dense_sgs_pass(n, loc, periodic, A, ap, rhs, x, cols, snap)
            if s == sweeps - 1 and not forward:
                change = np.max(np.abs(before - x))
    x = x.reshape(nz, ny, nx)
    L.seam_consistent(x[None], loc, periodic)
    return x, change


@pytest.mark.parametrize("n,periodic", CASES)
@pytest.mark.parametrize("loc", LOCS)
def test_lusgs_lagged_matches_dense_block_triangular_passes(n, periodic, loc):
    A, ap, rhs = L.random_system(n, loc, periodic, seed=31 + loc)
    for sweeps, alternate in ((1, 0), (2, 0), (2, 1), (3, 1)):
        ref, ref_ch = _dense_lusgs(n, loc, periodic, A, ap, rhs, sweeps, alternate)
        got, ch, _ = L.oracle_lusgs(n, loc, periodic, A, ap, rhs, sweeps, ghosts=0, alternate=alternate)
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(got - ref)) <= 1e-12 * scale, (sweeps, alternate)
        assert abs(ch - ref_ch) <= 1e-12 * scale


@pytest.mark.parametrize("n,periodic", CASES)
def test_lusgs_diagonal_order_equals_sequential_bitwise(n, periodic):
    for loc in LOCS:
        A, ap, rhs = L.random_system(n, loc, periodic, seed=51)
        for ghosts, overlap in ((0, 0), (1, 0), (1, 2)):
            for sweeps, alt in ((1, 0), (3, 0), (2, 1), (3, 1)):
                s, cs, _ = L.oracle_lusgs(n, loc, periodic, A, ap, rhs, sweeps, ghosts, overlap, 0, alt)
                ls = orc.last_lusgs_launches()
                d, cd, _ = L.oracle_lusgs(n, loc, periodic, A, ap, rhs, sweeps, ghosts, overlap, 1, alt)
                ld = orc.last_lusgs_launches()
                assert np.array_equal(s, d) and cs == cd
                assert ls == 2 * sweeps and ld > ls
                assert np.max(np.abs(s)) > 0.1


def _replay_zero(n, loc, periodic, A, ap, rhs, sweeps, n_o):
    """The box-local rule of Source/LineSweeps.H:19-43 for one box, written
    from the header (as the reference test's replay, test_LineSweeps.cpp:191-286)."""
    lo, hi = L.valid_box(n, loc)
    nz, ny, nx = L.shape_of(lo, hi)
    Lc = (n[0], n[1])
    t = (loc & 1, (loc >> 1) & 1)

    def wrap(i, j):
        if periodic[0]:
            i = i % Lc[0]
        if periodic[1]:
            j = j % Lc[1]
        return i, j

    own_lo, own_hi = [0, 0], [nx - 1, ny - 1]
    for d in range(2):
        if t[d] and periodic[d]:
            own_lo[d] += 1
    S_lo, S_hi = list(own_lo), list(own_hi)
    for d in range(2):
        if Lc[d] == 1:
            continue
        S_lo[d] -= n_o
        S_hi[d] += n_o
        if not periodic[d]:
            S_lo[d] = max(S_lo[d], 0)
            S_hi[d] = min(S_hi[d], (nx, ny)[d] - 1)
    selfx = periodic[0] and Lc[0] == 1
    selfy = periodic[1] and Lc[1] == 1
    W = {}

    def w(i, j, k):
        return W.get((i, j, k), 0.0)

    def column(i, j):
        cc, ee = [0.0] * nz, [0.0] * nz
        wi, wj = wrap(i, j)
        for k in range(nz):
            a = A[:, k, wj, wi]
            pw = ap[0, k, wj, wi]
            xn = yn = 0.0
            if selfx:
                pw -= a[L.XM] + a[L.XP]
            else:
                xn = a[L.XM] * w(i - 1, j, k) + a[L.XP] * w(i + 1, j, k)
            if selfy:
                pw -= a[L.YM] + a[L.YP]
            else:
                yn = a[L.YM] * w(i, j - 1, k) + a[L.YP] * w(i, j + 1, k)
            rk = rhs[0, k, wj, wi] + xn + yn
            lower, upper = -a[L.ZM], -a[L.ZP]
            if k == 0:
                cc[k] = upper / pw
                ee[k] = rk / pw
            else:
                m = pw - lower * cc[k - 1]
                cc[k] = upper / m
                ee[k] = (rk - lower * ee[k - 1]) / m
        W[(i, j, nz - 1)] = ee[nz - 1]
        for k in range(nz - 2, -1, -1):
            W[(i, j, k)] = ee[k] - cc[k] * W[(i, j, k + 1)]

    for _s in range(sweeps):
        for j in range(S_lo[1], S_hi[1] + 1):
            for i in range(S_lo[0], S_hi[0] + 1):
                column(i, j)
        for j in range(S_hi[1], S_lo[1] - 1, -1):
            for i in range(S_hi[0], S_lo[0] - 1, -1):
                column(i, j)
    out = np.zeros((nz, ny, nx))
    for k in range(nz):
        for j in range(own_lo[1], own_hi[1] + 1):
            for i in range(own_lo[0], own_hi[0] + 1):
                wi, wj = wrap(i, j)
                out[k, wj, wi] = w(i, j, k)
    # a non-owned copy compares with its owner's value
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                wi, wj = wrap(i, j)
                out[k, j, i] = out[k, wj, wi]
    return out


@pytest.mark.parametrize("n,periodic", [((8, 8, 4), (0, 0)), ((8, 8, 4), (1, 1)), ((7, 6, 4), (1, 0)),
                                        ((8, 1, 4), (1, 1))])
def test_lusgs_zero_ghosts_match_the_header_replay(n, periodic):
    for loc in LOCS:
        A, ap, rhs = L.random_system(n, loc, periodic, seed=11 + loc)
        for n_o in (0, 2):
            if n_o > min(x for x in (n[0], n[1]) if x > 1):
                continue
            for sweeps in (1, 2):
                got, _, _ = L.oracle_lusgs(n, loc, periodic, A, ap, rhs, sweeps, ghosts=1, overlap=n_o)
                ref = _replay_zero(n, loc, periodic, A, ap, rhs, sweeps, n_o)
                scale = np.max(np.abs(ref))
                assert scale > 0.1
                assert np.max(np.abs(got - ref)) <= 1e-13 * scale, (loc, n_o, sweeps)


def test_zero_and_lagged_are_one_sweep_on_a_nonperiodic_box():
    n = (8, 8, 4)
    for loc in LOCS:
        A, ap, rhs = L.random_system(n, loc, (0, 0), seed=41)
        z, cz, _ = L.oracle_lusgs(n, loc, (0, 0), A, ap, rhs, 2, ghosts=1)
        g, cg, _ = L.oracle_lusgs(n, loc, (0, 0), A, ap, rhs, 2, ghosts=0)
        assert np.array_equal(z, g) and cz == cg


def _residual(n, loc, periodic, A, ap, rhs, x):
    N, _, _ = L.dense_matrices(n, loc, periodic, A, ap)
    return rhs.reshape(-1) - N @ x.reshape(-1)


def test_alternating_ordering_solves_opposite_sign_upwinding_exactly():
    n = (8, 8, 4)
    A, ap, rhs = L.random_system(n, 0, (0, 0), seed=71, extra=0.2)
    A[L.XP] = 0.0
    A[L.YM] = 0.0
    eps = np.finfo(float).eps
    x0, _, _ = L.oracle_lusgs(n, 0, (0, 0), A, ap, rhs, 2)
    x1, _, _ = L.oracle_lusgs(n, 0, (0, 0), A, ap, rhs, 2, alternate=1)
    scale = np.max(np.abs(rhs))
    assert np.max(np.abs(_residual(n, 0, (0, 0), A, ap, rhs, x1))) <= 1000 * eps * scale
    assert np.max(np.abs(_residual(n, 0, (0, 0), A, ap, rhs, x0))) > 1e-3 * scale
    # one sweep is not changed by the alternation; two are
    a1, _, _ = L.oracle_lusgs(n, 0, (0, 0), A, ap, rhs, 1)
    b1, _, _ = L.oracle_lusgs(n, 0, (0, 0), A, ap, rhs, 1, alternate=1)
    assert np.array_equal(a1, b1)
    assert np.max(np.abs(x0 - x1)) > 1e-6


def _outer_solve(n, loc, periodic, A, ap, rhs, solver, rtol=1e-12, maxit=2000):
    x = np.zeros_like(rhs[0])
    r0 = np.max(np.abs(rhs))
    for it in range(maxit):
        res = _residual(n, loc, periodic, A, ap, rhs, x).reshape(rhs.shape)
        if np.max(np.abs(res)) <= rtol * r0:
            return it, x
        dx = solver(res)
        x = x + dx
    return -1, x


def test_overlaps_reach_the_same_state_on_a_periodic_tile():
    n, periodic = (16, 8, 4), (1, 1)
    for loc in (0, 1):
        A, ap, rhs = L.random_system(n, loc, periodic, seed=21, extra=0.05, xscale=4.0)
        states = []
        for n_o in (0, 2, 4):
            its, x = _outer_solve(n, loc, periodic, A, ap, rhs,
                                  lambda r: L.oracle_lusgs(n, loc, periodic, A, ap, r, 2, ghosts=1, overlap=n_o)[0])
            assert its > 0
            states.append(x)
        for x in states[1:]:
            assert np.max(np.abs(x - states[0])) <= 1e-10 * np.max(np.abs(states[0]))


def test_refusals_name_their_keys():
    n = (8, 8, 4)
    A, ap, rhs = L.random_system(n, 0, (1, 1), seed=61)
    for ghosts, overlap, alt, key in ((1, -1, 0, "steady.line_overlap"), (1, 9, 0, "steady.line_overlap"),
                                      (0, 2, 0, "steady.line_overlap"), (1, 0, 2, "steady.line_alternate")):
        with pytest.raises(orc.OracleError, match=key):
            L.oracle_lusgs(n, 0, (1, 1), A, ap, rhs, 1, ghosts=ghosts, overlap=overlap, alternate=alt)
    # overlap wider than the (one) box: refused, the widest allowed passes
    n2 = (4, 4, 4)
    A2, ap2, rhs2 = L.random_system(n2, 0, (1, 1), seed=62)
    with pytest.raises(orc.OracleError, match="narrowest"):
        L.oracle_lusgs(n2, 0, (1, 1), A2, ap2, rhs2, 1, ghosts=1, overlap=5)
    L.oracle_lusgs(n2, 0, (1, 1), A2, ap2, rhs2, 1, ghosts=1, overlap=4)
    with pytest.raises(orc.OracleError, match="sweeps must be >= 1"):
        L.oracle_jacobi(n, 0, (1, 1), A, ap, rhs, 0)
    # delta without its ghost cell
    d = np.zeros((1, 4, 8, 8))
    with pytest.raises(orc.OracleError, match="ghost"):
        orc.zline_jacobi(0, A, (0, 0, 0), ap, rhs, (0, 0, 0), 1, L.geom(n, (1, 1)), d, (0, 0, 0))


def test_case_grid_is_covered():
    # every location x periodicity combination above produces nonzero sweeps
    for (n, periodic), loc in itertools.product(CASES, LOCS):
        A, ap, rhs = L.random_system(n, loc, periodic, seed=1)
        x, ch, _ = L.oracle_jacobi(n, loc, periodic, A, ap, rhs, 2)
        assert ch > 0 and np.max(np.abs(x)) > 0.1


@pytest.mark.parametrize("n,periodic", [((6, 5, 4), (1, 1)), ((33, 17, 12), (1, 0)), ((20, 21, 9), (0, 0))])
@pytest.mark.parametrize("loc", LOCS)
def test_threaded_jacobi_is_the_serial_jacobi_bitwise(n, periodic, loc):
    """The CPU baseline's threaded sweeps (the columns split over OpenMP
    threads) give the serial oracle's bits and change."""
    A, ap, rhs = L.random_system(n, loc, periodic, seed=40 + loc)
    ref, ref_ch, ref_d = L.oracle_jacobi(n, loc, periodic, A, ap, rhs, 3)
    for threads in (2, 3, 8):
        d, db = L.new_delta(n, loc)
        ch = orc.zline_jacobi(loc, A, (0, 0, 0), ap, rhs, (0, 0, 0), 3, L.geom(n, periodic), d, db, threads=threads)
        assert np.array_equal(d, ref_d) and ch == ref_ch, threads
