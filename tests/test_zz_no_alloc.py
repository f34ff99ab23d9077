"""The z-line Jacobi entry point that allocates nothing
(gust_zline_jacobi_noalloc) against gust_zline_jacobi, through the C-ABI.

Its scratch is one fixed array per device that every call reuses, so the
checks run two calls in a row, the second on a smaller system: the second
must not see anything the first left behind (the change accumulator is
cleared by each call). FP64 z-line Jacobi is integer-exact work in the sense
of the parity bar: the results must be bit-identical to the allocating entry
point's, which the other GPU tests hold bit-identical to the oracle.
"""
import numpy as np
import pytest

from tests import linesys as L

pytestmark = pytest.mark.gpu


def _fabs(n, loc, periodic, seed):
    import torch
    import ablsteadyrans_amd as G
    A, ap, rhs = L.random_system(n, loc, periodic, seed=seed)
    geom = G.Geometry.of_size(*n, periodic=periodic)
    dev = [G.Fab(torch.from_numpy(np.ascontiguousarray(x)).cuda(), (0, 0, 0)) for x in (A, ap, rhs)]
    return geom, dev


def _both(n, loc, periodic, sweeps, seed):
    import ablsteadyrans_amd as G
    from ablsteadyrans_amd.line_sweeps import zline_jacobi_noalloc
    geom, (A, ap, rhs) = _fabs(n, loc, periodic, seed)
    d_ref = G.Fab.at_loc(geom, loc, 1, 1)
    d_ref.data.fill_(5.0)
    ch_ref = G.zline_jacobi(loc, A, ap, rhs, sweeps, geom, d_ref)
    d = G.Fab.at_loc(geom, loc, 1, 1)
    d.data.fill_(-3.0)   # the call must zero it, ghosts included
    ch = zline_jacobi_noalloc(loc, A, ap, rhs, sweeps, geom, d)
    return d_ref.numpy(), ch_ref, d.numpy(), ch


@pytest.mark.parametrize("loc", (0, 1, 2, 4))
def test_noalloc_equals_the_allocating_entry_point_twice_in_a_row(cuda, loc):
    # first a large system, then a smaller one on the same fixed scratch
    for n, periodic, sweeps, seed in (((64, 48, 40), (1, 0), 3, 11 + loc), ((9, 7, 6), (0, 1), 2, 23 + loc)):
        ref, ch_ref, got, ch = _both(n, loc, periodic, sweeps, seed)
        assert got.shape == ref.shape
        assert np.array_equal(got, ref), (n, loc, float(np.max(np.abs(got - ref))))
        assert ch == ch_ref, (n, loc, ch, ch_ref)


def test_noalloc_matches_the_oracle_on_the_second_call(cuda):
    # the smaller second call against the oracle itself as well
    from ablsteadyrans_amd.line_sweeps import zline_jacobi_noalloc
    import ablsteadyrans_amd as G
    _both((40, 40, 30), 1, (1, 1), 2, 7)
    n, loc, periodic = (12, 10, 8), 1, (1, 0)
    A, ap, rhs = L.random_system(n, loc, periodic, seed=99)
    geom, fa = _fabs(n, loc, periodic, 99)
    d = G.Fab.at_loc(geom, loc, 1, 1)
    ch = zline_jacobi_noalloc(loc, *fa, 2, geom, d)
    ref, ref_ch, _ = L.oracle_jacobi(n, loc, periodic, A, ap, rhs, 2)
    assert np.array_equal(d.numpy()[0, 1:-1, 1:-1, 1:-1], ref)
    assert ch == ref_ch


def test_noalloc_refuses_systems_beyond_its_scratch(cuda):
    import torch
    import ablsteadyrans_amd as G
    from ablsteadyrans_amd.line_sweeps import zline_jacobi_noalloc
    # 320 x 320 x 96 cells: 9.8 M unknowns, 79 MB of iterate > 64 MiB
    n = (320, 320, 96)
    geom = G.Geometry.of_size(*n, periodic=(0, 0))
    A = G.Fab(torch.zeros((7, n[2], n[1], n[0]), dtype=torch.float64, device="cuda"), (0, 0, 0))
    ap = G.Fab(torch.ones((1, n[2], n[1], n[0]), dtype=torch.float64, device="cuda"), (0, 0, 0))
    rhs = G.Fab(torch.zeros((1, n[2], n[1], n[0]), dtype=torch.float64, device="cuda"), (0, 0, 0))
    d = G.Fab.at_loc(geom, 0, 1, 1)
    with pytest.raises(G.GustError, match="more than the fixed"):
        zline_jacobi_noalloc(0, A, ap, rhs, 1, geom, d)
    del A, ap, rhs, d
    torch.cuda.empty_cache()
    # and it still works afterwards
    ref, ch_ref, got, ch = _both((16, 16, 12), 0, (1, 1), 2, 5)
    assert np.array_equal(got, ref) and ch == ch_ref
