"""GPU parity of the line solvers against the oracle (through the C-ABI).

FP64: bit-identical to the oracle (same operations in the same order, no
contraction on either side), checked with array_equal. FP32: the
mixed-precision preconditioner path, within 2e-5 of the FP64 oracle relative
to the field's max.
"""
import numpy as np
import pytest

from tests import linesys as L

pytestmark = pytest.mark.gpu

LOCS = (0, 1, 2, 4)
CASES = [((6, 5, 4), (0, 0)), ((6, 5, 4), (1, 1)), ((7, 4, 5), (1, 0)), ((5, 6, 3), (0, 1)),
         ((8, 1, 4), (1, 1)), ((1, 1, 9), (0, 0)), ((33, 17, 12), (1, 0)), ((64, 64, 48), (1, 0))]


@pytest.fixture(params=[1, 2, 4, 0], ids=["streaming", "staged", "register", "autotuned"])
def schedule(request, cuda):
    """Every z-line Jacobi kernel schedule (the streaming ring, the LDS-staged
    tile, the register kernel, and the autotuner's choice among them, the
    prefetch depths, global c' and c' in registers) must give the oracle's
    bits."""
    import ablsteadyrans_amd as LS
    LS.set_jacobi_schedule(request.param)
    yield request.param
    LS.set_jacobi_schedule(LS.JACOBI_AUTO)


def _dev(arr, begin, dtype=None):
    import torch
    import ablsteadyrans_amd as G
    t = torch.from_numpy(np.ascontiguousarray(arr)).to("cuda")
    if dtype is not None:
        t = t.to(dtype)
    return G.Fab(t.contiguous(), begin)


def _gpu_jacobi(n, loc, periodic, A, ap, rhs, sweeps, dtype=None, ngrow=1):
    import torch
    import ablsteadyrans_amd as G
    geom = G.Geometry.of_size(*n, periodic=periodic)
    dA, dap, drhs = _dev(A, (0, 0, 0), dtype), _dev(ap, (0, 0, 0), dtype), _dev(rhs, (0, 0, 0), dtype)
    d = G.Fab.at_loc(geom, loc, 1, ngrow, dtype=dtype or torch.float64)
    d.data.fill_(7.0)   # the call must zero it
    ch = G.zline_jacobi(loc, dA, dap, drhs, sweeps, geom, d)
    out = d.numpy().astype(np.float64)
    g = ngrow
    return out[0, g:-g, g:-g, g:-g], ch, out


@pytest.mark.parametrize("n,periodic", CASES)
@pytest.mark.parametrize("loc", LOCS)
def test_jacobi_fp64_is_bit_identical_to_the_oracle(schedule, n, periodic, loc):
    A, ap, rhs = L.random_system(n, loc, periodic, seed=100 + loc)
    for sweeps in (1, 2, 3):
        ref, ref_ch, ref_full = L.oracle_jacobi(n, loc, periodic, A, ap, rhs, sweeps)
        got, ch, full = _gpu_jacobi(n, loc, periodic, A, ap, rhs, sweeps)
        assert np.array_equal(got, ref), (sweeps, np.max(np.abs(got - ref)))
        assert ch == ref_ch
        # ghosts: zero outside the domain (periodic images are not kept)
        ghost = full.copy()
        ghost[0, 1:-1, 1:-1, 1:-1] = 0
        assert np.all(ghost == 0)


@pytest.mark.parametrize("loc", LOCS)
def test_jacobi_fp32_is_within_tolerance(schedule, loc):
    import torch
    n, periodic = (33, 17, 24), (1, 0)
    A, ap, rhs = L.random_system(n, loc, periodic, seed=7 + loc)
    for sweeps in (1, 2):
        ref, ref_ch, _ = L.oracle_jacobi(n, loc, periodic, A, ap, rhs, sweeps)
        got, ch, _ = _gpu_jacobi(n, loc, periodic, A, ap, rhs, sweeps, dtype=torch.float32)
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(got - ref)) <= 2e-5 * scale
        assert abs(ch - ref_ch) <= 2e-5 * scale


def test_jacobi_wider_ghosts_and_offset_domain(cuda):
    """A domain not at the origin, delta with two ghosts: the view plumbing."""
    import torch
    import ablsteadyrans_amd as G
    from oracle import oracle as orc
    lo, n, periodic = (3, -2, 1), (9, 7, 6), (1, 1)
    geom = G.Geometry(lo, (lo[0] + n[0] - 1, lo[1] + n[1] - 1, lo[2] + n[2] - 1), (1, 1, 0))
    for loc in LOCS:
        A, ap, rhs = L.random_system(n, loc, periodic, seed=5)
        vlo, vhi = geom.loc_box(loc)
        d = np.zeros((1, vhi[2] - vlo[2] + 5, vhi[1] - vlo[1] + 5, vhi[0] - vlo[0] + 5))
        db = tuple(v - 2 for v in vlo)
        og = orc.geom_c(geom.lo, geom.hi, geom.periodic)
        ch_ref = orc.zline_jacobi(loc, A, vlo, ap, rhs, vlo, 2, og, d, db)
        dA, dap, drhs = _dev(A, vlo), _dev(ap, vlo), _dev(rhs, vlo)
        dd = G.Fab(torch.ones((1,) + d.shape[1:], dtype=torch.float64, device="cuda"), db)
        ch = G.zline_jacobi(loc, dA, dap, drhs, 2, geom, dd)
        got = dd.numpy()
        assert np.array_equal(got[:, 2:-2, 2:-2, 2:-2], d[:, 2:-2, 2:-2, 2:-2]) and ch == ch_ref
        # the ghosts (two layers) are zero; the reference leaves stale periodic
        # images there, which no caller reads
        ghost = got.copy()
        ghost[:, 2:-2, 2:-2, 2:-2] = 0
        assert np.all(ghost == 0)


@pytest.mark.parametrize("n,periodic", CASES[:6])
@pytest.mark.parametrize("loc", LOCS)
def test_lusgs_is_bit_identical_to_the_oracle(cuda, n, periodic, loc):
    import torch
    import ablsteadyrans_amd as G
    A, ap, rhs = L.random_system(n, loc, periodic, seed=200 + loc)
    geom = G.Geometry.of_size(*n, periodic=periodic)
    dA, dap, drhs = _dev(A, (0, 0, 0)), _dev(ap, (0, 0, 0)), _dev(rhs, (0, 0, 0))
    small = min(x for x in (n[0], n[1]) if x > 1) if max(n[0], n[1]) > 1 else 0
    for ghosts, overlap in ((0, 0), (1, 0), (1, 2)):
        if overlap > small:
            continue
        for sweeps, alt in ((1, 0), (2, 0), (2, 1), (3, 1)):
            for order in (0, 1):
                ref, ref_ch, ref_full = L.oracle_lusgs(n, loc, periodic, A, ap, rhs, sweeps, ghosts, overlap, order,
                                                       alt)
                d = G.Fab.at_loc(geom, loc, 1, 1, dtype=torch.float64)
                ch = G.zline_lusgs(loc, dA, dap, drhs, sweeps, geom, d,
                                   G.LineSweepOptions(kind=1, ghosts=ghosts, overlap=overlap, order=order,
                                                      alternate=alt))
                got = d.numpy()[0, 1:-1, 1:-1, 1:-1]
                assert np.array_equal(got, ref), (ghosts, overlap, sweeps, alt, order)
                assert ch == ref_ch


@pytest.mark.parametrize("n,periodic", [((20, 12, 24), (1, 1)), ((9, 7, 40), (0, 1)), ((300, 260, 3), (1, 0))])
@pytest.mark.parametrize("loc", LOCS)
def test_lusgs_deep_and_wide_tiles_are_bit_identical(cuda, n, periodic, loc):
    """Lagged ghosts without alternation on columns of 17-32 levels (held in
    registers), of more than 32 (in chunks) and on diagonals of more than 256
    columns: the oracle's bits and change."""
    import torch
    import ablsteadyrans_amd as G
    A, ap, rhs = L.random_system(n, loc, periodic, seed=300 + loc)
    geom = G.Geometry.of_size(*n, periodic=periodic)
    dA, dap, drhs = _dev(A, (0, 0, 0)), _dev(ap, (0, 0, 0)), _dev(rhs, (0, 0, 0))
    for sweeps in (1, 2):
        ref, ref_ch, _ = L.oracle_lusgs(n, loc, periodic, A, ap, rhs, sweeps)
        d = G.Fab.at_loc(geom, loc, 1, 1, dtype=torch.float64)
        ch = G.zline_lusgs(loc, dA, dap, drhs, sweeps, geom, d, G.LineSweepOptions(kind=1))
        assert np.array_equal(d.numpy()[0, 1:-1, 1:-1, 1:-1], ref), (n, loc, sweeps)
        assert ch == ref_ch


def test_line_sweeps_dispatcher_routes_the_options(cuda):
    import torch
    import ablsteadyrans_amd as G
    n, periodic, loc = (8, 8, 4), (1, 1), 0
    A, ap, rhs = L.random_system(n, loc, periodic, seed=61)
    geom = G.Geometry.of_size(*n, periodic=periodic)
    dA, dap, drhs = _dev(A, (0, 0, 0)), _dev(ap, (0, 0, 0)), _dev(rhs, (0, 0, 0))
    outs = []
    for opt in (G.LineSweepOptions(kind=1, ghosts=1, overlap=2), G.LineSweepOptions(kind=1),
                G.LineSweepOptions(kind=0)):
        d = G.Fab.at_loc(geom, loc, 1, 1, dtype=torch.float64)
        G.line_sweeps(opt, loc, dA, dap, drhs, 2, geom, d)
        outs.append(d.numpy().copy())
    ref_z, _, _ = L.oracle_lusgs(n, loc, periodic, A, ap, rhs, 2, ghosts=1, overlap=2)
    ref_l, _, _ = L.oracle_lusgs(n, loc, periodic, A, ap, rhs, 2)
    ref_j, _, _ = L.oracle_jacobi(n, loc, periodic, A, ap, rhs, 2)
    for got, ref in zip(outs, (ref_z, ref_l, ref_j)):
        assert np.array_equal(got[0, 1:-1, 1:-1, 1:-1], ref)


def test_batched_jacobi_equals_the_single_calls(schedule):
    """Uneven tiles and all four locations in one launch per sweep."""
    import torch
    import ablsteadyrans_amd as G
    rng = np.random.default_rng(9)
    systems, refs = [], []
    for t in range(12):
        n = (int(rng.integers(3, 40)), int(rng.integers(1, 30)), int(rng.integers(2, 50)))
        periodic = (int(rng.integers(0, 2)), int(rng.integers(0, 2)))
        loc = LOCS[t % 4]
        A, ap, rhs = L.random_system(n, loc, periodic, seed=300 + t)
        ref, ref_ch, _ = L.oracle_jacobi(n, loc, periodic, A, ap, rhs, 2)
        geom = G.Geometry.of_size(*n, periodic=periodic)
        d = G.Fab.at_loc(geom, loc, 1, 1, dtype=torch.float64)
        systems.append(G.LineSystem(loc, geom, _dev(A, (0, 0, 0)), _dev(ap, (0, 0, 0)), _dev(rhs, (0, 0, 0)), d))
        refs.append((ref, ref_ch))
    changes = torch.zeros(len(systems), dtype=torch.float64, device="cuda")
    G.zline_jacobi_batched(systems, 2, changes)
    torch.cuda.synchronize()
    ch = changes.cpu().numpy()
    for s, (ref, ref_ch), c in zip(systems, refs, ch):
        assert np.array_equal(s.delta.numpy()[0, 1:-1, 1:-1, 1:-1], ref)
        assert c == ref_ch


def test_jacobi_on_tall_columns(schedule):
    """nz at the config sizes (48, 64, 96) and beyond (200 levels)."""
    for nz in (96, 200):
        n, periodic = (70, 3, nz), (0, 1)
        A, ap, rhs = L.random_system(n, 0, periodic, seed=nz)
        ref, ref_ch, _ = L.oracle_jacobi(n, 0, periodic, A, ap, rhs, 2)
        got, ch, _ = _gpu_jacobi(n, 0, periodic, A, ap, rhs, 2)
        assert np.array_equal(got, ref) and ch == ref_ch


def test_plan_and_graph_replays_match_the_one_shot_calls(cuda):
    """The plan API (setup once, kernels only) equals the one-shot entry point
    bit for bit, and a HIP graph of plan.run replays correctly several times
    with eager work (a changed rhs) between the replays."""
    import torch
    import ablsteadyrans_amd as G
    n, periodic = (20, 13, 9), (1, 0)
    geom = G.Geometry.of_size(*n, periodic=periodic)
    systems, host = [], []
    for t, loc in enumerate(LOCS):
        A, ap, rhs = L.random_system(n, loc, periodic, seed=400 + t)
        host.append((loc, A, ap, rhs))
        systems.append(G.LineSystem(loc, geom, _dev(A, (0, 0, 0)), _dev(ap, (0, 0, 0)), _dev(rhs, (0, 0, 0)),
                                    G.Fab.at_loc(geom, loc, 1, 1)))
    plan = G.JacobiPlan(systems)
    assert plan.unknowns == sum(int(np.prod(L.shape_of(*L.valid_box(n, loc)))) for loc in LOCS)
    plan.run(3)
    torch.cuda.synchronize()
    for s, (loc, A, ap, rhs), c in zip(systems, host, plan.changes.cpu().numpy()):
        ref, ref_ch, _ = L.oracle_jacobi(n, loc, periodic, A, ap, rhs, 3)
        assert np.array_equal(s.delta.numpy()[0, 1:-1, 1:-1, 1:-1], ref) and c == ref_ch
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.run(2)   # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        plan.run(2)
    for rep in range(4):
        for s in systems:
            s.rhs.data.mul_(1.0 + 0.25 * rep)
        g.replay()
        torch.cuda.synchronize()
        replayed = [s.delta.numpy().copy() for s in systems]
        ch_rep = plan.changes.cpu().numpy().copy()
        plan.run(2)
        torch.cuda.synchronize()
        for s, r in zip(systems, replayed):
            assert np.array_equal(s.delta.numpy(), r), rep
        assert np.array_equal(plan.changes.cpu().numpy(), ch_rep)
    plan.close()


def test_autotuner_records_and_reuses_its_choice(cuda):
    """The first batched solve of a set of shapes times every candidate that
    applies and keeps the fastest; the same shapes again add no record and
    give the same bits; the staged candidate is skipped where a tile does
    not fit the LDS."""
    import importlib
    import torch
    import ablsteadyrans_amd as G
    LSM = importlib.import_module("ablsteadyrans_amd.line_sweeps")
    LSM.set_jacobi_schedule(LSM.JACOBI_AUTO)
    LSM.jacobi_tune_reset()
    rng = np.random.default_rng(21)
    specs = [((48, 40, 64), (1, 0), 0), ((40, 48, 64), (0, 1), 1), ((48, 40, 65), (1, 0), 4)]
    outs = []
    for rep in range(2):
        systems = []
        for t, (n, periodic, loc) in enumerate(specs):
            A, ap, rhs = L.random_system(n, loc, periodic, seed=500 + t)
            geom = G.Geometry.of_size(*n, periodic=periodic)
            d = G.Fab.at_loc(geom, loc, 1, 1, dtype=torch.float64)
            systems.append(G.LineSystem(loc, geom, _dev(A, (0, 0, 0)), _dev(ap, (0, 0, 0)), _dev(rhs, (0, 0, 0)), d))
        G.zline_jacobi_batched(systems, 3)
        torch.cuda.synchronize()
        outs.append([s.delta.numpy().copy() for s in systems])
        recs = LSM.jacobi_tuned()
        print(f"autotuner after pass {rep}: {recs}", flush=True)
        assert len(recs) == 1
    rec = recs[0]
    timed = {k: v for k, v in rec["ms"].items() if v is not None}
    assert rec["choice"] == min(timed, key=timed.get) and len(timed) >= 4
    assert rec["sweeps"] == 3 and rec["systems"] == 3
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
    LSM.set_jacobi_schedule(1)
    systems = []
    for t, (n, periodic, loc) in enumerate(specs):
        A, ap, rhs = L.random_system(n, loc, periodic, seed=500 + t)
        geom = G.Geometry.of_size(*n, periodic=periodic)
        d = G.Fab.at_loc(geom, loc, 1, 1, dtype=torch.float64)
        systems.append(G.LineSystem(loc, geom, _dev(A, (0, 0, 0)), _dev(ap, (0, 0, 0)), _dev(rhs, (0, 0, 0)), d))
    G.zline_jacobi_batched(systems, 3)
    torch.cuda.synchronize()
    LSM.set_jacobi_schedule(LSM.JACOBI_AUTO)
    for a, s in zip(outs[0], systems):
        assert np.array_equal(a, s.delta.numpy())
