#!/usr/bin/env python3
"""Benchmark of the SIMPLEC hot path on one MI355X.

Headline (value): steady solves per second at a 6-order residual drop on the
Gaussian-hill configuration (h = 100 exp(-(r/250 m)^2) m in 4 x 4 km,
200 x 200 x 64 cells at dx = 20 m, 64 levels from dz1 = 1 m growing by 1.08;
the open-channel neutral ABL: log-law inflow, outflow, y periodic, surface
layer, slip top, Coriolis, the k-L closure) in the fast form: mixed
precision (FP32 z-line sweeps and FP32 multigrid V-cycles inside the FP64
defect correction: the residual, GMRES and the coarsest solve stay FP64)
with two levels of grid sequencing (the 100 x 100 level to the loose 2-order
stop, then the 200 x 200 level to the 6-order criterion, seeded with the
coarse level's references), with the solver settings of the reference's
3-D hill deck (16 sweeps, CFL 1 -> 4, rotation_relax 1). value = 1 / the
wall time of one whole sequenced solve from the coarse level's first guess
(hierarchy built before the timed region, state resident in HBM). One step
= one SIMPLEC iteration on the 2.56 M-cell tile (mixed): `warmup` driver
iterations from the first guess, then `steps` iterations from the first
guess timed between synchronisations (ms_per_step). The plain FP64 path
(one level, the drop-in form) and plain mixed precision are timed as whole
solves too (extra.c2_fp64, extra.c2_mixed); the per-field relative L2 of
each fast form's converged fields against the plain FP64 ones is reported
(extra.c2_field_error_*), and the ratio to the plain FP64 solves/s
(extra.fast_vs_fp64).

Also measured (extra): the reference's Neutral_Column deck (the previous
headline: 4 x 4 x 64, 2231 iterations) FP64 and mixed, and with the
Newton-Krylov finish (steady.stage = simplec_newton); the reference's own
steady decks with published CPU times (Askervein, Neutral_ABL_Flat,
Neutral_Hill_3D with z-line Jacobi, with line LU-SGS and with three levels of
grid sequencing); the z-line Jacobi sweep of the five variables on the
fractal-terrain tile (256 x 256 x 96, 65,536 columns; the roofline object:
algorithmic bytes per launch / launch time from HIP events on the launch
stream, 10 reals per unknown); the V(2,2) cycle. Serving and batched legs
(extra.serving, extra.batched). cpu_baseline: the oracle on the host cores
(the composite iteration on one thread, with the reference's published CPU
rate beside it; extra.cpu_sweep_baseline: the compiled oracle sweep on one
thread and on the host's threads).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NX, NY, NZ = 256, 256, 96
VARS = (("u", 1), ("v", 2), ("w", 4), ("theta", 0), ("k", 0))
REALS_PER_UNKNOWN = 10


def _systems(dtype, seed, periodic=(0, 0)):
    import torch
    import ablsteadyrans_amd as G
    geom = G.Geometry.of_size(NX, NY, NZ, periodic=periodic, dx=(25.0, 25.0, 1.0))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    systems = []
    for _name, loc in VARS:
        lo, hi = geom.loc_box(loc)
        shp = (hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
        A = torch.empty((G.N_COEFS,) + shp, dtype=torch.float64, device="cuda")
        nb = torch.rand((6,) + shp, generator=gen, dtype=torch.float64, device="cuda") * 0.9 + 0.1
        nb[0:2] *= 3.0
        A[1:] = nb
        A[0] = nb.sum(0)
        ap = (A[0] + 0.3).unsqueeze(0).contiguous()
        rhs = (torch.rand((1,) + shp, generator=gen, dtype=torch.float64, device="cuda") * 2 - 1)
        del nb
        d = G.Fab.at_loc(geom, loc, 1, 1, dtype=dtype)
        systems.append(G.LineSystem(loc, geom, G.Fab(A.to(dtype).contiguous(), lo),
                                    G.Fab(ap.to(dtype).contiguous(), lo), G.Fab(rhs.to(dtype).contiguous(), lo), d))
    return systems


def _time_sweeps(plan, steps, warmup):
    """Time `steps` steady-state sweep launches (ping-pong) with HIP events on
    the launch stream; returns (ms per sweep from events, ms per sweep wall)."""
    import torch
    stream = torch.cuda.current_stream()
    plan.run(2)  # a nonzero iterate in delta and the scratch iterate
    flip = 1
    for _ in range(warmup):
        plan.sweep(False, False, flip)
        flip ^= 1
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        plan.sweep(False, False, flip)
        flip ^= 1
    e1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return e0.elapsed_time(e1) / steps, (t1 - t0) * 1e3 / steps


_CHOICE_KERNELS = {
    "staged": "zjacobi_staged_kernel<{T},false,false>",
    "streaming-p4": "zjacobi_sweep_kernel<{T},false,false,true,4>",
    "streaming-p2": "zjacobi_sweep_kernel<{T},false,false,true,2>",
    "streaming-p8": "zjacobi_sweep_kernel<{T},false,false,true,8>",
    "streaming-global-cprime": "zjacobi_sweep_kernel<{T},false,false,false,4>",
    "register-e": "zjacobi_regs_kernel<{T},false,false,97,97>",
    "register-ce": "zjacobi_regs_kernel<{T},false,false,97,0>",
    "register-split": "zjacobi_regs_kernel<{T},false,false,97,80>",
}


def _rule_kernel(dtype_name):
    """The kernel the fixed schedule rule launches for a middle sweep of the
    five fractal-tile systems (columns of 96 and 97 levels): the register
    kernel, c' split between LDS and registers for FP64, in registers for
    FP32 (the profiles under profiles/ are of the same kernels)."""
    T = "double" if dtype_name == "f64" else "float"
    return _CHOICE_KERNELS["register-split" if dtype_name == "f64" else "register-ce"].format(T=T)


def _time_run(plan, sweeps, reps):
    import torch
    stream = torch.cuda.current_stream()
    plan.run(sweeps)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        plan.run(sweeps)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _cpu_threads():
    """The host threads the CPU baselines use: OMP_NUM_THREADS when set (the
    GPU hosts set their share), else the logical CPUs."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return max(1, os.cpu_count() or 1)


def _sweep_sample(nx, ny, rng):
    """The five variables' z-line systems on an nx x ny x NZ tile (diagonally
    dominant random rows), host arrays for the oracle."""
    import numpy as np
    from oracle import oracle as orc
    g = orc.geom_c((0, 0, 0), (nx - 1, ny - 1, NZ - 1), (0, 0, 0))
    data, unknowns = [], 0
    for _name, loc in VARS:
        sx, sy, sz = nx + (loc & 1), ny + ((loc >> 1) & 1), NZ + ((loc >> 2) & 1)
        A = np.zeros((7, sz, sy, sx))
        nb = rng.uniform(0.1, 1.0, size=(6, sz, sy, sx))
        nb[0:2] *= 3.0
        A[1:] = nb
        A[0] = nb.sum(0)
        ap = (A[0] + 0.3)[None].copy()
        rhs = rng.uniform(-1, 1, size=(1, sz, sy, sx))
        d = np.zeros((1, sz + 2, sy + 2, sx + 2))
        data.append((loc, A, ap, rhs, d))
        unknowns += sx * sy * sz
    return g, data, unknowns


def _time_sweep_sample(g, data, unknowns, threads, target_s):
    from oracle import oracle as orc
    calls = 0
    t0 = time.perf_counter()
    while True:
        for loc, A, ap, rhs, d in data:
            orc.zline_jacobi(loc, A, (0, 0, 0), ap, rhs, (0, 0, 0), 2, g, d, (-1, -1, -1), threads=threads)
        calls += 1
        if time.perf_counter() - t0 >= target_s:
            break
    dt = time.perf_counter() - t0
    per_sweep = dt / (2 * calls)
    return REALS_PER_UNKNOWN * 8 * unknowns / per_sweep / 1e9, per_sweep, 2 * calls, dt


def _cpu_baseline(target_s=8.0):
    """The oracle's z-line Jacobi (compiled C, -O3 -march=x86-64-v3, two
    sweeps per call; the Python loop only issues one call per variable) on
    the host: one thread on a 128 x 128 x 96 sample, and the host's threads
    (OpenMP over columns) on the fractal tile's own 256 x 256 x 96 systems;
    reported per sweep in the metric's unit (algorithmic GB/s)."""
    import numpy as np
    rng = np.random.default_rng(20264203)
    g1, d1, u1 = _sweep_sample(128, 128, rng)
    gbs1, ps1, n1, dt1 = _time_sweep_sample(g1, d1, u1, 1, target_s)
    del d1
    threads = _cpu_threads()
    gN, dN, uN = _sweep_sample(NX, NY, rng)
    gbsN, psN, nN, dtN = _time_sweep_sample(gN, dN, uN, threads, target_s)
    del dN
    model, ncpu = _cpu_info()
    return {"value": round(gbsN, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "cpu_model": model, "logical_cpus": ncpu,
            "sample": f"oracle zline_jacobi FP64 (compiled C, OpenMP over columns), 5 variables on {NX}x{NY}x{NZ} "
                      f"({uN} unknowns per sweep), {nN} sweeps in {dtN:.1f} s = {psN * 1e3:.1f} ms per sweep on "
                      f"{threads} threads; one thread on 128x128x{NZ} ({u1} unknowns): {n1} sweeps in {dt1:.1f} s = "
                      f"{ps1 * 1e3:.1f} ms per sweep, {gbs1:.3f} GB/s",
            "ms_per_sweep": round(psN * 1e3, 3), "single_thread_gbs": round(gbs1, 3),
            "single_thread_ms_per_sweep_128x128": round(ps1 * 1e3, 3)}


def _fractal_tile():
    """The fractal-terrain tile (decks.fractal_tile_256)."""
    from ablsteadyrans_amd import decks as D
    return D.fractal_tile_256()


def _time_vcycle(tile, precision, reps):
    """One V(2,2) cycle of the pressure-correction multigrid on the tile, with
    SIMPLEC-like face coefficients; returns (ms per cycle, levels)."""
    import torch
    from ablsteadyrans_amd import poisson as PS
    dm = PS.DeviceMesh(tile)
    S = PS.PoissonSolver(dm, PS.PoissonOptions(precision=precision))
    n = NX * NY * NZ
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    b = torch.rand(n, generator=gen, dtype=torch.float64, device="cuda") - 0.5
    x = torch.zeros_like(b)
    bc = (1, 1, 1, 2, 1, 2)   # neumann sides and ground, open x-hi and top
    S.vcycle(b, x, None, bc, setup=True)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        S.vcycle(b, x, None, bc, setup=False)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps, S.levels


def _mg_levels(nx, ny, nz):
    """Cells per multigrid level of a tile with no periodic side: x/y halved
    (odd sizes to (n + 1) / 2) while >= 4, until nx*ny <= 16 (or 12 levels)."""
    cells = []
    for lev in range(64):
        cells.append(nx * ny * nz)
        cx = 2 if nx >= 4 else 1
        cy = 2 if ny >= 4 else 1
        if (cx == 1 and cy == 1) or lev >= 12 or nx * ny <= 16:
            break
        nx = (nx + cx - 1) // cx
        ny = (ny + cy - 1) // cy
    return cells


def _vcycle_bytes(s, pre=2, post=2, coarse=16):
    """Algorithmic bytes of one preconditioner call (V(pre,post) with the
    explicit 15-point level stencils), reals of s bytes per cell:
      zebra sweep       18: 15 stencil coefficients (own column as lower,
                            1/pivot, c'; 12 neighbour), b, x of the other
                            colour (read once), x written;
      residual+restrict 18.25: 15 coefficients, b, x, J, 1/4 coarse b;
      prolongation       2.25: fine x read and written, 1/4 coarse x;
    coarse_sweeps sweeps on the coarsest level, and the FP64 vector's
    conversion into and out of the level-0 layout (8 + s each way)."""
    cells = _mg_levels(NX, NY, NZ)
    b = 0.0
    for n in cells[:-1]:
        b += n * s * ((pre + post) * 18 + 18.25 + 2.25)
    b += cells[-1] * s * coarse * 18
    b += cells[0] * (16 + 2 * s)
    return b


COLUMN_ITERATIONS_REF = 2231   # Tests/Canonical_RANS/Neutral_Column/README.md:26


def _vcycle_alg_bytes(s, pre=2, post=2):
    """The V-cycle's compulsory bytes with x/y semi-coarsening (levels
    Nc / 4^l): (4/3) Nc [(pre + post) 8 s + 8 s + 3.5 s] -- per smoothing
    pass the iterate read and written, rhs and the five mesh arrays; the
    residual pass; the transfers. 2.92 GB FP64 / 1.46 GB FP32 for V(2,2) on
    256 x 256 x 96."""
    nc = NX * NY * NZ
    return 4.0 / 3.0 * nc * ((pre + post) * 8 * s + 8 * s + 3.5 * s)


def _askervein_grid():
    """extern_erf/askervein.txt (the reference's terrain grid), as the test
    fixture holds it."""
    import numpy as np
    from ablsteadyrans_amd import mesh as M
    d = np.load(os.path.join(ROOT, "tests", "golden", "askervein_grid.npz"))
    return M.ErfGrid(d["x"], d["y"], d["z"])


def _time_solves(precision, steps, warmup):
    """Whole steady solves of the column deck from its first guess: returns
    (seconds per solve, iterations, status, u*, the solves' times)."""
    import torch
    from ablsteadyrans_amd import decks as D
    deck = D.prepare(D.neutral_column(), precision)
    times, last = [], None
    for rep in range(warmup + steps):
        D.reinitialize(deck)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = deck.solve()
        torch.cuda.synchronize()
        if rep >= warmup:
            times.append(time.perf_counter() - t0)
        last = r
    return sum(times) / len(times), int(last.iterations), int(last.status), float(last.mon[2]), times


def _time_newton_solves(steps, warmup):
    """Whole steady solves of the column deck with steady.stage =
    simplec_newton (SIMPLEC, the switch at r <= 1e-4 after 20 iterations with
    the monitors settled, then pseudo-transient Newton-Krylov with the
    SIMPLEC preconditioner, GMRES(60)): seconds per solve and the counts."""
    import torch
    from ablsteadyrans_amd import decks as D
    from ablsteadyrans_amd import newton as NW
    deck = D.prepare(D.neutral_column(), 0)
    nk = NW.NewtonStage(deck.simplec, NW.NewtonOptions(gmres_restart=60, start_cfl=deck.driver.cfl_max))
    times, last = [], None
    for rep in range(warmup + steps):
        D.reinitialize(deck)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = nk.solve(deck.state, deck.driver, NW.SwitchOptions())
        torch.cuda.synchronize()
        if rep >= warmup:
            times.append(time.perf_counter() - t0)
        last = r
    sec = sum(times) / len(times)
    return {"steady_solves_per_s": round(1.0 / sec, 4), "seconds_per_solve": round(sec, 4),
            "iterations": last["iterations"], "simplec_iterations": last["simplec_iterations"],
            "newton_iterations": last["newton_iterations"], "newton_steps": last["newton_steps"],
            "first_switch": last["first_switch"], "status": last["status"],
            "relative_residual": last["relative_residual"]}


# The Gaussian-hill benchmark's solver settings: those of the reference's
# 3-D hill deck (Tests/Canonical_RANS/Neutral_Hill_3D/inputs_steady_hill3d:
# 16 z-line sweeps, CFL 1 -> 4, rotation_relax 1). With 2 sweeps and CFL 10
# the hill's wake stalls the relative residual near 0.4.
C2_SETTINGS = dict(sweeps=16, cfl_max=4.0, rotation_relax=1.0)
C2_WORKLOAD = ("steady solve to a 6-order residual drop (residual_tol 5e-5, monitors flat over 50) of the "
               "Gaussian hill h = 100 exp(-(r/250 m)^2) m centred in 4 x 4 km, 200x200x64 cells (dx = dy = 20 m, "
               "dz1 = 1 m growing by 1.08, top 1709.5 m, BTF); neutral ABL: log-law inflow (u* 0.4 m/s, z0 "
               "0.03 m) at xlo, outflow at xhi, y periodic, surface layer z0 0.03 m, slip top, Coriolis f = "
               "1e-4 1/s, k-L closure; 16 z-line Jacobi sweeps, CFL 1 -> 4, rotation_relax 1, pressure reltol "
               "1e-3 (GMRES + x/y semi-coarsening multigrid)")


def _c2_run(precision, steps, warmup):
    """The Gaussian-hill steady solve at one precision (0 FP64; 2 mixed:
    FP32 sweeps and V-cycles inside the FP64 defect correction): warmup
    driver iterations from the first guess, then `steps` timed iterations
    from the first guess (ms per step), then one whole steady solve from
    the first guess timed end to end (the driver loop in the library).
    Returns a dict and the converged fields (host arrays)."""
    import torch
    from ablsteadyrans_amd import decks as D
    from ablsteadyrans_amd import physics as GP
    deck = D.prepare(D.gaussian_hill_200(**C2_SETTINGS), precision)
    full = deck.driver.max_iterations
    if warmup > 0:
        deck.driver.max_iterations = warmup
        deck.solve()
        D.reinitialize(deck)
    deck.driver.max_iterations = steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    deck.solve()
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / steps
    D.reinitialize(deck)
    deck.driver.max_iterations = full
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = deck.solve()
    torch.cuda.synchronize()
    solve_s = time.perf_counter() - t0
    fields = {n: deck.state[n].numpy()[0, 2:-2, 2:-2, 2:-2].copy() for n in GP.FIELDS}
    out = {"iterations": int(r.iterations), "status": int(r.status), "seconds_per_solve": round(solve_s, 3),
           "steady_solves_per_s": round(1.0 / solve_s, 5), "ms_per_step": round(ms_step, 3),
           "ms_per_iteration_whole_solve": round(solve_s * 1e3 / max(int(r.iterations), 1), 3),
           "relative_residual": float(r.relative_residual), "mean_ustar": round(float(r.mon[2]), 6)}
    del deck
    torch.cuda.empty_cache()
    return out, fields


def _c2_sequenced_run(precision, levels=2):
    """The Gaussian-hill steady solve with grid sequencing: the coarse
    level(s) from the initializer to the loose 2-order stop, then level 0
    (the 200x200x64 tile) from the prolonged state to the strict 6-order
    criterion seeded with the coarsest level's references. Timed end to end
    (hierarchy prepared before the timed region). Returns a dict and level
    0's converged fields."""
    import torch
    from ablsteadyrans_amd import decks as D
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd import sequencing as SQ
    deck = D.gaussian_hill_200(**C2_SETTINGS)
    sd = SQ.prepare_sequenced(deck, levels, precision=precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res, summ = SQ.solve_sequenced(sd, SQ.SequencingOptions(enabled=True))
    torch.cuda.synchronize()
    solve_s = time.perf_counter() - t0
    fields = {n: deck.state[n].numpy()[0, 2:-2, 2:-2, 2:-2].copy() for n in GP.FIELDS}
    out = {"levels": levels, "status": int(res.driver.status), "seconds_per_solve": round(solve_s, 3),
           "steady_solves_per_s": round(1.0 / solve_s, 5),
           "relative_residual": float(res.driver.relative_residual),
           "per_level": [{"level": x.level, "cells": [x.cells_x, x.cells_y], "iterations": x.iterations,
                          "seconds": round(x.wall_seconds, 3), "stop": x.stop} for x in summ]}
    del sd, deck
    torch.cuda.empty_cache()
    return out, fields


def _hill3d_sequenced():
    """The reference's Neutral_Hill_3D deck with three levels of grid
    sequencing (docs/results.rst:658-676: 54 s on four ranks, 706 level-0
    iterations after 383 and 51)."""
    import torch
    from ablsteadyrans_amd import decks as D
    from ablsteadyrans_amd import sequencing as SQ
    deck = D.neutral_hill_3d()
    sd = SQ.prepare_sequenced(deck, 3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res, summ = SQ.solve_sequenced(sd, SQ.SequencingOptions(enabled=True))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"deck": "Neutral_Hill_3D, 3 levels of grid sequencing", "status": int(res.driver.status),
            "iterations_per_level": [x.iterations for x in summ], "iterations_reference": [383, 51, 706],
            "seconds": round(dt, 3), "seconds_reference": 54.0, "reference_ranks": 4,
            "speedup_vs_reference": round(54.0 / dt, 2)}


def _field_errors(a, b):
    """Per-field relative L2 of a against b (valid points)."""
    import numpy as np
    out = {}
    for n in a:
        den = float(np.linalg.norm(b[n]))
        out[n] = float(np.linalg.norm(a[n] - b[n]) / den) if den > 0 else float(np.linalg.norm(a[n]))
    return out


def _time_reference_decks():
    """The reference's own steady decks with published CPU times (Askervein:
    175 iterations, 2 s on two ranks; Neutral_ABL_Flat: about 2850
    iterations, 13 s on two ranks; Neutral_Hill_3D: 2092 iterations, 148 s
    on four ranks; the same with line LU-SGS: 823 iterations, 51 s on four
    ranks, 860 on one box), each from its initializer state."""
    import torch
    from ablsteadyrans_amd import decks as D
    out = []
    for deck in (D.askervein(_askervein_grid()), D.neutral_abl_flat(), D.neutral_hill_3d(),
                 D.neutral_hill_3d_lusgs()):
        D.prepare(deck)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = deck.solve()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out.append({"deck": deck.name, "cells": list(deck.tile.geom.n), "status": int(r.status),
                    "iterations": int(r.iterations), "seconds": round(dt, 3),
                    "iterations_reference": deck.iterations_reference, "seconds_reference": deck.seconds_reference,
                    "reference_ranks": deck.ranks_reference,
                    "speedup_vs_reference": round(deck.seconds_reference / dt, 2),
                    "ms_per_iteration": round(dt * 1e3 / max(int(r.iterations), 1), 3)})
    return out


def _profile_deck(name, iterations):
    """iterations driver iterations of one reference deck after 10 warm-up
    iterations (the state back at the first guess in between)."""
    import torch
    from ablsteadyrans_amd import decks as D
    make = {"column": D.neutral_column, "askervein": lambda: D.askervein(_askervein_grid()),
            "abl": D.neutral_abl_flat, "hill3d": D.neutral_hill_3d, "hill3d_lusgs": D.neutral_hill_3d_lusgs}[name]
    deck = D.prepare(make())
    deck.driver.max_iterations = 10
    deck.solve()
    D.reinitialize(deck)
    deck.driver.max_iterations = iterations
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = deck.solve()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"profile_deck": deck.name, "iterations": int(r.iterations), "seconds": round(dt, 4),
            "ms_per_iteration": round(dt * 1e3 / max(int(r.iterations), 1), 4)}


def _serving_leg(n=16):
    """The serving configuration: n requests arriving as a Poisson process at
    20 per second (exponential gaps, seed 20265203), each a tile of the
    batched distribution (serving.random_tile: 32..128 squared, 32..80
    levels from 1 m by 1.08, 1-5 Gaussian hills, a random wind direction),
    half of them (alternating) the mass-consistent wind and half a converged
    steady field at a 6-order residual drop (at most 6000 iterations),
    served with 8 slots in the mixed-precision form, admitted by HBM budget
    (75 % of the free HBM, each tile charged its cells times the measured
    bytes per cell); then the same requests one after another, each alone
    on one stream in FP64 (the plain per-request path). Returns both
    summaries and the measured bytes per cell."""
    import numpy as np
    from ablsteadyrans_amd import serving as S
    rng = np.random.default_rng(20265203)
    arrivals = S.poisson_arrivals(n, 20.0, rng)
    reqs = []
    for i in range(n):
        spec = S.random_tile(np.random.default_rng(20265203 + 1 + i))
        reqs.append(S.Request(i, "steady" if i % 2 else "init", spec, arrival=float(arrivals[i]), precision=2,
                              max_iterations=6000))
    server = S.Server.measured(slots=8)
    served, wall = server.run(reqs)
    plain = [S.Request(r.rid, r.kind, r.spec, 0.0, 0, r.max_iterations) for r in reqs]
    seq, wall_seq = S.Server(slots=1).run(plain)
    return S.summary(served, wall), S.summary(seq, wall_seq), round(server.bytes_per_cell, 1)


def _batched_leg(iterations=20):
    """64 independent tiles of the batched distribution (seeds 20264203 +
    tile) submitted at once, each prepared (mesh on the device, the
    mass-consistent first guess) and advanced `iterations` mixed-precision
    SIMPLEC iterations: 8 concurrent streams vs one stream; tiles per
    second and cell-iterations per second."""
    import numpy as np
    from ablsteadyrans_amd import serving as S
    reqs = [S.Request(i, "steady", S.random_tile(np.random.default_rng(20264203 + i)), precision=2,
                      max_iterations=iterations) for i in range(64)]
    res8, wall8 = S.Server(slots=8).run(reqs)
    res1, wall1 = S.Server(slots=1).run(reqs)
    cells = sum(r.spec.cells for r in reqs)
    return {"tiles": 64, "cells": cells, "iterations_per_tile": iterations,
            "errors": sum(1 for r in res8 + res1 if r.error),
            "tiles_per_s_8_streams": round(64 / wall8, 3), "tiles_per_s_1_stream": round(64 / wall1, 3),
            "cell_iterations_per_s_8_streams": round(cells * iterations / wall8, 1),
            "speedup": round(wall1 / wall8, 3)}


def _cpu_column_baseline(target_s=15.0):
    """The oracle's SIMPLEC iteration composite on the column deck (FP64 C
    pieces, an exact sparse pressure solve; one host thread), iterated from
    the deck's first guess at local dtau with CFL 1 for about target_s;
    reported as steady solves/s with the deck's 2231 iterations."""
    import numpy as np
    from oracle import oracle as orc
    from oracle import physics as P
    from tests import physics_helpers as PHY
    from tests import stage_helpers as STH
    from ablsteadyrans_amd import decks as D
    tile = D.neutral_column().tile
    m = PHY.omesh(tile)
    g = m.geom
    q = PHY.zeros_state(tile.geom)
    _, _, zu = PHY.positions(tile, 1)
    q["u"][0] = 0.25 / 0.41 * np.log((np.maximum(zu, 0.0) + 0.1) / 0.1)
    q["theta"][:] = 300.0
    q["tke"][:] = 0.25 * 0.25 / (0.5562 * 0.5562)
    phys = P.Physics(closure=P.ClosureOptions(max_geom_lscale=1.0e12),
                     bc=P.Boundary(zlo=P.SURFACE_LAYER, zhi=P.SLIP_WALL),
                     surface=P.SurfaceOptions(z0=0.1), forcing=P.Forcing(pressure_gradient=(1.225e-4, 0.0, 0.0)),
                     log_face_mean=True)
    iters = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < target_s:
        q1 = {n: a[:, 1:-1, 1:-1, 1:-1].copy() for n, a in q.items()}
        dtau = orc.local_dtau(q1["u"], q1["v"], q1["w"], (-1, -1, -1), m.z, (-1, -1, -1), m.J, (-1, -1, -1), g, 1.0)
        STH.simplec_iteration(tile, m, q, phys, dtau, sweeps=2)
        iters += 1
    dt = time.perf_counter() - t0
    per_it = dt / iters
    return {"value": round(1.0 / (per_it * COLUMN_ITERATIONS_REF), 5), "unit": "steady solves/s", "cores": 1,
            "kind": "port",
            "sample": f"the oracle's SIMPLEC iteration composite on the Neutral_Column deck (4x4x64), {iters} "
                      f"iterations in {dt:.1f} s = {per_it * 1e3:.1f} ms per iteration, 1 thread, scaled by the "
                      f"deck's {COLUMN_ITERATIONS_REF} iterations to steady",
            "ms_per_iteration": round(per_it * 1e3, 3)}


def _cpu_info():
    """(CPU model, logical CPUs) of the host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def _cpu_c2_baseline(n=99, fp64_iterations=None):
    """The oracle's SIMPLEC iteration composite (the C oracle's FP64 pieces,
    numpy in between; the pressure correction's Poisson solve by GMRES with
    a z-line preconditioner on the oracle operator at the deck's reltol
    1e-3, tests/cpu_poisson.py) on one host thread, one iteration from the
    first guess on an n x n x 64 tile of the Gaussian-hill workload (the same
    dx, levels, hill and physics, the hill centred); reported as steady
    solves/s of the 200 x 200 x 64 solve: the iteration time scaled by the
    cell ratio and by the plain FP64 GPU solve's iteration count."""
    from ablsteadyrans_amd import decks as D
    from ablsteadyrans_amd import mesh as M
    from oracle import initializer as OI
    from oracle import oracle as orc
    from tests import config_helpers as CH
    from tests import cpu_poisson as CP
    from tests import physics_helpers as PHY
    from tests import stage_helpers as STH
    deck = D.open_abl(f"gaussian_hill_{n}x{n}x64", (n, n, 64), (20.0 * n, 20.0 * n), 1.0,
                      M.Terrain(kind="hills", hills=[(10.0 * n, 10.0 * n, 100.0, 250.0)]), **C2_SETTINGS)
    tile = deck.tile
    m = PHY.omesh(tile)
    phys = CH.open_abl_physics(deck)
    q = PHY.zeros_state(PHY.P_geom(tile))
    gi = deck.inputs.inflow
    OI.first_guess_state(q, m, OI.InflowProfile(kind=OI.LOG_LAW, ustar=gi.ustar, z0=gi.z0, direction=gi.direction),
                         OI.InitOptions(first_guess=1))
    solve = CP.LineGMRES(tile, rtol=deck.stage.pressure_reltol)
    B1 = (-1, -1, -1)
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1):   # numpy's BLAS on one thread too
        t0 = time.perf_counter()
        q1 = {k: a[:, 1:-1, 1:-1, 1:-1].copy() for k, a in q.items()}
        dtau = orc.local_dtau(q1["u"], q1["v"], q1["w"], B1, m.z, B1, m.J, B1, m.geom, 1.0)
        STH.simplec_iteration(tile, m, q, phys, dtau, sweeps=C2_SETTINGS["sweeps"], tke_floor=0.0,
                              relax=C2_SETTINGS["rotation_relax"], solve=solve)
        dt = time.perf_counter() - t0
    scale = (200.0 * 200.0) / (n * n)
    its = fp64_iterations or 868
    per_it = dt * scale
    model, ncpu = _cpu_info()
    return {"value": float(f"{1.0 / (per_it * its):.4g}"), "unit": "steady solves/s", "cores": 1, "kind": "port",
            "sample": f"one SIMPLEC iteration of the oracle's composite (C pieces, numpy between them, GMRES with "
                      f"a z-line preconditioner for the pressure: {solve.iterations} iterations) on a {n}x{n}x64 "
                      f"tile of the Gaussian-hill workload in {dt:.1f} s on 1 thread, scaled by the cell ratio "
                      f"{scale:g} and the {its} iterations of the plain FP64 solve; host CPU {model}, "
                      f"{ncpu} logical CPUs",
            "s_per_iteration_200x200x64": round(per_it, 2), "cpu_model": model, "host_logical_cpus": ncpu,
            "reference_published": _reference_cpu_rate(its)}


def _reference_cpu_rate(its):
    """The reference's own CPU throughput as its documentation publishes it
    (docs/results.rst:200-201: Neutral_Hill_3D, 64 x 64 x 20 cells, 16
    z-line Jacobi sweeps, 2092 iterations in 148 s on four ranks), in cell
    iterations per second per rank, and what it implies for the 200 x 200 x
    64 solve of `its` iterations on one rank and on this host's threads
    (perfect scaling assumed: an upper bound for the reference)."""
    rate = 2092 * 64 * 64 * 20 / (148.0 * 4)
    cells = 200 * 200 * 64
    threads = _cpu_threads()
    return {"cell_iterations_per_s_per_rank": round(rate), "source": "docs/results.rst:200-201",
            "c2_steady_solves_per_s_1_rank": float(f"{rate / (cells * its):.4g}"),
            f"c2_steady_solves_per_s_{threads}_ranks": float(f"{rate * threads / (cells * its):.4g}")}


def _traffic_from_profile():
    """HBM bytes per launch of the FP64 sweep from the committed counter
    summary (its 'hbm_bytes_per_launch:' line), or None."""
    p = os.path.join(ROOT, "profiles", "zjacobi_sweep_fractal_256x256x96_pmc.md")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed SIMPLEC iterations of the Gaussian-hill steady solve per precision")
    ap.add_argument("--warmup", type=int, default=5, help="untimed iterations before them")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving and batched legs")
    ap.add_argument("--profile-only", action="store_true", help="sweeps only (for rocprofv3 runs)")
    ap.add_argument("--profile-deck", choices=("column", "askervein", "abl", "hill3d", "hill3d_lusgs", "c2", "c2mixed"),
                    help="only N driver iterations of one reference deck (for rocprofv3 runs)")
    ap.add_argument("--iterations", type=int, default=100, help="driver iterations of --profile-deck")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only a plain 1 GiB device copy, repeated (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (a tile is solved on one device); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import ablsteadyrans_amd as G

    torch.cuda.set_device(0)
    sweep_steps, sweep_warmup = 200, 20
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(sweep_steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8,
                          "launches": sweep_steps}))
        return
    if args.profile_deck in ("c2", "c2mixed"):
        from ablsteadyrans_amd import decks as D
        deck = D.prepare(D.gaussian_hill_200(**C2_SETTINGS), 2 if args.profile_deck == "c2mixed" else 0)
        deck.driver.max_iterations = 10
        deck.solve()
        D.reinitialize(deck)
        deck.driver.max_iterations = args.iterations
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = deck.solve()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        print(json.dumps({"profile_deck": deck.name, "precision": args.profile_deck, "iterations": int(r.iterations),
                          "ms_per_iteration": round(dt * 1e3 / max(int(r.iterations), 1), 4)}))
        return
    if args.profile_deck:
        print(json.dumps(_profile_deck(args.profile_deck, args.iterations)))
        return
    if args.profile_only:
        # the fixed rule (no autotuning): every profiling pass launches the
        # same kernels (the register kernel, c' split for FP64, in registers
        # for FP32)
        import importlib
        importlib.import_module("ablsteadyrans_amd.line_sweeps").set_jacobi_schedule(3)
        sys64 = _systems(torch.float64, 20264203)
        plan64 = G.JacobiPlan(sys64)
        ms_ev, _ = _time_sweeps(plan64, sweep_steps, sweep_warmup)
        sys32 = _systems(torch.float32, 20264203)
        plan32 = G.JacobiPlan(sys32)
        _time_sweeps(plan32, sweep_steps, sweep_warmup)
        del plan64, plan32, sys64, sys32
        tile = _fractal_tile()
        v64, _ = _time_vcycle(tile, 0, 3)
        v32, _ = _time_vcycle(tile, 1, 3)
        print(json.dumps({"profile_only": True, "ms_per_sweep_fp64": ms_ev, "vcycle_ms": [v64, v32]}))
        return

    # 1. the headline: the Gaussian-hill steady solve, mixed and FP64
    print("bench: steady solve of the 200x200x64 Gaussian hill (mixed)", file=sys.stderr, flush=True)
    c2m, fm = _c2_run(2, args.steps, args.warmup)
    print("bench: steady solve of the 200x200x64 Gaussian hill (FP64)", file=sys.stderr, flush=True)
    c2d, fd = _c2_run(0, args.steps, args.warmup)
    print("bench: steady solve of the 200x200x64 Gaussian hill (mixed, two levels of grid sequencing)",
          file=sys.stderr, flush=True)
    c2s, fs = _c2_sequenced_run(2)
    import importlib
    _LS = importlib.import_module("ablsteadyrans_amd.line_sweeps")   # the package exports a function of that name
    tuned = [r for r in _LS.jacobi_tuned() if r["systems"] >= 2][:8]
    errs = _field_errors(fm, fd)
    errs_seq = _field_errors(fs, fd)
    del fm, fd, fs
    # 2. the reference's column deck, whole solves (the previous headline)
    print("bench: steady solves of the Neutral_Column deck (FP64, mixed)", file=sys.stderr, flush=True)
    s64, it64, st64, us64, t64 = _time_solves(0, 3, 1)
    s32, it32, st32, us32, t32 = _time_solves(2, 3, 1)
    newton = _time_newton_solves(3, 1)
    # 3. the reference's own steady decks with published CPU times
    print("bench: the reference's steady decks (Askervein, Neutral_ABL_Flat, Neutral_Hill_3D)", file=sys.stderr,
          flush=True)
    decks = _time_reference_decks()
    decks.append(_hill3d_sequenced())
    # 4. the z-line sweep (the roofline kernel) and the V-cycle on the fractal tile
    print("bench: z-line sweeps and V-cycles on the 256x256x96 fractal tile", file=sys.stderr, flush=True)
    # the fixed schedule rule for the roofline legs: the kernels the committed
    # profiles measured (the autotuner picks among candidates within ~1 %)
    _LS.set_jacobi_schedule(3)
    sys64 = _systems(torch.float64, 20264203)
    plan64 = G.JacobiPlan(sys64)
    unknowns = plan64.unknowns
    bytes64 = REALS_PER_UNKNOWN * 8 * unknowns
    ms_ev, _ = _time_sweeps(plan64, sweep_steps, sweep_warmup)
    kernel64 = _rule_kernel("f64")
    run64 = _time_run(plan64, 2, 20)
    del plan64, sys64
    tile = _fractal_tile()
    vc64, levels = _time_vcycle(tile, 0, 10)
    vc32, _ = _time_vcycle(tile, 1, 10)
    sys32 = _systems(torch.float32, 20264203)
    plan32 = G.JacobiPlan(sys32)
    ms32, _ = _time_sweeps(plan32, sweep_steps, sweep_warmup)
    kernel32 = _rule_kernel("f32")
    run32 = _time_run(plan32, 2, 20)
    _LS.set_jacobi_schedule(0)
    bytes32 = REALS_PER_UNKNOWN * 4 * unknowns
    gbs64 = bytes64 / (ms_ev * 1e-3) / 1e9
    gbs32 = bytes32 / (ms32 * 1e-3) / 1e9
    serving = batched = None
    if not args.no_serving:
        print("bench: serving (Poisson arrivals at 20/s) and 64 batched tiles", file=sys.stderr, flush=True)
        served, plain, bpc = _serving_leg()
        serving = {"served_mixed_8_slots": served, "plain_fp64_one_at_a_time": plain,
                   "measured_bytes_per_cell": bpc}
        batched = _batched_leg()
    cpu = cpu_sweep = cpu_column = None
    if not args.no_cpu:
        print("bench: CPU baselines (oracle)", file=sys.stderr, flush=True)
        cpu = _cpu_c2_baseline(fp64_iterations=c2d["iterations"])
        cpu_column = _cpu_column_baseline()
        cpu_sweep = _cpu_baseline()
    value = c2s["steady_solves_per_s"]
    line = {
        "metric": "steady solves/s at 6-order residual drop, mixed vs FP64; sweep HBM GB/s; p99",
        "value": value,
        "unit": "steady solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": c2m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sweeps and V-cycles in f64 defect correction",
        "data": "synthetic: the analytic Gaussian-hill terrain and the log-law first guess made divergence-free, "
                "resident in HBM",
        "config": {"workload": C2_WORKLOAD, "cells": [200, 200, 64],
                   "form": "mixed precision, two levels of grid sequencing (100x100 then 200x200)",
                   "step": "one SIMPLEC iteration on the 200x200x64 tile (residual, momentum predictor with 16 "
                           "sweeps, pressure correction, theta and k sweeps), mixed precision",
                   "iterations_per_level": [x["iterations"] for x in c2s["per_level"]], "status": c2s["status"],
                   "seconds_per_solve": c2s["seconds_per_solve"]},
        "roofline": {"bound": "hbm", "achieved": round(gbs64, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs64 / HBM_PEAK_GBS, 4), "traffic": _traffic_from_profile(),
                     "kernel": kernel64 + " (the metric's sweep HBM GB/s: 5 variables on 256x256x96, one "
                                          "launch)",
                     "algorithmic_bytes_per_launch": bytes64, "avg_launch_ms": round(ms_ev, 5)},
        "cpu_baseline": cpu,
        "extra": {"c2_mixed_sequenced": c2s, "c2_mixed": c2m, "c2_fp64": c2d,
                  "fast_vs_fp64": round(c2s["steady_solves_per_s"] / c2d["steady_solves_per_s"], 3),
                  "mixed_vs_fp64": round(c2m["steady_solves_per_s"] / c2d["steady_solves_per_s"], 3),
                  "c2_field_error_mixed_sequenced_vs_fp64_rel_l2": {k: float(f"{v:.3e}")
                                                                    for k, v in errs_seq.items()},
                  "c2_field_error_mixed_vs_fp64_rel_l2": {k: float(f"{v:.3e}") for k, v in errs.items()},
                  "column_fp64_steady_solves_per_s": round(1.0 / s64, 4), "column_fp64_iterations": it64,
                  "column_fp64_status": st64, "column_ustar": round(us64, 8),
                  "column_mixed_steady_solves_per_s": round(1.0 / s32, 4), "column_mixed_iterations": it32,
                  "column_ms_per_iteration_fp64": round(s64 * 1e3 / max(it64, 1), 3),
                  "column_simplec_newton": newton,
                  "column_simplec_newton_vs_simplec": round(newton["steady_solves_per_s"] * s64, 3),
                  "sweep_fp64_gbs": round(gbs64, 2), "sweep_fp64_ms": round(ms_ev, 5),
                  "sweep_fp32_gbs": round(gbs32, 2), "sweep_fp32_ms": round(ms32, 5),
                  "sweep_fp32_frac": round(gbs32 / HBM_PEAK_GBS, 4),
                  "sweep_fp32_kernel": kernel32,
                  "fp64_line_solve_2_sweeps_ms": round(run64, 4), "fp32_line_solve_2_sweeps_ms": round(run32, 4),
                  "vcycle_levels": levels,
                  "fp64_vcycle_ms": round(vc64, 4), "fp32_vcycle_ms": round(vc32, 4),
                  "fp64_vcycle_gbs_algorithmic": round(_vcycle_alg_bytes(8) / (vc64 * 1e-3) / 1e9, 1),
                  "fp32_vcycle_gbs_algorithmic": round(_vcycle_alg_bytes(4) / (vc32 * 1e-3) / 1e9, 1),
                  "fp64_vcycle_frac_algorithmic": round(_vcycle_alg_bytes(8) / (vc64 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                  "fp64_vcycle_gbs_moved": round(_vcycle_bytes(8) / (vc64 * 1e-3) / 1e9, 1),
                  "fp32_vcycle_gbs_moved": round(_vcycle_bytes(4) / (vc32 * 1e-3) / 1e9, 1),
                  "jacobi_autotuner_c2": tuned,
                  "reference_decks": decks,
                  "serving": serving,
                  "serving_p99_s": serving["served_mixed_8_slots"]["all"]["p99_s"] if serving else None,
                  "batched": batched,
                  "cpu_sweep_baseline": cpu_sweep,
                  "cpu_column_baseline": cpu_column},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
