#!/usr/bin/env python3
"""Benchmark of the SIMPLEC hot path on one MI355X.

Headline (value): steady solves per second at a 6-order residual drop, FP64,
on the reference's Neutral_Column deck (Tests/Canonical_RANS/Neutral_Column/
inputs_steady_column: 4 x 4 x 64 cells over 2560 x 2560 x 1000 m, periodic
sides, surface layer z0 = 0.1 m, slip top, the driving force F = 1.225e-4
m/s^2, the uncapped closure with the logarithmic face mean, two z-line
sweeps, CFL ramp to 50, steady at residual_tol 1e-8 and a 6-order drop; the
reference measured 2231 iterations). One step = one whole steady solve from
the deck's first guess (the log law with u* = 0.25 m/s and its equilibrium
k), the driver loop on the device (gust_steady_solve); the state is resident
in HBM before the timed region. The mixed form (FP32 multigrid V-cycles
under the FP64 GMRES of every pressure correction) is timed the same way.

Also measured (extra): the reference's own steady decks with published CPU
times (Askervein, Neutral_ABL_Flat, Neutral_Hill_3D: iterations to the
verdict and wall time, beside the reference's); one SIMPLEC iteration on the
Gaussian-hill tile
(200 x 200 x 64, dz1 = 1 m, FP64 and mixed); the z-line Jacobi sweep of the
five variables on the fractal-terrain tile (256 x 256 x 96, 65,536 columns;
the roofline object: algorithmic bytes per launch / launch time from HIP
events on the launch stream, 10 reals per unknown); the V(2,2) cycle.
Serving (extra.serving, extra.serving_p99_s): 14 requests arriving at 20 per
second (mass-consistent winds on random terrain tiles, every seventh a
converged steady field on a tile of the reference's 3-D hill deck) served
with 8 concurrent streams in mixed precision, beside the same requests one
at a time in FP64; p50/p99 latency from arrival to result. Batched
(extra.batched): 64 random tiles at once, 8 streams vs one.
cpu_baseline: the oracle's SIMPLEC iteration composite (FP64 C pieces and an
exact sparse pressure solve, one host thread) on the same column deck for a
bounded number of iterations, scaled by the deck's 2231 iterations.
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


def _cpu_baseline(target_s=12.0):
    """The oracle's zline_jacobi (2 sweeps per call, one thread) over the five
    variables of a 128 x 128 x 96 sample; reported per sweep in the metric's
    unit (algorithmic GB/s)."""
    import numpy as np
    from oracle import oracle as orc
    nx, ny, nz = 128, 128, NZ
    g = orc.geom_c((0, 0, 0), (nx - 1, ny - 1, nz - 1), (0, 0, 0))
    rng = np.random.default_rng(20264203)
    data = []
    unknowns = 0
    for _name, loc in VARS:
        sx, sy, sz = nx + (loc & 1), ny + ((loc >> 1) & 1), nz + ((loc >> 2) & 1)
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
    calls = 0
    t0 = time.perf_counter()
    while True:
        for loc, A, ap, rhs, d in data:
            orc.zline_jacobi(loc, A, (0, 0, 0), ap, rhs, (0, 0, 0), 2, g, d, (-1, -1, -1))
        calls += 1
        if time.perf_counter() - t0 >= target_s:
            break
    dt = time.perf_counter() - t0
    per_sweep = dt / (2 * calls)
    gbs = REALS_PER_UNKNOWN * 8 * unknowns / per_sweep / 1e9
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"oracle zline_jacobi FP64, 5 variables on 128x128x96 (same nz, {unknowns} unknowns per sweep), "
                      f"{2 * calls} sweeps in {dt:.1f} s = {per_sweep * 1e3:.1f} ms per sweep, 1 thread",
            "ms_per_sweep": round(per_sweep * 1e3, 3)}


def _fractal_tile():
    """The fractal-terrain tile: 256 x 256 x 96 cells at dx = 25 m, a seeded
    k^-2 random surface (rms 80 m, clipped to 0..400 m), first layer 2 m with
    the ratio solved so the top is 2000 m, BTF."""
    from ablsteadyrans_amd import mesh as M
    lo, hi = (0.0, 0.0, 0.0), (NX * 25.0, NY * 25.0, 2000.0)
    geom = M.geometry((NX, NY, NZ), lo, hi)
    zl = M.make_zlevels(NZ, 2000.0, dz0=2.0, stretch_ratio=0.0)
    h = M.fractal_heights(NX, NY, seed=20264203, rms=80.0, slope=-2.0, clip=(0.0, 400.0))
    return M.build_mesh(geom, lo, hi, zl, M.Terrain(kind="fractal", heights=h))


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
    """Cells per multigrid level: x/y halved while even and >= 4, until
    nx*ny <= 16 (or 12 levels)."""
    cells = []
    for lev in range(64):
        cells.append(nx * ny * nz)
        cx = 2 if (nx % 2 == 0 and nx >= 4) else 1
        cy = 2 if (ny % 2 == 0 and ny >= 4) else 1
        if (cx == 1 and cy == 1) or lev >= 12 or nx * ny <= 16:
            break
        nx //= cx
        ny //= cy
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


def _time_reference_decks():
    """The reference's own steady decks with published CPU times (Askervein:
    175 iterations, 2 s on two ranks; Neutral_ABL_Flat: about 2850
    iterations, 13 s on two ranks; Neutral_Hill_3D: 2092 iterations, 148 s
    on four ranks), each from its initializer state."""
    import torch
    from ablsteadyrans_amd import decks as D
    out = []
    for deck in (D.askervein(_askervein_grid()), D.neutral_abl_flat(), D.neutral_hill_3d()):
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
            "abl": D.neutral_abl_flat, "hill3d": D.neutral_hill_3d}[name]
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


def _serving_leg():
    """The serving configuration: 14 requests arriving as a Poisson process
    at 20 per second, every seventh a converged steady field on a tile of
    the reference's 3-D hill deck (random wind direction and hill position;
    the deck's criterion), the others the mass-consistent wind on a random
    terrain tile (32..128 squared, 32..80 levels, 1-5 Gaussian hills),
    served with 8 slots in mixed precision; then the same requests one after
    another, each alone on one stream in FP64 (the plain per-request path).
    Returns both summaries."""
    import numpy as np
    from ablsteadyrans_amd import serving as S
    rng = np.random.default_rng(20261203 + 5000)
    arrivals = S.poisson_arrivals(14, 20.0, rng)
    reqs = []
    for i in range(14):
        spec = S.hill_tile(rng) if i % 7 == 6 else S.random_tile(rng)
        reqs.append(S.Request(i, "steady" if i % 7 == 6 else "init", spec, arrival=float(arrivals[i]),
                              precision=1, max_iterations=6000))
    served, wall = S.Server(slots=8).run(reqs)
    plain = [S.Request(r.rid, r.kind, r.spec, 0.0, 0, r.max_iterations) for r in reqs]
    seq, wall_seq = S.Server(slots=1).run(plain)
    return S.summary(served, wall), S.summary(seq, wall_seq)


def _batched_leg():
    """64 independent random terrain tiles (mass-consistent wind) submitted
    at once: 8 concurrent streams vs one stream; tiles per second."""
    import numpy as np
    from ablsteadyrans_amd import serving as S
    rng = np.random.default_rng(20261203 + 4000)
    reqs = [S.Request(i, "init", S.random_tile(rng)) for i in range(64)]
    res8, wall8 = S.Server(slots=8).run(reqs)
    res1, wall1 = S.Server(slots=1).run(reqs)
    cells = sum(r.spec.cells for r in reqs)
    return {"tiles": 64, "cells": cells, "errors": sum(1 for r in res8 + res1 if r.error),
            "tiles_per_s_8_streams": round(64 / wall8, 3), "tiles_per_s_1_stream": round(64 / wall1, 3),
            "speedup": round(wall1 / wall8, 3)}


def _hill_tile():
    """The Gaussian hill h = 100 exp(-(r/250 m)^2) m centred in 4 x 4 km at
    dx = 20 m, 64 levels from dz1 = 1 m growing by 1.08, periodic sides."""
    from ablsteadyrans_amd import mesh as M
    zl = M.make_zlevels(64, 1.0, dz0=1.0, stretch_ratio=1.08)
    top = float(zl[-1])
    geom = M.geometry((200, 200, 64), (0.0, 0.0, 0.0), (4000.0, 4000.0, top), (1, 1))
    t = M.Terrain(kind="hills", hills=[(2000.0, 2000.0, 100.0, 250.0)])
    return M.build_mesh(geom, (0.0, 0.0, 0.0), (4000.0, 4000.0, top), zl, t), top


def _time_hill_iterations(precision, iters, warmup=3):
    """ms per SIMPLEC iteration on the hill tile (neutral ABL with rotation,
    the log-law first guess, local dtau at CFL 2), HIP events on the stream."""
    import torch
    from ablsteadyrans_amd import initializer as GI
    from ablsteadyrans_amd import physics as GP
    from ablsteadyrans_amd.fab import Fab
    from ablsteadyrans_amd.poisson import DeviceMesh
    tile, top = _hill_tile()
    dm = DeviceMesh(tile)
    inputs = GP.ResidualInputs(closure=GP.ClosureOptions(tke_floor=1e-6), surface=GP.SurfaceLayerOptions(z0=0.03),
                               forcing=GP.ForcingOptions(coriolis=True, coriolis_3d=False, latitude=90.0,
                                                         rotation_period=125663.7061435917, geostrophic=True,
                                                         ug=10.0))
    stage = GP.SimplecStage(dm, inputs, GP.SimplecOptions(sweeps=2, tke_floor=1e-6, height=top,
                                                          poisson_precision=precision))
    q = GP.DeviceState(tile.geom)
    GI.first_guess_state(dm, GI.InflowProfile(kind=GI.LOG_LAW, ustar=0.4, z0=0.03), GI.InitOptions(first_guess=1), q)
    dtau = Fab.at_loc(tile.geom, 0, 1, 0)
    stage.init(q)
    import ctypes as C
    from ablsteadyrans_amd import _lib
    g = tile.geom.c()

    def step():
        nu = stage.nu_view()
        _lib.check(_lib.lib.gust_local_dtau(C.byref(g), C.byref(q["u"].array4()), C.byref(q["v"].array4()),
                                            C.byref(q["w"].array4()), C.byref(dm.z_nd.array4()),
                                            C.byref(dm.detJ.array4()), C.byref(nu), 2.0, C.byref(dtau.array4()),
                                            C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        stage.iterate(q, dtau)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(iters):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / iters


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
    ap.add_argument("--steps", type=int, default=3, help="timed steady solves per precision")
    ap.add_argument("--warmup", type=int, default=1, help="untimed steady solves per precision")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving and batched legs")
    ap.add_argument("--profile-only", action="store_true", help="sweeps only (for rocprofv3 runs)")
    ap.add_argument("--profile-deck", choices=("column", "askervein", "abl", "hill3d", "hill200"),
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
    if args.profile_deck == "hill200":
        ms = _time_hill_iterations(0, args.iterations)
        print(json.dumps({"profile_deck": "gaussian_hill_200x200x64", "iterations": args.iterations,
                          "ms_per_iteration": round(ms, 4)}))
        return
    if args.profile_deck:
        print(json.dumps(_profile_deck(args.profile_deck, args.iterations)))
        return
    if args.profile_only:
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

    # 1. the headline: whole steady solves of the column deck, FP64 and mixed
    print("bench: steady solves of the Neutral_Column deck (FP64)", file=sys.stderr, flush=True)
    s64, it64, st64, us64, t64 = _time_solves(0, args.steps, args.warmup)
    print("bench: steady solves of the Neutral_Column deck (mixed)", file=sys.stderr, flush=True)
    s32, it32, st32, us32, t32 = _time_solves(1, args.steps, args.warmup)
    # 2. one SIMPLEC iteration on the hill tile
    print("bench: SIMPLEC iterations on the 200x200x64 hill tile", file=sys.stderr, flush=True)
    hill64 = _time_hill_iterations(0, 10)
    hill32 = _time_hill_iterations(1, 10)
    # 3. the reference's own steady decks with published CPU times
    print("bench: the reference's steady decks (Askervein, Neutral_ABL_Flat, Neutral_Hill_3D)", file=sys.stderr,
          flush=True)
    decks = _time_reference_decks()
    # 4. the z-line sweep (the roofline kernel) and the V-cycle on the fractal tile
    print("bench: z-line sweeps and V-cycles on the 256x256x96 fractal tile", file=sys.stderr, flush=True)
    sys64 = _systems(torch.float64, 20264203)
    plan64 = G.JacobiPlan(sys64)
    unknowns = plan64.unknowns
    bytes64 = REALS_PER_UNKNOWN * 8 * unknowns
    ms_ev, _ = _time_sweeps(plan64, sweep_steps, sweep_warmup)
    run64 = _time_run(plan64, 2, 20)
    del plan64, sys64
    tile = _fractal_tile()
    vc64, levels = _time_vcycle(tile, 0, 10)
    vc32, _ = _time_vcycle(tile, 1, 10)
    sys32 = _systems(torch.float32, 20264203)
    plan32 = G.JacobiPlan(sys32)
    ms32, _ = _time_sweeps(plan32, sweep_steps, sweep_warmup)
    run32 = _time_run(plan32, 2, 20)
    bytes32 = REALS_PER_UNKNOWN * 4 * unknowns
    gbs64 = bytes64 / (ms_ev * 1e-3) / 1e9
    gbs32 = bytes32 / (ms32 * 1e-3) / 1e9
    serving = batched = None
    if not args.no_serving:
        print("bench: serving (Poisson arrivals at 20/s) and 64 batched tiles", file=sys.stderr, flush=True)
        served, plain = _serving_leg()
        serving = {"served_mixed_8_slots": served, "plain_fp64_one_at_a_time": plain}
        batched = _batched_leg()
    cpu = cpu_sweep = None
    if not args.no_cpu:
        print("bench: CPU baselines (oracle, one thread)", file=sys.stderr, flush=True)
        cpu = _cpu_column_baseline()
        cpu_sweep = _cpu_baseline()
    value = 1.0 / s64
    line = {
        "metric": "steady solves/s at 6-order residual drop, mixed vs FP64; sweep HBM GB/s; p99",
        "value": round(value, 4),
        "unit": "steady solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s64 * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: the reference's Neutral_Column deck (its first guess), resident in HBM",
        "config": {"workload": "steady solve to a 6-order residual drop (residual_tol 1e-8) of the reference's "
                               "Neutral_Column deck: 4x4x64 cells, 2560x2560x1000 m, periodic sides, surface "
                               "layer z0 0.1 m, F = 1.225e-4 m/s^2, 2 z-line sweeps, CFL 1 -> 50",
                   "iterations": it64, "iterations_reference": COLUMN_ITERATIONS_REF, "status": st64,
                   "ustar": round(us64, 8), "seconds_per_solve": [round(t, 3) for t in t64]},
        "roofline": {"bound": "hbm", "achieved": round(gbs64, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs64 / HBM_PEAK_GBS, 4), "traffic": _traffic_from_profile(),
                     "kernel": "zjacobi_sweep_kernel<double,false,false> (the metric's sweep HBM GB/s: 5 "
                               "variables on 256x256x96, one launch)",
                     "algorithmic_bytes_per_launch": bytes64, "avg_launch_ms": round(ms_ev, 5)},
        "cpu_baseline": cpu,
        "extra": {"mixed_steady_solves_per_s": round(1.0 / s32, 4), "mixed_iterations": it32, "mixed_status": st32,
                  "mixed_ustar": round(us32, 8), "mixed_vs_fp64": round(s64 / s32, 3),
                  "ms_per_iteration_column_fp64": round(s64 * 1e3 / max(it64, 1), 3),
                  "hill_200x200x64_ms_per_iteration_fp64": round(hill64, 3),
                  "hill_200x200x64_ms_per_iteration_mixed": round(hill32, 3),
                  "sweep_fp64_gbs": round(gbs64, 2), "sweep_fp64_ms": round(ms_ev, 5),
                  "sweep_fp32_gbs": round(gbs32, 2), "sweep_fp32_ms": round(ms32, 5),
                  "sweep_fp32_frac": round(gbs32 / HBM_PEAK_GBS, 4),
                  "fp64_line_solve_2_sweeps_ms": round(run64, 4), "fp32_line_solve_2_sweeps_ms": round(run32, 4),
                  "vcycle_levels": levels,
                  "fp64_vcycle_ms": round(vc64, 4), "fp64_vcycle_gbs": round(_vcycle_bytes(8) / (vc64 * 1e-3) / 1e9, 1),
                  "fp32_vcycle_ms": round(vc32, 4), "fp32_vcycle_gbs": round(_vcycle_bytes(4) / (vc32 * 1e-3) / 1e9, 1),
                  "reference_decks": decks,
                  "serving": serving,
                  "serving_p99_s": serving["served_mixed_8_slots"]["all"]["p99_s"] if serving else None,
                  "batched": batched,
                  "cpu_sweep_baseline": cpu_sweep},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
