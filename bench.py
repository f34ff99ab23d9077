#!/usr/bin/env python3
"""Benchmark of the SIMPLEC hot path on one MI355X.

Workload (the kernel-level configuration of the path): the fractal-terrain
tile of 256 x 256 x 96 cells, 65,536 columns, the five sweep variables of an
iteration -- u on x-faces, v on y-faces, w on z-faces, theta and k on cells --
each a seeded random M-matrix line system (coupling recipe of
Tests/Unit_Tests/test_LineSweeps.cpp:76-115). One step = one z-line Jacobi
sweep over all five systems in one launch (a steady-state sweep: the input
iterate is the previous sweep's), FP64. Inputs are resident in HBM before the
timed region.

value = algorithmic HBM bytes of the step / step time, with 10 reals per
unknown (6 off-diagonal coefficients, a_P, rhs, the input and the output
iterate). The roofline object carries the same figure for the sweep kernel
timed with HIP events on its own stream; traffic comes from the committed
rocprofv3 counter summary when present. cpu_baseline: the oracle (FP64 C
restatement of the reference's zline_jacobi, one host thread) on a bounded
sample of the same workload.
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


def _traffic_from_profile():
    p = os.path.join(ROOT, "profiles", "zjacobi_sweep_c3_fp64_pmc.json")
    if os.path.exists(p):
        try:
            return json.load(open(p)).get("hbm_bytes_per_launch")
        except Exception:
            return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--profile-only", action="store_true", help="sweeps only (for rocprofv3 runs)")
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
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(max(1, args.steps)):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8,
                          "launches": max(1, args.steps)}))
        return
    sys64 = _systems(torch.float64, 20264203)
    plan64 = G.JacobiPlan(sys64)
    unknowns = plan64.unknowns
    bytes64 = REALS_PER_UNKNOWN * 8 * unknowns
    ms_ev, ms_wall = _time_sweeps(plan64, args.steps, args.warmup)
    if args.profile_only:
        sys32 = _systems(torch.float32, 20264203)
        plan32 = G.JacobiPlan(sys32)
        _time_sweeps(plan32, args.steps, args.warmup)
        del plan64, plan32, sys64, sys32
        tile = _fractal_tile()
        v64, _ = _time_vcycle(tile, 0, 3)
        v32, _ = _time_vcycle(tile, 1, 3)
        print(json.dumps({"profile_only": True, "ms_per_sweep_fp64": ms_ev, "vcycle_ms": [v64, v32]}))
        return
    run64 = _time_run(plan64, 2, 20)
    del plan64
    tile = _fractal_tile()
    vc64, levels = _time_vcycle(tile, 0, 10)
    vc32, _ = _time_vcycle(tile, 1, 10)
    sys32 = _systems(torch.float32, 20264203)
    plan32 = G.JacobiPlan(sys32)
    ms32, _ = _time_sweeps(plan32, args.steps, args.warmup)
    run32 = _time_run(plan32, 2, 20)
    bytes32 = REALS_PER_UNKNOWN * 4 * unknowns
    gbs64 = bytes64 / (ms_ev * 1e-3) / 1e9
    gbs32 = bytes32 / (ms32 * 1e-3) / 1e9
    cpu = None if args.no_cpu else _cpu_baseline()
    line = {
        "metric": "steady solves/s at 6-order residual drop, mixed vs FP64; sweep HBM GB/s; p99",
        "value": round(gbs64, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_wall, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random M-matrix line systems, resident in HBM)",
        "config": {"workload": "z-line Jacobi sweep, 5 variables (u, v, w, theta, k) on 256x256x96 "
                               "(65,536 columns x 96 levels), one batched launch per sweep",
                   "unknowns_per_sweep": unknowns, "reals_per_unknown": REALS_PER_UNKNOWN,
                   "sweep_kind": "steady-state sweep (input iterate nonzero)"},
        "roofline": {"bound": "hbm", "achieved": round(gbs64, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs64 / HBM_PEAK_GBS, 4), "traffic": _traffic_from_profile(),
                     "kernel": "zjacobi_sweep_kernel<double,false,false>",
                     "algorithmic_bytes_per_launch": bytes64, "avg_launch_ms": round(ms_ev, 5)},
        "cpu_baseline": cpu,
        "extra": {"fp32_sweep_gbs": round(gbs32, 2), "fp32_ms_per_sweep": round(ms32, 5),
                  "fp32_frac": round(gbs32 / HBM_PEAK_GBS, 4),
                  "fp64_line_solve_2_sweeps_ms": round(run64, 4), "fp32_line_solve_2_sweeps_ms": round(run32, 4),
                  "vcycle_levels": levels,
                  "fp64_vcycle_ms": round(vc64, 4), "fp64_vcycle_gbs": round(_vcycle_bytes(8) / (vc64 * 1e-3) / 1e9, 1),
                  "fp32_vcycle_ms": round(vc32, 4), "fp32_vcycle_gbs": round(_vcycle_bytes(4) / (vc32 * 1e-3) / 1e9, 1),
                  "note": "steady solves/s and p99 need the full SIMPLEC iteration (not yet on this path)"},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
