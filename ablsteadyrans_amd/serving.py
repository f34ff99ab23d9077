"""Steady-wind requests served on one GPU.

A request is a terrain tile (size, Gaussian hills, wind direction) asking for
either the mass-consistent wind (the initializer alone, Source/Initializer.cpp)
or the converged steady RANS field (the initializer, then SIMPLEC iterations
until the steady criterion holds, Source/Driver.cpp:403-492). Requests arrive
over time; :class:`Server` admits each one as soon as a slot and enough of
its HBM budget are free, and runs it on its own HIP stream from its own host
thread. The library's calls release the GIL and keep their scratch per
stream, so the tiles' kernels overlap on the device while each request's
arithmetic is exactly what it is alone (the same kernels in the same order
on its stream). Per request it records the arrival, start and finish times;
:func:`summary` gives p50/p99 latency and throughput.

The batched configuration (many independent tiles at once) is the same
server with every request arriving at time zero.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import decks as D
from . import initializer as GI
from . import mesh as M
from . import physics as GP
from . import reductions as RD

# Device bytes per cell of a served tile: the state, the stage's residual,
# coefficient, flux and closure fields, the Poisson hierarchy and the
# GMRES(50) Krylov basis (about 100 FP64 vectors per cell).
BYTES_PER_CELL = 3200


@dataclass
class TileSpec:
    """A periodic terrain tile: nx x ny x nz cells of dx metres, Gaussian hills
    (xc, yc, H, L) with h = sum H exp(-r^2 / L^2), the wind from `direction`
    degrees (counter-clockwise from +x), the top at `top` metres."""
    nx: int
    ny: int
    nz: int
    dx: float = 50.0
    hills: list = field(default_factory=list)
    direction: float = 0.0
    top: float = 1000.0
    dz0: float = 0.0     # 0: uniform levels

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz


@dataclass
class Request:
    rid: int
    kind: str            # "init" or "steady"
    spec: TileSpec
    arrival: float = 0.0   # seconds after the server's start
    precision: int = 1     # 0 FP64 V-cycles, 1 FP32 V-cycles under FP64 GMRES
    max_iterations: int = 4000
    drop_orders: float = 2.0


@dataclass
class Result:
    rid: int
    kind: str
    cells: int
    arrival: float
    start: float
    finish: float
    iterations: int = 0
    status: int = 0
    divergence_after: float = 0.0
    fields: dict | None = None
    error: str | None = None

    @property
    def latency(self) -> float:
        return self.finish - self.arrival


def random_tile(rng: np.random.Generator, max_slope: float = 0.3) -> TileSpec:
    """The batched/served tile distribution: nx, ny uniform in 32..128, nz in
    32..80, 1-5 Gaussian hills with H in 20-300 m and L in 100-600 m (H at
    most max_slope L, drawn again otherwise: steeper hills make the steady
    iteration diverge on these meshes), a random wind direction in 0-360
    degrees."""
    nx, ny = int(rng.integers(32, 129)), int(rng.integers(32, 129))
    nz = int(rng.integers(32, 81))
    dx = 50.0
    hills = []
    for _ in range(int(rng.integers(1, 6))):
        while True:
            H, L = float(rng.uniform(20.0, 300.0)), float(rng.uniform(100.0, 600.0))
            if H <= max_slope * L:
                break
        hills.append((float(rng.uniform(0, nx * dx)), float(rng.uniform(0, ny * dx)), H, L))
    return TileSpec(nx, ny, nz, dx, hills, float(rng.uniform(0.0, 360.0)))


def hill_tile(rng: np.random.Generator) -> TileSpec:
    """A tile of the reference's Neutral_Hill_3D deck with a random wind
    direction and hill position: 64 x 64 x 20 cells of 80 m under an 800 m
    top, one Gaussian hill (100 m high, 500 m wide) within the middle half."""
    lx = 64 * 80.0
    xc, yc = float(rng.uniform(0.25 * lx, 0.75 * lx)), float(rng.uniform(0.25 * lx, 0.75 * lx))
    return TileSpec(64, 64, 20, 80.0, [(xc, yc, 100.0, 500.0)], float(rng.uniform(0.0, 360.0)), top=800.0)


def tile_deck(spec: TileSpec, kind: str, max_iterations: int = 4000, sweeps: int = 16, drop_orders: float = 2.0,
              ref_iterations: int = 100) -> D.Deck:
    """The deck of one request: periodic sides, the surface layer (z0 = 0.1
    m), a slip top with 0.03 K/m over the sounding (300 K to 700 m), the
    geostrophic wind of 10 m/s from the request's direction at latitude 90,
    Rayleigh damping of w and theta above 0.6875 of the top (550 m of 800 m
    in the reference's 3-D hill deck), the closure with a k
    floor of 1e-6, sixteen z-line sweeps, CFL 1 -> 4 (the settings of the
    reference's Neutral_Hill_3D deck); the first guess the log law (u* = 0.35
    m/s) along the wind made divergence-free; steady at a drop of
    drop_orders below the residual reference taken over the first
    ref_iterations iterations (the 3-D hill deck's criterion: 2 orders over
    100; residual_tol 5e-5, Source/SteadyCriterion.H:16-32)."""
    lx, ly = spec.nx * spec.dx, spec.ny * spec.dx
    zl = M.make_zlevels(spec.nz, spec.top, dz0=spec.dz0)
    top = float(zl[-1])
    lo, hi = (0.0, 0.0, 0.0), (lx, ly, top)
    geom = M.geometry((spec.nx, spec.ny, spec.nz), lo, hi, (1, 1))
    terrain = M.Terrain(kind="hills", hills=list(spec.hills)) if spec.hills else M.Terrain(kind="flat")
    tile = M.tile_frame(geom, lo, hi, zl, terrain)   # the mesh is built on the device
    a = math.radians(spec.direction)
    sounding = dict(theta_ref_z=(0.0, 700.0, 1000.0), theta_ref=(300.0, 300.0, 309.0))
    prof = GI.InflowProfile(kind=GI.LOG_LAW, ustar=0.35, z0=0.1, direction=spec.direction, **sounding)
    damping = GP.Damping(rayleigh_zdamp=0.6875 * top, rayleigh_dampcoef=0.2, fields=("w", "theta"), ztop=top,
                         prob_lo=lo[:2], prob_hi=hi[:2])
    inputs = GP.ResidualInputs(closure=GP.ClosureOptions(tke_floor=1e-6), surface=GP.SurfaceLayerOptions(z0=0.1),
                               forcing=GP.ForcingOptions(coriolis=True, coriolis_3d=False, latitude=90.0,
                                                         rotation_period=125663.7061435917, geostrophic=True,
                                                         ug=10.0 * math.cos(a), vg=10.0 * math.sin(a)),
                               zhi_theta_grad=0.03, damping=damping, **sounding)
    height = RD.residual_scale_height(tile.zlevels, 0.0, top)
    return D.Deck(f"tile{spec.nx}x{spec.ny}x{spec.nz}", tile, inputs, prof, GI.InitOptions(first_guess=1),
                  GP.SimplecOptions(sweeps=sweeps, tke_floor=1e-6, height=height, rotation_relax=1.0),
                  GP.DriverOptions(max_iterations=max_iterations, cfl0=1.0, cfl_max=4.0,
                                   residual_drop_orders=drop_orders, residual_ref_iterations=ref_iterations), 0)


def run_request(req: Request, keep_fields: bool = False, stream=None) -> Result:
    """One request on the given torch stream (default: a new one), start to
    result; the times are time.perf_counter() values."""
    import torch
    st = stream if stream is not None else torch.cuda.Stream()
    start = time.perf_counter()
    res = Result(req.rid, req.kind, req.spec.cells, req.arrival, start, start)
    try:
        with torch.cuda.stream(st):
            deck = D.prepare(tile_deck(req.spec, req.kind, req.max_iterations, drop_orders=req.drop_orders),
                              req.precision)
            if req.kind == "steady":
                r = deck.solve()
                res.iterations, res.status = int(r.iterations), int(r.status)
            if keep_fields:
                res.fields = {n: deck.state[n].numpy().copy() for n in GP.FIELDS}
            st.synchronize()
    except Exception as e:   # the request fails, the server goes on
        res.error = str(e)
    res.finish = time.perf_counter()
    return res


class Server:
    """Serves requests with `slots` concurrent tiles at most, each admitted
    when its HBM estimate (cells x BYTES_PER_CELL) fits what the running
    ones leave of `hbm_budget` bytes. Requests are taken in arrival order; a
    request that does not fit waits (later, smaller ones do not overtake
    it)."""

    def __init__(self, slots: int = 8, hbm_budget: float = 64e9):
        if slots < 1:
            raise ValueError("slots must be >= 1")
        self.slots, self.hbm_budget = int(slots), float(hbm_budget)

    def run(self, requests: list[Request], keep_fields: bool = False) -> tuple[list[Result], float]:
        """Serve the requests as they arrive (Request.arrival seconds after
        the call); returns the results (in request order) and the wall time
        from the call to the last result."""
        import torch
        for r in requests:
            if r.spec.cells * BYTES_PER_CELL > self.hbm_budget:
                raise ValueError(f"request {r.rid} needs more than the whole HBM budget")
        order = sorted(range(len(requests)), key=lambda i: (requests[i].arrival, i))
        lock = threading.Condition()
        state = {"next": 0, "in_use": 0.0}
        results: list[Result | None] = [None] * len(requests)
        t0 = time.perf_counter()
        device = torch.cuda.current_device()

        def worker():
            torch.cuda.set_device(device)
            stream = torch.cuda.Stream()
            try:
                serve(stream)
            finally:
                # the library's scratch and caches of this stream go with it
                _lib.release_stream(stream)

        def serve(stream):
            while True:
                with lock:
                    while True:
                        if state["next"] >= len(order):
                            return
                        i = order[state["next"]]
                        req = requests[i]
                        need = req.spec.cells * BYTES_PER_CELL
                        wait = t0 + req.arrival - time.perf_counter()
                        if wait <= 0 and state["in_use"] + need <= self.hbm_budget:
                            state["next"] += 1
                            state["in_use"] += need
                            break
                        lock.wait(timeout=max(min(wait, 0.05), 0.001) if wait > 0 else 0.05)
                r = run_request(req, keep_fields, stream)
                r.arrival += t0
                results[i] = r
                with lock:
                    state["in_use"] -= need
                    lock.notify_all()

        threads = [threading.Thread(target=worker, daemon=True) for _ in range(min(self.slots, len(requests)))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        wall = time.perf_counter() - t0
        return [r for r in results], wall


def poisson_arrivals(n: int, rate: float, rng: np.random.Generator) -> np.ndarray:
    """Arrival times (s) of n requests of a Poisson process of `rate` per second."""
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def percentile(values, q: float) -> float:
    v = np.sort(np.asarray(values, dtype=np.float64))
    if v.size == 0:
        return float("nan")
    return float(np.percentile(v, q))


def summary(results: list[Result], wall: float) -> dict:
    """p50/p99 latency (s) per request kind and overall, throughput, errors."""
    out = {"requests": len(results), "wall_s": round(wall, 4),
           "throughput_per_s": round(len(results) / wall, 4) if wall > 0 else None,
           "errors": sum(1 for r in results if r.error)}
    for kind in ("init", "steady", None):
        sel = [r for r in results if (kind is None or r.kind == kind) and not r.error]
        if not sel:
            continue
        lat = [r.latency for r in sel]
        key = kind or "all"
        out[key] = {"n": len(sel), "p50_s": round(percentile(lat, 50), 4), "p99_s": round(percentile(lat, 99), 4),
                    "mean_s": round(float(np.mean(lat)), 4)}
        if kind == "steady":
            out[key]["steady"] = sum(1 for r in sel if r.status == 1)
            out[key]["mean_iterations"] = round(float(np.mean([r.iterations for r in sel])), 1)
    return out
