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

# Device bytes per cell of a served tile before any measurement: the state,
# the stage's residual, coefficient, flux and closure fields, the Poisson
# hierarchy and the GMRES(50) Krylov basis (about 100 FP64 vectors per
# cell). Server measures the real figure on a calibration tile
# (measure_bytes_per_cell) and admits by that.
BYTES_PER_CELL = 3200


def measure_bytes_per_cell(spec: "TileSpec | None" = None, precision: int = 2) -> float:
    """Device memory one steady request holds, per cell: the free-memory drop
    across preparing its deck and two SIMPLEC iterations on a fresh stream
    (the library's scratch for that stream included), on an otherwise idle
    device. The stream's scratch is released afterwards."""
    import torch
    spec = spec or TileSpec(96, 96, 56, hills=[(960.0, 960.0, 60.0, 300.0)])
    st = torch.cuda.Stream()
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    with torch.cuda.stream(st):
        deck = D.prepare(tile_deck(spec, "steady", 2), precision)
        deck.solve()
        st.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        held = _lib.stream_bytes_held(st)
    del deck
    _lib.release_stream(st)
    torch.cuda.synchronize()
    used = max(free0 - free1, held)
    return used / spec.cells


@dataclass
class TileSpec:
    """A periodic terrain tile: nx x ny x nz cells of dx metres, Gaussian hills
    (xc, yc, H, L) with h = sum H exp(-r^2 / L^2), the wind from `direction`
    degrees (counter-clockwise from +x); levels from dz0 growing by
    `stretch` (the top free), or uniform up to `top` when dz0 = 0."""
    nx: int
    ny: int
    nz: int
    dx: float = 20.0
    hills: list = field(default_factory=list)
    direction: float = 0.0
    top: float = 1000.0
    dz0: float = 1.0
    stretch: float = 1.08

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def zlevels(self):
        if self.dz0 > 0:
            return M.make_zlevels(self.nz, 1.0, dz0=self.dz0, stretch_ratio=self.stretch)
        return M.make_zlevels(self.nz, self.top)


@dataclass
class Request:
    rid: int
    kind: str            # "init" or "steady"
    spec: TileSpec
    arrival: float = 0.0   # seconds after the server's start
    precision: int = 2     # 0 FP64; 1 FP32 V-cycles; 2 mixed (FP32 sweeps and V-cycles in FP64 defect correction)
    max_iterations: int = 4000
    drop_orders: float = 6.0
    ref_iterations: int = 5


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


# A hill's height at most this fraction of its width L (its steepest slope
# 0.86 H / L, about 23 degrees): spikes like H = 270 m over L = 110 m leave
# no positive friction velocity in the first guess on 20 m cells.
MAX_ASPECT = 0.5


def random_tile(rng: np.random.Generator) -> TileSpec:
    """The batched/served tile distribution: nx, ny uniform in 32..128, nz in
    32..80, dx = dy = 20 m, levels from dz1 = 1 m growing by 1.08, 1-5
    Gaussian hills with H uniform in 20-300 m (clipped to MAX_ASPECT L) and
    L in 100-600 m at uniform centres, a wind direction uniform in 0-360
    degrees. A tile's top is
    (1.08^nz - 1) / 0.08 m (134 m at 32 levels, 5.9 km at 80): when the
    hills' summed heights exceed half of it they are scaled down together to
    that, so the terrain-following levels stay ordered."""
    nx, ny = int(rng.integers(32, 129)), int(rng.integers(32, 129))
    nz = int(rng.integers(32, 81))
    dx = 20.0
    hills = []
    for _ in range(int(rng.integers(1, 6))):
        H, L = float(rng.uniform(20.0, 300.0)), float(rng.uniform(100.0, 600.0))
        H = min(H, MAX_ASPECT * L)
        hills.append([float(rng.uniform(0, nx * dx)), float(rng.uniform(0, ny * dx)), H, L])
    spec = TileSpec(nx, ny, nz, dx, [], float(rng.uniform(0.0, 360.0)))
    top = float(spec.zlevels()[-1])
    total = sum(h[2] for h in hills)
    scale = min(1.0, 0.5 * top / total)
    spec.hills = [(xc, yc, H * scale, L) for xc, yc, H, L in hills]
    return spec


def tile_deck(spec: TileSpec, kind: str, max_iterations: int = 4000, sweeps: int = 16, drop_orders: float = 6.0,
              ref_iterations: int = 5) -> D.Deck:
    """The deck of one request: the open-channel neutral ABL of the
    benchmark configurations (decks.open_abl) over the tile's Gaussian hills
    -- the log-law inflow (u* = 0.4 m/s, z0 = 0.03 m) from the request's
    direction on the side the wind enters, outflow opposite, the other pair
    periodic, the surface layer, a slip top, Coriolis (f = 1e-4 1/s), the
    k-L closure -- with the solver settings of the reference's 3-D hill
    deck (sixteen z-line sweeps, CFL 1 -> 4, rotation_relax 1); steady at a
    drop of drop_orders below the residual reference of the first
    ref_iterations iterations (residual_tol 5e-5, Source/SteadyCriterion.H:
    16-32). The first guess is the inflow profile made divergence-free."""
    terrain = M.Terrain(kind="hills", hills=list(spec.hills)) if spec.hills else M.Terrain(kind="flat")
    return D.open_abl(f"tile{spec.nx}x{spec.ny}x{spec.nz}", (spec.nx, spec.ny, spec.nz),
                      (spec.nx * spec.dx, spec.ny * spec.dx), spec.dz0, terrain, sweeps=sweeps, cfl_max=4.0,
                      drop_orders=drop_orders, max_iterations=max_iterations, rotation_relax=1.0,
                      ref_iterations=ref_iterations, direction=spec.direction, zlevels=spec.zlevels())


def run_request(req: Request, keep_fields: bool = False, stream=None) -> Result:
    """One request on the given torch stream (default: a new one), start to
    result; the times are time.perf_counter() values."""
    import torch
    st = stream if stream is not None else torch.cuda.Stream()
    start = time.perf_counter()
    res = Result(req.rid, req.kind, req.spec.cells, req.arrival, start, start)
    try:
        with torch.cuda.stream(st):
            deck = D.prepare(tile_deck(req.spec, req.kind, req.max_iterations, drop_orders=req.drop_orders,
                                       ref_iterations=req.ref_iterations), req.precision)
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
    when its HBM estimate (cells x bytes_per_cell) fits what the running
    ones leave of `hbm_budget` bytes. Requests are taken in arrival order; a
    request that does not fit waits (later, smaller ones do not overtake
    it)."""

    def __init__(self, slots: int = 8, hbm_budget: float = 64e9, bytes_per_cell: float | None = None):
        if slots < 1:
            raise ValueError("slots must be >= 1")
        self.slots, self.hbm_budget = int(slots), float(hbm_budget)
        self.bytes_per_cell = float(bytes_per_cell) if bytes_per_cell else BYTES_PER_CELL

    @classmethod
    def measured(cls, slots: int = 8, budget_fraction: float = 0.75, margin: float = 1.15) -> "Server":
        """A server whose budget is budget_fraction of the free HBM and whose
        admission charges each tile the measured bytes per cell (times a
        margin for the scratch that grows with the largest tile a stream
        serves)."""
        import torch
        bpc = measure_bytes_per_cell() * margin
        free, _ = torch.cuda.mem_get_info()
        return cls(slots, budget_fraction * free, bpc)

    def run(self, requests: list[Request], keep_fields: bool = False) -> tuple[list[Result], float]:
        """Serve the requests as they arrive (Request.arrival seconds after
        the call); returns the results (in request order) and the wall time
        from the call to the last result."""
        import torch
        for r in requests:
            if r.spec.cells * self.bytes_per_cell > self.hbm_budget:
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
                        need = req.spec.cells * self.bytes_per_cell
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
