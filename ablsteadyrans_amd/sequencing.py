"""Grid sequencing over the horizontal grid hierarchy (Source/Sequencing.H,
Source/Transfer.H, Source/MeshHierarchy.H), over the library's C-ABI:

1. on the coarsest level L-1 the initializer on that level's mesh, then the
   SIMPLEC stage under its own driver until the loose stop;
2. on each level l = L-2 ... 1 the prolongation from level l+1 (bilinear per
   zeta-level, k clipped, the level's boundary conditions, the projection),
   then SIMPLEC until the loose stop;
3. on level 0 (the run's mesh) the prolongation from level 1, then the run's
   driver until the strict stop.

Every finer level starts its residual references from the coarsest level's
(seed_reference), so a drop is measured from the initial field's residual;
with ``hold = 0`` (the default) the seeded levels skip the reference hold,
and with ``cfl_previous`` they start at the previous level's CFL. The loose
stop and the input checks are the library's (``gust_loose_stop``,
``gust_check_sequencing_options``); the transfers are HIP kernels
(``gust_prolong_fields``, ``gust_restrict_state``, ``gust_prolong_state``).
The coarse meshes are built on the host (:func:`ablsteadyrans_amd.mesh.hierarchy`).
"""
from __future__ import annotations

import copy
import ctypes as C
import time
from dataclasses import dataclass, field

from . import _lib
from . import initializer as GI
from . import mesh as M
from . import physics as GP
from . import poisson as PS
from .steady import SteadyCriterion, SteadyParams

GIVEN_BITS = {"sequencing_drop_orders": 1, "sequencing_monitor_tol": 2, "sequencing_hold": 4,
              "sequencing_cfl_start": 8}


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


@dataclass
class SequencingOptions:
    """Source/Sequencing.H:85-92 (steady.sequencing*)."""

    enabled: bool = False
    drop_orders: float = 2.0
    monitor_tol: float = 0.0
    hold: int = 0
    cfl_previous: bool = False
    given: tuple = ()

    def c(self) -> _lib.SequencingOptionsC:
        mask = 0
        for k in self.given:
            mask |= GIVEN_BITS[k]
        return _lib.SequencingOptionsC(int(bool(self.enabled)), float(self.drop_orders), float(self.monitor_tol),
                                       int(self.hold), int(bool(self.cfl_previous)), mask)


def _params(strict) -> _lib.SteadyParamsC:
    if isinstance(strict, SteadyParams):
        return strict.c()
    return _lib.SteadyParamsC(strict.residual_tol, strict.residual_drop_orders, strict.residual_ref_iterations,
                              strict.monitor_rel_tol, strict.monitor_window)


def check_sequencing_options(opt: SequencingOptions, strict, levels: int, n_cell) -> list:
    """Source/Sequencing.cpp:118-176: the problems, each naming its key."""
    oc, pc = opt.c(), _params(strict)
    n = C.c_int32(0)
    buf = C.create_string_buffer(4096)
    _lib.check(_lib.lib.gust_check_sequencing_options(C.byref(oc), C.byref(pc), int(levels), int(n_cell[0]),
                                                      int(n_cell[1]), C.byref(n), buf, len(buf)))
    return buf.value.decode().split("\n") if n.value else []


def loose_stop(c: SteadyCriterion, opt: SequencingOptions, hold_iterations: int):
    """(stop, reason) after the last iteration c recorded
    (Source/Sequencing.cpp:179-234)."""
    oc = opt.c()
    stop = C.c_int32(0)
    buf = C.create_string_buffer(256)
    _lib.check(_lib.lib.gust_loose_stop(c._h, C.byref(oc), int(hold_iterations), C.byref(stop), buf, len(buf)))
    return bool(stop.value), buf.value.decode()


# --- transfers -------------------------------------------------------------------------------------------------

def prolong_fields(fine_geom, coarse_geom, ratio: int, coarse: GP.DeviceState, fine: GP.DeviceState,
                   stream=None) -> None:
    """Source/Transfer.cpp:324-352 (gust_prolong_fields)."""
    fg, cg = fine_geom.c(), coarse_geom.c()
    cc, fc = coarse.c(), fine.c()
    _lib.check(_lib.lib.gust_prolong_fields(C.byref(fg), C.byref(cg), int(ratio), C.byref(cc), C.byref(fc),
                                            _stream(stream)))


def restrict_state(fine_mesh: PS.DeviceMesh, fine: GP.DeviceState, coarse_mesh: PS.DeviceMesh, ratio: int,
                   coarse: GP.DeviceState, omega=None, stream=None) -> None:
    """Source/Transfer.cpp:263-322 (gust_restrict_state); fine's u and v need
    one ghost filled."""
    fg, cg = fine_mesh.geom.c(), coarse_mesh.geom.c()
    fm, _k1 = GP._mesh_c(fine_mesh)
    cm, _k2 = GP._mesh_c(coarse_mesh)
    fc, cc = fine.c(), coarse.c()
    om = omega.array4() if omega is not None else None
    _lib.check(_lib.lib.gust_restrict_state(C.byref(fg), C.byref(fm), C.byref(fc), C.byref(cg), C.byref(cm),
                                            int(ratio), C.byref(cc), C.byref(om) if om is not None else None,
                                            _stream(stream)))


def prolong_state(fine_stage: GP.SimplecStage, solver: PS.PoissonSolver, coarse_geom, ratio: int,
                  coarse: GP.DeviceState, fine: GP.DeviceState, tke_floor: float, stream=None) -> _lib.ProlongResultC:
    """Source/Transfer.cpp:354-403 (gust_prolong_state): the interpolation, k
    clipped at the floor, the fine stage's boundary conditions, the
    projection with solver (p restored)."""
    cg = coarse_geom.c()
    cc, fc = coarse.c(), fine.c()
    res = _lib.ProlongResultC()
    _lib.check(_lib.lib.gust_prolong_state(fine_stage._h, solver._h, C.byref(cg), int(ratio), C.byref(cc),
                                           C.byref(fc), float(tke_floor), C.byref(res), _stream(stream)))
    return res


# --- the level loop --------------------------------------------------------------------------------------------

LEVEL_MAX_ITERATIONS, LEVEL_STEADY, LEVEL_LOOSE, LEVEL_DIVERGED = 0, 1, 2, 3


def solve_level(stage: GP.SimplecStage, q: GP.DeviceState, opt: GP.DriverOptions, loose: SequencingOptions | None,
                seed_ref=None, release_hold: bool = False, start_cfl: float = 0.0, stream=None) -> _lib.LevelResultC:
    """gust_steady_solve_level: the level's seeding and start, then the steady
    loop until Steady, Diverged, the loose stop (loose given) or
    max_iterations (Source/Sequencing.cpp:361-390)."""
    lc = _lib.LevelControlC()
    keep = None
    if loose is not None:
        keep = loose.c()
        lc.loose = C.pointer(keep)
    if seed_ref is not None:
        lc.seeded = 1
        for n in range(6):
            lc.seed_ref[n] = float(seed_ref[n])
    lc.release_hold = int(bool(release_hold))
    lc.start_cfl = float(start_cfl)
    qc, oc = q.c(), opt.c()
    res = _lib.LevelResultC()
    _lib.check(_lib.lib.gust_steady_solve_level(stage._h, C.byref(qc), C.byref(oc), C.byref(lc), C.byref(res),
                                                _stream(stream)))
    del keep
    return res


@dataclass
class LevelSummary:
    """Source/Sequencing.H:302-314: what sequencing did on one level."""

    level: int
    cells_x: int
    cells_y: int
    iterations: int = 0
    wall_seconds: float = 0.0
    relative_residual: float = 0.0
    stop: str = ""
    prolonged: bool = False
    divergence_before: float = 0.0
    divergence_after: float = 0.0
    tke_clipped: int = 0


@dataclass
class Sequenced:
    """A prepared deck's hierarchy: per level (coarsest first) the deck (its
    mesh, stage and state), level 0 being the run's own."""

    levels: list = field(default_factory=list)


def _level_deck(deck, tile, precision: int):
    d = copy.copy(deck)
    d.tile = tile
    d.stage = copy.deepcopy(deck.stage)
    d.mesh = PS.DeviceMesh(tile)
    d.state = GP.DeviceState(tile.geom)
    d.stage.poisson_precision = 1 if precision >= 1 else 0
    d.stage.sweep_precision = 1 if precision >= 2 else 0
    d.simplec = GP.SimplecStage(d.mesh, d.inputs, d.stage)
    return d


def prepare_sequenced(deck, levels: int, passes: int = 1, smoothing: int = 0, gamma_m: float = 0.5,
                      precision: int = 0) -> Sequenced:
    """The decks of levels L-1 ... 0 on the hierarchy of deck's mesh (each
    with the run's inputs, stage and driver settings on its own mesh, its own
    boundary conditions and Poisson solver); level 0 is deck itself, prepared
    with precision unless it is already."""
    tiles = M.hierarchy(deck.tile, levels, passes, smoothing, gamma_m)
    out = Sequenced()
    for lev in range(levels - 1, 0, -1):
        out.levels.append(_level_deck(deck, tiles[lev], precision))
    if deck.simplec is None:
        deck.mesh = PS.DeviceMesh(deck.tile)
        deck.state = GP.DeviceState(deck.tile.geom)
        deck.stage.poisson_precision = 1 if precision >= 1 else 0
        deck.stage.sweep_precision = 1 if precision >= 2 else 0
        deck.simplec = GP.SimplecStage(deck.mesh, deck.inputs, deck.stage)
    out.levels.append(deck)
    return out


def solve_sequenced(seq_decks: Sequenced, seq: SequencingOptions, stream=None):
    """Steps 1-3 of the module's docstring (sequence_coarse_levels and the
    run's driver on level 0, Source/Sequencing.cpp:329-420). Returns (level
    0's LevelResultC, [LevelSummary] coarsest first, level 0 last). A coarse
    level that diverges raises GustError naming the level."""
    import torch
    L = len(seq_decks.levels)
    if L < 2:
        raise ValueError("gust: sequence_coarse_levels needs mesh.hierarchy_levels >= 2")
    top = seq_decks.levels[-1]
    err = check_sequencing_options(seq, top.driver, L, top.tile.geom.n)
    if not seq.enabled:
        err.append("steady.sequencing = 0: nothing to sequence")
    if err:
        raise _lib.GustError("gust: invalid inputs:\n  " + "\n  ".join(err))
    ref = None
    cfl_prev = 0.0
    coarse = None
    summaries = []
    for idx, d in enumerate(seq_decks.levels):
        lev = L - 1 - idx
        nx, ny, _ = d.tile.geom.n
        s = LevelSummary(lev, nx, ny)
        t0 = time.perf_counter()
        if coarse is None:
            solver = PS.PoissonSolver(d.mesh, PS.PoissonOptions(reltol=d.init_reltol))
            if d.init.projection:
                GI.initialize_state(d.mesh, d.profile, d.init, d.inputs.faces, solver, d.state)
            else:
                GI.first_guess_state(d.mesh, d.profile, d.init, d.state)
        else:
            solver = PS.PoissonSolver(d.mesh, PS.PoissonOptions(reltol=d.init_reltol))
            pr = prolong_state(d.simplec, solver, coarse.tile.geom, 2, coarse.state, d.state, d.stage.tke_floor,
                               stream)
            s.prolonged = True
            s.divergence_before = pr.projection.divergence_before
            s.divergence_after = pr.projection.divergence_after
            s.tke_clipped = int(pr.tke_clipped)
        seeded = ref is not None
        res = solve_level(d.simplec, d.state, d.driver, seq if lev > 0 else None, seed_ref=ref,
                          release_hold=seeded and seq.hold == 0,
                          start_cfl=cfl_prev if (seeded and seq.cfl_previous) else 0.0, stream=stream)
        torch.cuda.synchronize()
        s.wall_seconds = time.perf_counter() - t0
        s.iterations = int(res.driver.iterations)
        s.relative_residual = float(res.driver.relative_residual)
        if res.stop == LEVEL_DIVERGED:
            raise _lib.GustError(f"gust: grid sequencing: level {lev} diverged")
        s.stop = {LEVEL_STEADY: "steady", LEVEL_LOOSE: res.reason.decode(),
                  LEVEL_MAX_ITERATIONS: "steady.max_iterations"}[res.stop]
        summaries.append(s)
        if ref is None:
            ref = list(res.res_ref)
        cfl_prev = float(res.driver.cfl)
        coarse = d
        if lev == 0:
            return res, summaries
    raise AssertionError("unreachable")
