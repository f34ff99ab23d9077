"""The z-line solvers, mirroring the reference's call signatures.

* ``zline_jacobi(loc, A, ap, rhs, sweeps, geom, delta)`` --
  Source/Predictor.H:267-268 (returns max |change over the last sweep|).
* ``zline_lusgs(loc, A, ap, rhs, sweeps, geom, delta, opt)`` --
  Source/LineSweeps.H:159-161.
* ``line_sweeps(opt, loc, A, ap, rhs, sweeps, geom, delta)`` --
  Source/Predictor.H:257-259.

The ``Mesh`` argument of the reference is not needed: the checks it serves
(layout, whole columns) are the FABs' boxes against ``geom``. Refusals raise
:class:`GustError` carrying the reference's "gust: <op>: ..." message, as
``amrex::Abort`` does under ``amrex.throw_exception=1``. Every call runs on the
HIP library; there is no other path.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from .fab import Fab, Geometry, LineSweepOptions, LUSGS


def _stream(stream):
    if stream is not None:
        return C.c_void_p(int(stream))
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _is_f32(fab: Fab) -> bool:
    import torch
    return fab.data.dtype == torch.float32


def zline_jacobi(loc: int, A: Fab, ap: Fab, rhs: Fab, sweeps: int, geom: Geometry, delta: Fab,
                 stream=None) -> float:
    change = C.c_double(0.0)
    f = _lib.lib.gust_zline_jacobi_f32 if _is_f32(delta) else _lib.lib.gust_zline_jacobi
    a, p, r, d, g = A.array4(), ap.array4(), rhs.array4(), delta.array4(), geom.c()
    _lib.check(f(int(loc), C.byref(a), C.byref(p), C.byref(r), int(sweeps), C.byref(g), C.byref(d),
                 C.byref(change), _stream(stream)))
    return change.value


def zline_jacobi_noalloc(loc: int, A: Fab, ap: Fab, rhs: Fab, sweeps: int, geom: Geometry, delta: Fab,
                         stream=None) -> float:
    """zline_jacobi through the entry point that allocates nothing (its
    scratch is a fixed array per device; FP64 only)."""
    change = C.c_double(0.0)
    a, p, r, d, g = A.array4(), ap.array4(), rhs.array4(), delta.array4(), geom.c()
    _lib.check(_lib.lib.gust_zline_jacobi_noalloc(int(loc), C.byref(a), C.byref(p), C.byref(r), int(sweeps),
                                                  C.byref(g), C.byref(d), C.byref(change), _stream(stream)))
    return change.value


def zline_lusgs(loc: int, A: Fab, ap: Fab, rhs: Fab, sweeps: int, geom: Geometry, delta: Fab,
                opt: LineSweepOptions | None = None, stream=None) -> float:
    opt = opt or LineSweepOptions(kind=LUSGS)
    change = C.c_double(0.0)
    a, p, r, d, g, o = A.array4(), ap.array4(), rhs.array4(), delta.array4(), geom.c(), opt.c()
    _lib.check(_lib.lib.gust_zline_lusgs(int(loc), C.byref(a), C.byref(p), C.byref(r), int(sweeps), C.byref(g),
                                         C.byref(d), C.byref(o), C.byref(change), _stream(stream)))
    return change.value


def line_sweeps(opt: LineSweepOptions, loc: int, A: Fab, ap: Fab, rhs: Fab, sweeps: int, geom: Geometry,
                delta: Fab, stream=None) -> float:
    if opt.kind == LUSGS:
        return zline_lusgs(loc, A, ap, rhs, sweeps, geom, delta, opt, stream)
    return zline_jacobi(loc, A, ap, rhs, sweeps, geom, delta, stream)


def check_line_options(opt: LineSweepOptions) -> list[str]:
    """Source/LineSweeps.cpp:315-329: [] if valid, else the message (naming steady.<key>)."""
    o = opt.c()
    if _lib.lib.gust_check_line_options(C.byref(o)) == 0:
        return []
    return [_lib.lib.gust_last_error().decode()]


JACOBI_AUTO, JACOBI_STREAMING, JACOBI_STAGED, JACOBI_RULE, JACOBI_REGISTER = 0, 1, 2, 3, 4


def set_jacobi_schedule(schedule: int) -> None:
    """The z-line Jacobi kernel schedule for plans made afterwards (process
    wide): JACOBI_AUTO (autotuned per set of system shapes: the first line
    solve times the candidates and the fastest is kept), JACOBI_STREAMING,
    JACOBI_STAGED, JACOBI_RULE (few tiles staged in LDS, many streamed) or
    JACOBI_REGISTER (e' held in registers, columns of up to 97 levels; taller
    ones stream). Every schedule gives the same bits."""
    _lib.check(_lib.lib.gust_jacobi_set_schedule(int(schedule)))


TUNE_CHOICES = ("staged", "streaming-p4", "streaming-p2", "streaming-p8", "streaming-global-cprime", "register-e",
                "register-ce", "register-split")


def jacobi_tuned() -> list:
    """The autotuner's records: dicts of dtype, sweeps, systems, the chosen
    candidate and every candidate's time in ms (None: did not apply)."""
    n = C.c_int32(0)
    _lib.check(_lib.lib.gust_jacobi_tuned(-1, C.byref(n), None, None, None, None, None, None, 0))
    out = []
    for i in range(n.value):
        dt, sw, ns, ch = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        ms = (C.c_float * len(TUNE_CHOICES))()
        _lib.check(_lib.lib.gust_jacobi_tuned(i, None, C.byref(dt), C.byref(sw), C.byref(ns), C.byref(ch), ms,
                                              None, 0))
        out.append({"dtype": "f64" if dt.value == 0 else "f32", "sweeps": sw.value, "systems": ns.value,
                    "choice": TUNE_CHOICES[ch.value],
                    "ms": {TUNE_CHOICES[c]: (round(float(ms[c]), 4) if ms[c] >= 0 else None) for c in range(len(TUNE_CHOICES))}})
    return out


def jacobi_tune_reset() -> None:
    """Forgets the autotuner's records."""
    _lib.check(_lib.lib.gust_jacobi_tune_reset())


class LineSystem:
    """One system of a batched launch: a field at ``loc`` on a tile."""

    def __init__(self, loc: int, geom: Geometry, A: Fab, ap: Fab, rhs: Fab, delta: Fab):
        self.loc, self.geom, self.A, self.ap, self.rhs, self.delta = loc, geom, A, ap, rhs, delta

    def c(self) -> _lib.LineSystemC:
        s = _lib.LineSystemC()
        s.loc = int(self.loc)
        s.geom = self.geom.c()
        s.A, s.ap, s.rhs, s.delta = self.A.array4(), self.ap.array4(), self.rhs.array4(), self.delta.array4()
        return s


def zline_jacobi_batched(systems: list[LineSystem], sweeps: int, changes=None, stream=None) -> None:
    """All systems in one launch per sweep; ``changes`` (a float64 device tensor
    of len(systems)) receives each system's last-sweep change. Asynchronous."""
    n = len(systems)
    arr = (_lib.LineSystemC * n)(*[s.c() for s in systems])
    f32 = _is_f32(systems[0].delta)
    f = _lib.lib.gust_zline_jacobi_batched_async_f32 if f32 else _lib.lib.gust_zline_jacobi_batched_async
    ptr = C.c_void_p(changes.data_ptr()) if changes is not None else C.c_void_p(0)
    _lib.check(f(n, arr, int(sweeps), ptr, _stream(stream)))


class JacobiPlan:
    """A batched Jacobi plan (``gust_jacobi_plan_*``): set up once, then
    ``run`` enqueues kernels only (graph-capturable). Keeps the systems'
    tensors alive for its lifetime."""

    def __init__(self, systems: list[LineSystem], stream=None):
        import torch
        self.systems = list(systems)
        n = len(self.systems)
        self._arr = (_lib.LineSystemC * n)(*[s.c() for s in self.systems])
        dtype = 1 if _is_f32(self.systems[0].delta) else 0
        h = C.c_void_p(0)
        _lib.check(_lib.lib.gust_jacobi_plan_create(n, self._arr, dtype, C.byref(h), _stream(stream)))
        self._h = h
        u, t = C.c_int64(0), C.c_int32(0)
        _lib.check(_lib.lib.gust_jacobi_plan_info(self._h, C.byref(u), C.byref(t)))
        self.unknowns, self.tiles = u.value, t.value
        self.changes = torch.zeros(n, dtype=torch.float64, device=self.systems[0].delta.data.device)

    def run(self, sweeps: int, stream=None) -> None:
        _lib.check(_lib.lib.gust_jacobi_plan_run(self._h, int(sweeps), C.c_void_p(self.changes.data_ptr()),
                                                 _stream(stream)))

    def sweep(self, first: bool, last: bool, into_scratch: bool, stream=None) -> None:
        _lib.check(_lib.lib.gust_jacobi_plan_sweep(self._h, int(first), int(last), int(into_scratch),
                                                   _stream(stream)))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib.gust_jacobi_plan_destroy(self._h)
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
