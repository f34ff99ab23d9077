"""The Newton-Krylov finish of the steady solve (Source/Newton.H,
Source/Continuation.H, Source/NewtonPreconditioner.H, Source/NewtonStage.H,
the driver's switch rules Source/Driver.cpp:335-400) through the C-ABI
(``gust_newton_*``, ``gust_continuation_*``, ``gust_steady_solve_newton``).

* :func:`fill_ghosts_homogeneous` -- BoundaryConditions::fill_ghosts with
  homogeneous = true (the ghost fill of a perturbation);
* :func:`jvp` -- Residual.cpp:324-396, J v by the forward difference of the
  full residual;
* :class:`NewtonStage` -- NewtonKrylov::step, residual_norm, one
  continuation step, and the driver loop over the SIMPLEC and Newton stages.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

from . import _lib
from .fab import Fab
from .physics import FIELDS, LOCS, DeviceState, DriverOptions, ResidualInputs, SimplecStage, _mesh_c
from .poisson import DeviceMesh, _stream

IDENTITY, SIMPLEC = 0, 1
# the SIMPLEC preconditioner's pressure denominators (PressureDenominator,
# Source/NewtonPreconditioner.H:116-118)
STAGE_DENOMINATOR, DIAGONAL_DENOMINATOR = 0, 1


@dataclass
class NewtonOptions:
    """steady.newton_* and the continuation's keys (Source/Newton.H:78-99,
    Source/Continuation.H:63-79); cfl_max is steady.newton_cfl_max, start_cfl
    the SIMPLEC stage's steady.cfl_max (the continuation's first CFL and the
    preconditioner's CFL when a step has none)."""
    gmres_restart: int = 30
    gmres_max_iters: int = 2000
    fd_eps_scale: float = math.sqrt(2.220446049250313e-16)
    krylov_rtol: float = 1.0e-2
    preconditioner: int = SIMPLEC
    cfl_max: float = 128.0
    start_cfl: float = 10.0
    ser_exponent: float = 1.0
    ser_max_growth: float = 2.0
    armijo_c: float = 1.0e-4
    backtrack: float = 0.5
    lambda_min: float = 1.0e-2
    cfl_cut: float = 0.25
    max_retries: int = 4
    dtau_inf_at_cap: int = 1
    ptc_growth: float = 1.0
    pc_denominator: int = DIAGONAL_DENOMINATOR

    def c(self) -> _lib.NewtonOptionsC:
        o = _lib.NewtonOptionsC()
        for name, _t in _lib.NewtonOptionsC._fields_:
            setattr(o, name, getattr(self, name))
        return o


def check_newton_options(opt: NewtonOptions) -> list:
    """check_newton_options / check_continuation_options: one message per
    violated range, naming steady.<key>."""
    buf = C.create_string_buffer(4096)
    n = C.c_int32(0)
    _lib.check(_lib.lib.gust_check_newton_options(C.byref(opt.c()), buf, len(buf), C.byref(n)))
    return [line for line in buf.value.decode().splitlines() if line][:n.value]


@dataclass
class SwitchOptions:
    """The driver's switch rules (Source/Driver.H:67-97)."""
    switch_window: int = 20
    switch_residual: float = 1.0e-4
    switch_back_growth: float = 10.0
    switch_monitor_tol: float = 1.0e-2
    newton_fallback_failures: int = 2

    def c(self) -> _lib.SwitchOptionsC:
        return _lib.SwitchOptionsC(int(self.switch_window), self.switch_residual, self.switch_back_growth,
                                   self.switch_monitor_tol, int(self.newton_fallback_failures))


def fill_ghosts_homogeneous(mesh: DeviceMesh, q: DeviceState, inputs: ResidualInputs, stream=None) -> None:
    """BoundaryConditions::fill_ghosts(q, ..., homogeneous = true)
    (Source/BoundaryConditions.cpp:418-540, 597-666): every prescribed value,
    inflow profile and z gradient replaced by 0."""
    g = mesh.geom.c()
    qc, pc = q.c(), inputs.c()
    mc, _keep = _mesh_c(mesh)
    _lib.check(_lib.lib.gust_fill_ghosts_homogeneous(C.byref(g), C.byref(qc), C.byref(mc), C.byref(pc),
                                                     _stream(stream)))


def residual_fields(geom, **kw) -> dict:
    """Fabs for R or J v: the six fields' valid points."""
    return {n: Fab.at_loc(geom, loc, 1, 0, **kw) for n, loc in zip(FIELDS, LOCS)}


def _res_c(R: dict):
    views = [R[n].array4() for n in FIELDS]
    rc = _lib.ResidualsC()
    for k in range(6):
        rc.R[k] = C.pointer(views[k])
    return rc, views


def jvp(stage: SimplecStage, q: DeviceState, v: DeviceState, R0: dict, Jv: dict, stream=None) -> float:
    """Source/Residual.cpp:324-396: Jv = (R(fill(q + eps v)) - R0) / eps on
    the stage's tile, with eps = step (1 + mean|q/s|) / rms(v/s); q and v
    with their ghosts filled (v homogeneously), R0 = R(q). Returns eps (0:
    v = 0, and Jv = 0)."""
    qc, vc = q.c(), v.c()
    r0, k0 = _res_c(R0)
    jv, k1 = _res_c(Jv)
    eps = C.c_double(0)
    _lib.check(_lib.lib.gust_jvp(stage._h, C.byref(qc), C.byref(vc), C.byref(r0), C.byref(jv), C.byref(eps),
                                 _stream(stream)))
    return float(eps.value)


def _rec(r: _lib.ContinuationRecordC) -> dict:
    out = {name: getattr(r, name) for name, _t in _lib.ContinuationRecordC._fields_}
    out["lambda"] = out.pop("lambda_")
    out["outcome"] = ("accepted", "skipped", "rejected")[out["outcome"]]
    return out


class NewtonStage:
    """NewtonKrylov, Continuation, SimplecPreconditioner and NewtonStage on a
    SIMPLEC stage's tile (the stage must outlive this object)."""

    def __init__(self, stage: SimplecStage, opt: NewtonOptions | None = None, stream=None):
        self.stage, self.opt = stage, opt or NewtonOptions()
        h = C.c_void_p(0)
        _lib.check(_lib.lib.gust_newton_create(stage._h, C.byref(self.opt.c()), C.byref(h), _stream(stream)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib.lib is not None:
            _lib.lib.gust_newton_destroy(h)
            self._h = None

    def step(self, q: DeviceState, dtau: Fab | None = None, stream=None) -> dict:
        """NewtonKrylov::step (Source/Newton.cpp:531-539): q += dq from
        (D + J) dq = -R(q); dtau None: plain Newton."""
        qc = q.c()
        info = _lib.NewtonStepInfoC()
        _lib.check(_lib.lib.gust_newton_step(self._h, C.byref(qc), C.byref(dtau.array4()) if dtau is not None
                                             else None, C.byref(info), _stream(stream)))
        return {name: getattr(info, name) for name, _t in _lib.NewtonStepInfoC._fields_}

    def residual_norm(self, q: DeviceState, stream=None) -> float:
        """|T^-1 R(fill(q))|_2 (Source/Newton.cpp:553-565)."""
        qc = q.c()
        r = C.c_double(0)
        _lib.check(_lib.lib.gust_newton_residual_norm(self._h, C.byref(qc), C.byref(r), _stream(stream)))
        return float(r.value)

    def pc_setup(self, q: DeviceState, dtau: Fab | None = None, stream=None) -> None:
        """SimplecPreconditioner::setup (Source/NewtonPreconditioner.cpp:
        114-148): the linearisation point q (ghosts filled) and the cell step
        (None: the local step at start_cfl)."""
        qc = q.c()
        _lib.check(_lib.lib.gust_newton_pc_setup(self._h, C.byref(qc), C.byref(dtau.array4()) if dtau is not None
                                                 else None, _stream(stream)))

    def pc_apply(self, r: dict, dq: dict, stream=None) -> None:
        """SimplecPreconditioner::apply (Source/NewtonPreconditioner.cpp:
        150-224): dq = one SIMPLEC iteration on (D + J) dq = r; r and dq as
        :func:`residual_fields` gives them."""
        rc, k0 = _res_c(r)
        dc, k1 = _res_c(dq)
        _lib.check(_lib.lib.gust_newton_pc_apply(self._h, C.byref(rc), C.byref(dc), _stream(stream)))

    def continuation_step(self, q: DeviceState, stream=None):
        """Continuation::step (Source/Continuation.cpp:339-427): returns
        (accepted, the last attempt's record)."""
        qc = q.c()
        st = C.c_int32(0)
        rec = _lib.ContinuationRecordC()
        _lib.check(_lib.lib.gust_continuation_step(self._h, C.byref(qc), C.byref(st), C.byref(rec),
                                                   _stream(stream)))
        return st.value == 0, _rec(rec)

    @property
    def continuation_state(self) -> dict:
        s = _lib.ContinuationStateC()
        _lib.check(_lib.lib.gust_continuation_get(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _t in _lib.ContinuationStateC._fields_}

    def restore(self, steps: int, cfl: float, infinite: bool, r0: float, r_last: float) -> None:
        """Continuation::restore (Source/Continuation.cpp:291-305)."""
        s = _lib.ContinuationStateC(int(steps), cfl, 1 if infinite else 0, r0, r_last)
        _lib.check(_lib.lib.gust_continuation_restore(self._h, C.byref(s)))

    def solve(self, q: DeviceState, opt: DriverOptions | None = None, switch: SwitchOptions | None = None,
              stream=None) -> dict:
        """steady.stage = simplec_newton: the driver loop over the SIMPLEC
        stage and this stage with the switch rules (Source/Driver.cpp:335-492)."""
        qc = q.c()
        oc = (opt or DriverOptions()).c()
        sc = (switch or SwitchOptions()).c()
        r = _lib.NewtonRunResultC()
        _lib.check(_lib.lib.gust_steady_solve_newton(self.stage._h, self._h, C.byref(qc), C.byref(oc), C.byref(sc),
                                                     C.byref(r), _stream(stream)))
        d = r.driver
        return {"iterations": d.iterations, "status": d.status, "relative_residual": d.relative_residual,
                "res": list(d.res), "mon": list(d.mon[:d.nmon]), "simplec_iterations": r.simplec_iterations,
                "newton_iterations": r.newton_iterations, "switches": r.switches, "first_switch": r.first_switch,
                "newton_steps": r.newton_steps, "stage": r.stage, "handed_back": r.handed_back}
