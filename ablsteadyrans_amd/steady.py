"""The steady criterion (Source/SteadyCriterion.H/.cpp) and the two driver
steps that read it (Source/Driver.cpp:325-334, 421-437), over the host-side
C++ in ``libgust_amd.so``. Names, arguments and refusals mirror the
reference: an invalid criterion, a size mismatch or a foreign history raises
:class:`~ablsteadyrans_amd._lib.GustError` with the reference's text."""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field

from . import _lib


@dataclass
class SteadyParams:
    """Source/SteadyCriterion.H:13-33 (the defaults are the reference's)."""

    residual_tol: float = 5.0e-5
    residual_drop_orders: float = 4.0
    residual_ref_iterations: int = 5
    monitor_rel_tol: float = 1.0e-3
    monitor_window: int = 50

    def c(self) -> _lib.SteadyParamsC:
        return _lib.SteadyParamsC(float(self.residual_tol), float(self.residual_drop_orders),
                                  int(self.residual_ref_iterations), float(self.monitor_rel_tol),
                                  int(self.monitor_window))


@dataclass
class MonitorSpec:
    """Source/SteadyCriterion.H:38-48: a monitored quantity and its scale (no default)."""

    name: str
    scale: float


class SteadyStatus(enum.IntEnum):
    """Source/SteadyCriterion.H:50-55."""

    NotSteady = 0
    Steady = 1
    Diverged = 2


@dataclass
class SteadyVerdict:
    """Source/SteadyCriterion.H:57-64."""

    status: SteadyStatus = SteadyStatus.NotSteady
    iteration: int = -1
    reasons: list = field(default_factory=list)
    residuals_met: bool = False
    monitors_met: bool = False


def validate(p: SteadyParams) -> list:
    """Source/SteadyCriterion.cpp:23-48: one message per violated range."""
    n = C.c_int32(0)
    buf = C.create_string_buffer(4096)
    pc = p.c()
    _lib.check(_lib.lib.gust_steady_validate(C.byref(pc), C.byref(n), buf, len(buf)))
    return buf.value.decode().split("\n") if n.value else []


def _doubles(xs):
    xs = [float(x) for x in xs]
    return (C.c_double * max(1, len(xs)))(*xs), len(xs)


class SteadyCriterion:
    """Source/SteadyCriterion.H:88-132: records residual and monitor
    histories and decides whether the solution is steady."""

    def __init__(self, p: SteadyParams, residual_names, monitors):
        self.params = p
        self.residual_names = list(residual_names)
        self.monitors = [m if isinstance(m, MonitorSpec) else MonitorSpec(*m) for m in monitors]
        rn = (C.c_char_p * max(1, len(self.residual_names)))(*[s.encode() for s in self.residual_names])
        mn = (C.c_char_p * max(1, len(self.monitors)))(*[m.name.encode() for m in self.monitors])
        ms, _ = _doubles([m.scale for m in self.monitors])
        h = C.c_void_p(0)
        pc = p.c()
        _lib.check(_lib.lib.gust_steady_create(C.byref(pc), len(self.residual_names), rn, len(self.monitors), mn, ms,
                                               C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib.lib is not None:
            _lib.lib.gust_steady_destroy(h)
            self._h = None

    def add(self, residuals, monitors) -> None:
        """Source/SteadyCriterion.cpp:96-131."""
        r, nr = _doubles(residuals)
        m, nm = _doubles(monitors)
        _lib.check(_lib.lib.gust_steady_add(self._h, r, nr, m, nm))

    def verdict(self) -> SteadyVerdict:
        """Source/SteadyCriterion.cpp:133-227."""
        v = _lib.SteadyResultC()
        _lib.check(_lib.lib.gust_steady_verdict(self._h, C.byref(v), None, 0))
        buf = C.create_string_buffer(int(v.reasons_len) + 1)
        _lib.check(_lib.lib.gust_steady_verdict(self._h, C.byref(v), buf, len(buf)))
        reasons = buf.value.decode().split("\n") if v.nreasons else []
        return SteadyVerdict(SteadyStatus(v.status), v.iteration, reasons, bool(v.residuals_met),
                             bool(v.monitors_met))

    def iterations(self) -> int:
        n = C.c_int32(0)
        _lib.check(_lib.lib.gust_steady_iterations(self._h, C.byref(n)))
        return n.value

    def residual_thresholds(self) -> list:
        """Source/SteadyCriterion.cpp:229-238."""
        out = (C.c_double * len(self.residual_names))()
        _lib.check(_lib.lib.gust_steady_residual_thresholds(self._h, out, len(self.residual_names)))
        return list(out)

    def seed_reference(self, ref) -> None:
        """Source/SteadyCriterion.cpp:85-94."""
        r, n = _doubles(ref)
        _lib.check(_lib.lib.gust_steady_seed_reference(self._h, r, n))

    def references(self) -> list:
        """The residual references (SteadyHistory::res_ref)."""
        out = (C.c_double * len(self.residual_names))()
        _lib.check(_lib.lib.gust_steady_references(self._h, out, len(self.residual_names)))
        return list(out)

    def history(self) -> bytes:
        """The recorded history (SteadyHistory) as the library's byte image."""
        size = C.c_int64(0)
        _lib.check(_lib.lib.gust_steady_history_save(self._h, None, 0, C.byref(size)))
        buf = C.create_string_buffer(max(1, size.value))
        _lib.check(_lib.lib.gust_steady_history_save(self._h, buf, len(buf), C.byref(size)))
        return buf.raw[:size.value]

    def restore(self, image: bytes) -> None:
        """Source/SteadyCriterion.cpp:72-83."""
        _lib.check(_lib.lib.gust_steady_history_restore(self._h, image, len(image)))

    def relative_residual(self, residuals, hold: int, previous: float = 0.0) -> float:
        """Driver::step's r_n (Source/Driver.cpp:421-437)."""
        r, n = _doubles(residuals)
        out = C.c_double(0)
        _lib.check(_lib.lib.gust_steady_relative_residual(self._h, r, n, int(hold), float(previous), C.byref(out)))
        return out.value

    def next_cfl(self, hold: int, cfl: float, r: float, cfl0: float, cfl_max: float, ser_exponent: float,
                 max_growth: float) -> float:
        """Driver::next_cfl (Source/Driver.cpp:325-334)."""
        out = C.c_double(0)
        _lib.check(_lib.lib.gust_steady_next_cfl(self._h, int(hold), float(cfl), float(r), float(cfl0),
                                                 float(cfl_max), float(ser_exponent), float(max_growth),
                                                 C.byref(out)))
        return out.value
