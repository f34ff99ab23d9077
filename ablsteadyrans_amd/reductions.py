"""Reductions of the SIMPLEC control loop over device FABs
(Source/Reductions.cpp, Source/Residual.cpp:101-150,
Source/Simplec.cpp:374-445): the column-ordered deterministic sum, the
NaN-reporting maximum, the scaled sums of the Newton step, the stage's
scaled residual norms and monitors, and the confidence share. Each call
forwards to ``libgust_amd.so`` and returns host numbers (it synchronises the
stream, as the reference's reductions do)."""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

from . import _lib
from .fab import CELL, Fab, Geometry
from .poisson import _stream


@dataclass
class JvpScales:
    """Source/Residual.H:72-83: typical magnitudes of the fields."""

    velocity: float = 10.0
    theta: float = 10.0
    tke: float = 0.1
    pressure: float = 100.0
    step: float = math.sqrt(2.220446049250313e-16)

    def c(self) -> _lib.JvpScalesC:
        return _lib.JvpScalesC(self.velocity, self.theta, self.tke, self.pressure, self.step)


def _res(R) -> tuple:
    views = [f.array4() for f in R]
    rc = _lib.ResidualsC()
    for n in range(6):
        rc.R[n] = C.pointer(views[n])
    return rc, views


def deterministic_sum(geom: Geometry, a: Fab, comp: int = 0, stream=None) -> float:
    """Source/Reductions.cpp:11-53 on one box: columns in k order on the
    device, column sums in storage order on the host."""
    out = C.c_double(0)
    g = geom.c()
    _lib.check(_lib.lib.gust_deterministic_sum(C.byref(g), C.byref(a.array4()), int(comp), C.byref(out),
                                               _stream(stream)))
    return out.value


def global_max(geom: Geometry, a: Fab, loc: int = CELL, comp: int = 0, stream=None) -> float:
    """Source/Reductions.cpp:55-66: NaN if any valid value is NaN; +0 for zero."""
    out = C.c_double(0)
    g = geom.c()
    _lib.check(_lib.lib.gust_global_max(C.byref(g), int(loc), C.byref(a.array4()), int(comp), C.byref(out),
                                        _stream(stream)))
    return out.value


def scaled_sums(geom: Geometry, loc: int, q: Fab, v: Fab, s: float, offset: Fab | None = None,
                stream=None) -> tuple:
    """Source/Residual.cpp:101-150 on one box: (sum |(q - offset)/s|,
    sum (v/s)^2, count) over the owned points at loc."""
    out = (C.c_double * 3)()
    g = geom.c()
    off = C.byref(offset.array4()) if offset is not None else None
    _lib.check(_lib.lib.gust_scaled_sums(C.byref(g), int(loc), C.byref(q.array4()), C.byref(v.array4()), off,
                                         float(s), out, _stream(stream)))
    return out[0], out[1], out[2]


def residual_scales(scales: JvpScales, height: float) -> list:
    """Source/Simplec.cpp:374-382: U^2/H (u, v, w), U/H, U theta_s/H, U k_s/H."""
    out = (C.c_double * 6)()
    sc = scales.c()
    _lib.check(_lib.lib.gust_residual_scales(C.byref(sc), float(height), out))
    return list(out)


def residual_scale_height(zlevels, prob_lo_z: float, prob_hi_z: float) -> float:
    """Source/Mesh.cpp:332-340."""
    z = (C.c_double * len(zlevels))(*[float(x) for x in zlevels])
    out = C.c_double(0)
    _lib.check(_lib.lib.gust_residual_scale_height(z, len(zlevels), float(prob_lo_z), float(prob_hi_z),
                                                   C.byref(out)))
    return out.value


def stage_norms(geom: Geometry, R, u: Fab, v: Fab, w: Fab, tke: Fab, scales: JvpScales, height: float,
                ustar: Fab | None = None, stream=None) -> tuple:
    """Source/Simplec.cpp:385-409: (res[6], monitors) with R = (u, v, w, p,
    theta, k) residuals; monitors = [max speed, mean k(, mean u*)]."""
    rc, _keep = _res(R)
    res = (C.c_double * 6)()
    mon = (C.c_double * 3)()
    nmon = C.c_int32(0)
    g = geom.c()
    sc = scales.c()
    us = C.byref(ustar.array4()) if ustar is not None else None
    _lib.check(_lib.lib.gust_stage_norms(C.byref(g), C.byref(rc), C.byref(u.array4()), C.byref(v.array4()),
                                         C.byref(w.array4()), C.byref(tke.array4()), us, C.byref(sc), float(height),
                                         res, mon, C.byref(nmon), _stream(stream)))
    return list(res), list(mon[:nmon.value])


def confidence(geom: Geometry, R, z_nd: Fab, scales: JvpScales, height: float, thresholds,
               stream=None) -> tuple:
    """Source/Simplec.cpp:411-445: (share of converged cells, largest height
    above the ground of the others)."""
    if len(thresholds) != 6:
        raise _lib.GustError("gust: confidence: needs six thresholds")
    rc, _keep = _res(R)
    th = (C.c_double * 6)(*[float(x) for x in thresholds])
    share, top = C.c_double(0), C.c_double(0)
    g = geom.c()
    sc = scales.c()
    _lib.check(_lib.lib.gust_confidence(C.byref(g), C.byref(rc), C.byref(z_nd.array4()), C.byref(sc), float(height),
                                        th, C.byref(share), C.byref(top), _stream(stream)))
    return share.value, top.value
