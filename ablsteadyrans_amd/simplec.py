"""SIMPLEC steps (Source/Simplec.H) over device FABs: the pressure correction
(Source/Simplec.cpp:134-220). Each call forwards to ``libgust_amd.so``."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import _lib
from .fab import Fab
from .poisson import DeviceMesh, PoissonResult, PoissonSolver, _bc, _stream


@dataclass
class PressureCorrectionResult:
    """Source/Simplec.H:116-123."""

    poisson: PoissonResult
    divergence_before: float
    divergence_after: float
    max_dp: float
    max_du: float
    max_dv: float
    max_dw: float


def pressure_correction(solver: PoissonSolver, mesh: DeviceMesh, u: Fab, v: Fab, w: Fab, p: Fab, d_u: Fab, d_v: Fab,
                        d_w: Fab, bc, alpha_p: float = 1.0, source: Fab | None = None,
                        project_net_source: bool = False, stream=None) -> PressureCorrectionResult:
    """Source/Simplec.cpp:134-220: corrects u, v, w, p in place."""
    g = mesh.geom.c()
    res = _lib.PressureCorrectionResultC()
    views = [f.array4() for f in (u, v, w, p, mesh.z_nd, mesh.ax, mesh.ay, mesh.az, mesh.detJ, d_u, d_v, d_w)]
    src = C.byref(source.array4()) if source is not None else None
    _lib.check(_lib.lib.gust_pressure_correction(solver._h, C.byref(g), *[C.byref(a) for a in views], _bc(bc),
                                                 float(alpha_p), src, int(bool(project_net_source)), C.byref(res),
                                                 _stream(stream)))
    pr = res.poisson
    return PressureCorrectionResult(
        PoissonResult(pr.iterations, pr.initial_residual, pr.final_residual, bool(pr.singular), pr.removed_mean),
        res.divergence_before, res.divergence_after, res.max_dp, res.max_du, res.max_dv, res.max_dw)
