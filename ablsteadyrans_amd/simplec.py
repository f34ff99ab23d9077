"""SIMPLEC steps (Source/Simplec.H, Source/Predictor.H) over device FABs:
the momentum predictor (Source/Predictor.cpp:129-405), the scalar sweep
(Source/Simplec.cpp:222-281) and the pressure correction
(Source/Simplec.cpp:134-220). Each call forwards to ``libgust_amd.so``."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

from . import _lib
from .fab import Fab, LineSweepOptions
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


@dataclass
class PredictorOptions:
    """Source/Predictor.H:242-252 and simplec.* (Source/Simplec.H:26-40)."""

    sweeps: int = 2
    rotation_relax: float = 0.3
    w_vnn_cap: float = 0.0
    w_vnn_slope_margin: float = 4.0
    coriolis_f: float = 0.0
    gravity: float = 0.0
    theta0: float = 300.0
    sweep: LineSweepOptions = field(default_factory=LineSweepOptions)
    sweep_precision: int = 0   # 1: FP32 sweeps inside the FP64 defect correction

    def c(self):
        o = _lib.PredictorOptionsC()
        o.sweeps = int(self.sweeps)
        o.rotation_relax, o.w_vnn_cap, o.w_vnn_slope_margin = self.rotation_relax, self.w_vnn_cap, \
            self.w_vnn_slope_margin
        o.coriolis_f, o.gravity, o.theta0 = self.coriolis_f, self.gravity, self.theta0
        o.sweep = self.sweep.c()
        o.sweep_precision = int(self.sweep_precision)
        return o


@dataclass
class MomentumPredictorResult:
    """Source/Predictor.H:214-220."""

    max_du: float
    max_dv: float
    max_dw: float
    last_sweep_du: float
    last_sweep_dv: float
    last_sweep_dw: float


@dataclass
class ScalarSweepResult:
    """Source/Simplec.H:144-149."""

    max_dq: float
    last_sweep: float
    clipped: int


def _faces(fabs, keep):
    """A gust_faces over three FABs (None entries stay null); the views are
    kept alive in `keep`."""
    f = _lib.FacesC()
    for c in range(3):
        if fabs[c] is not None:
            a = fabs[c].array4()
            keep.append(a)
            f.c[c] = C.pointer(a)
    return f


def momentum_predictor(mesh: DeviceMesh, vel, R, A, dtau_faces, opt: PredictorOptions, ap, d, theta: Fab = None,
                       R_theta: Fab = None, A_theta: Fab = None, mask=None, stream=None) -> MomentumPredictorResult:
    """Source/Predictor.cpp:129-405: updates vel = (u, v, w) in place; ap, d
    receive a_P and the SIMPLEC denominators."""
    g = mesh.geom.c()
    keep = []
    fv, fR, fA, fdt, fap, fd = (_faces(x, keep) for x in (vel, R, A, dtau_faces, ap, d))
    fm = _faces(mask, keep) if mask is not None else None
    res = _lib.MomentumPredictorResultC()
    o = opt.c()
    ref = lambda f: C.byref(f.array4()) if f is not None else None
    _lib.check(_lib.lib.gust_momentum_predictor(
        C.byref(g), C.byref(fv), ref(theta), C.byref(mesh.z_nd.array4()), C.byref(fR), ref(R_theta), C.byref(fA),
        ref(A_theta), C.byref(fdt), C.byref(fm) if fm is not None else None, C.byref(o), C.byref(fap), C.byref(fd),
        C.byref(res), _stream(stream)))
    return MomentumPredictorResult(res.max_du, res.max_dv, res.max_dw, res.last_sweep_du, res.last_sweep_dv,
                                   res.last_sweep_dw)


THETA, TKE = 0, 1


def scalar_sweep(geom, eq: int, fld: Fab, R_eq: Fab, A: Fab, dtau: Fab, sweeps: int = 2, floor: float = 0.0,
                 clip: bool = True, sweep: LineSweepOptions | None = None, stream=None) -> ScalarSweepResult:
    """Source/Simplec.cpp:222-281: fld += dq in place."""
    g = geom.c()
    res = _lib.ScalarSweepResultC()
    o = (sweep or LineSweepOptions()).c()
    _lib.check(_lib.lib.gust_scalar_sweep(C.byref(g), int(eq), C.byref(fld.array4()), C.byref(R_eq.array4()),
                                          C.byref(A.array4()), C.byref(dtau.array4()), int(sweeps), float(floor),
                                          int(bool(clip)), C.byref(o), C.byref(res), _stream(stream)))
    return ScalarSweepResult(res.max_dq, res.last_sweep, int(res.clipped))
