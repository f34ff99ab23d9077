"""The mass-consistent initializer, mirroring Source/Initializer.H and
Source/InflowProfile.H over the HIP library:

* :class:`InflowProfile` -- InflowProfileOptions (Source/InflowProfile.H:
  42-66) with the file kind's table given in memory, and the sounding;
* :class:`InitOptions` -- InitOptions (Source/Initializer.H:63-98), the
  profile first guess;
* :func:`inflow_eval` -- InflowProfile::View::eval at device heights;
* :func:`initializer_poisson_bc`, :func:`obrien_omega`,
  :func:`first_guess_state`, :func:`initialize_state` -- the functions of
  Source/Initializer.H:108-157.

Face types are the residual's codes (``physics.PERIODIC`` .. ``physics.GEOSTROPHIC``: periodic, inflow,
outflow, slip_wall, no_slip_wall, surface_layer, geostrophic) in the order
xlo, ylo, zlo, xhi, yhi, zhi. Refusals raise GustError with the reference's
"gust: <op>: ..." text; no path runs without the HIP library.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .fab import Fab, Geometry
from .physics import DeviceState, _mesh_c
from .poisson import DeviceMesh, PoissonResult, PoissonSolver, _stream

LOG_LAW, POWER_LAW, TABLE = 0, 1, 2


def _dbl(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class InflowProfile:
    """bc.inflow_* (Source/InflowProfile.H:42-66; the defaults are the
    reference's). ``table`` = (z, u, v, theta, k or None) for the table kind;
    the sounding is bc.theta_profile_z / bc.theta_profile."""
    kind: int = LOG_LAW
    ustar: float = 0.4
    z0: float = 0.1
    speed_ref: float = 10.0
    z_ref: float = 10.0
    alpha: float = 0.14
    direction: float = 0.0
    Cmu0: float = 0.5562
    kappa: float = 0.41
    unbounded: bool = False
    l_max: float = 30.0
    table: tuple | None = None
    theta_ref_z: tuple = (0.0,)
    theta_ref: tuple = (300.0,)

    def c(self):
        p = _lib.InflowProfileC()
        p.kind = int(self.kind)
        p.ustar, p.z0 = float(self.ustar), float(self.z0)
        p.speed_ref, p.z_ref, p.alpha = float(self.speed_ref), float(self.z_ref), float(self.alpha)
        p.direction, p.Cmu0, p.kappa = float(self.direction), float(self.Cmu0), float(self.kappa)
        p.unbounded = 1 if self.unbounded else 0
        p.l_max = float(self.l_max)
        keep = []
        if self.table is not None:
            cols = list(self.table)
            n = len(cols[0])
            for name, col in zip(("z", "u", "v", "theta", "k"), cols):
                if col is None:
                    continue
                if len(col) != n:
                    raise _lib.GustError("gust: InflowProfile: the table's columns need the same length")
                arr, ptr = _dbl(col)
                keep.append(arr)
                setattr(p, name, ptr)
            p.n = n
        n = len(self.theta_ref_z)
        if n != len(self.theta_ref) or n > 64:
            raise _lib.GustError("gust: InflowProfile: reference profile: z and theta need the same number of "
                                 "points, at least one (at most 64)")
        p.n_theta_ref = n
        for m in range(n):
            p.theta_ref_z[m] = float(self.theta_ref_z[m])
            p.theta_ref[m] = float(self.theta_ref[m])
        return p, keep


@dataclass
class InitOptions:
    """init.* (Source/Initializer.H:63-98) for the profile first guess."""
    first_guess: int = 0
    projection: bool = True
    z_agl_min: float = 0.0
    alpha_h: float = 1.0
    alpha_v: float = 1.0
    k_decay_height: float = 0.0
    obrien: bool = False
    column_theta: bool = False

    def c(self):
        return _lib.InitOptionsC(int(self.first_guess), 1 if self.projection else 0, float(self.z_agl_min),
                                 float(self.alpha_h), float(self.alpha_v), float(self.k_decay_height),
                                 1 if self.obrien else 0, 1 if self.column_theta else 0)


@dataclass
class InitResult:
    """Source/Initializer.H:110-120."""
    poisson: PoissonResult
    divergence_before: float
    divergence_after: float
    obrien_top_residual: float


def check_inflow_profile(p: InflowProfile) -> None:
    pc, keep = p.c()
    _lib.check(_lib.lib.gust_check_inflow_profile(C.byref(pc)))
    del keep


def check_init_options(o: InitOptions) -> None:
    oc = o.c()
    _lib.check(_lib.lib.gust_check_init_options(C.byref(oc)))


def inflow_eval(p: InflowProfile, z, stream=None):
    """(u, v, theta, k, nu_t) at the heights z (a 1-D device tensor), each a
    new device tensor."""
    import torch
    pc, keep = p.c()
    z = z.contiguous().to(torch.float64)
    out = [torch.empty_like(z) for _ in range(5)]
    _lib.check(_lib.lib.gust_inflow_eval(C.byref(pc), int(z.numel()), C.c_void_p(z.data_ptr()),
                                         *[C.c_void_p(o.data_ptr()) for o in out], _stream(stream)))
    del keep
    return tuple(out)


def initializer_poisson_bc(geom: Geometry, faces) -> tuple:
    g = geom.c()
    f = (C.c_int32 * 6)(*[int(x) for x in faces])
    out = (C.c_int32 * 6)()
    _lib.check(_lib.lib.gust_initializer_poisson_bc(C.byref(g), f, out))
    return tuple(out)


def obrien_omega(mesh: DeviceMesh, u: Fab, v: Fab, omega: Fab, stream=None) -> float:
    """Omega (z-faces) from the column continuity; returns max |Omega_top|
    before the removal."""
    g = mesh.geom.c()
    m, keep = _mesh_c(mesh)
    a = [u.array4(), v.array4(), omega.array4()]
    r = C.c_double(0)
    _lib.check(_lib.lib.gust_obrien_omega(C.byref(g), C.byref(m), *[C.byref(x) for x in a], C.byref(r),
                                          _stream(stream)))
    del keep
    return r.value


def first_guess_state(mesh: DeviceMesh, prof: InflowProfile, opt: InitOptions, q: DeviceState, stream=None) -> float:
    g = mesh.geom.c()
    m, keep = _mesh_c(mesh)
    pc, pkeep = prof.c()
    oc = opt.c()
    qc = q.c()
    r = C.c_double(0)
    _lib.check(_lib.lib.gust_first_guess_state(C.byref(g), C.byref(m), C.byref(pc), C.byref(oc), C.byref(qc),
                                               C.byref(r), _stream(stream)))
    del keep, pkeep
    return r.value


def initialize_state(mesh: DeviceMesh, prof: InflowProfile, opt: InitOptions, faces, solver: PoissonSolver | None,
                     q: DeviceState, lam: Fab | None = None, stream=None) -> InitResult:
    """The first guess and the correction by ``solver`` (a PoissonSolver on
    the same mesh; its options are the solver.poisson_* controls)."""
    g = mesh.geom.c()
    m, keep = _mesh_c(mesh)
    pc, pkeep = prof.c()
    oc = opt.c()
    qc = q.c()
    f = (C.c_int32 * 6)(*[int(x) for x in faces])
    la = lam.array4() if lam is not None else None
    res = _lib.InitResultC()
    h = solver._h if solver is not None else None
    _lib.check(_lib.lib.gust_initialize_state(C.byref(g), C.byref(m), C.byref(pc), C.byref(oc), f, h, C.byref(qc),
                                              C.byref(la) if la is not None else None, C.byref(res),
                                              _stream(stream)))
    del keep, pkeep
    p = res.poisson
    return InitResult(PoissonResult(p.iterations, p.initial_residual, p.final_residual, bool(p.singular),
                                    p.removed_mean), res.divergence_before, res.divergence_after,
                      res.obrien_top_residual)
