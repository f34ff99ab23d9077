"""Convection and diffusion over the C-grid's control volumes
(Source/Convection.H, Source/Diffusion.H, Source/CVFluxes.H) over device
FABs; each call forwards to ``libgust_amd.so``."""
from __future__ import annotations

import ctypes as C

from . import _lib
from .fab import Fab
from .poisson import DeviceMesh, _stream

VAN_LEER, VAN_ALBADA, UPWIND = 0, 1, 2   # Source/Limiter_K.H:19


def _r(*fabs):
    return [C.byref(f.array4()) for f in fabs]


def cv_flux_box(geom, loc: int, d: int):
    """(lo, hi) inclusive of the control-volume faces along d (Source/CVFluxes.cpp:54-67)."""
    lo, hi = list(geom.lo), [geom.hi[a] + ((loc >> a) & 1) for a in range(3)]
    if (loc >> d) & 1:
        lo[d], hi[d] = geom.lo[d] - 1, geom.hi[d] + 1
    else:
        hi[d] = geom.hi[d] + 1
    return tuple(lo), tuple(hi)


def alloc_cv_fluxes(geom, loc: int, **kw):
    return [Fab.alloc(*cv_flux_box(geom, loc, d), 1, 0, **kw) for d in range(3)]


def convective_fluxes(loc, mesh: DeviceMesh, q: Fab, F, G, limiter=VAN_LEER, smooth_eps=1e-5, stream=None):
    g = mesh.geom.c()
    _lib.check(_lib.lib.gust_convective_fluxes(int(loc), C.byref(g), *_r(q, *F), int(limiter), float(smooth_eps),
                                               *_r(*G), _stream(stream)))


def flux_divergence(loc, mesh: DeviceMesh, G, out: Fab, stream=None):
    g = mesh.geom.c()
    _lib.check(_lib.lib.gust_flux_divergence(int(loc), C.byref(g), *_r(*G, mesh.detJ, out), _stream(stream)))


def convection(loc, mesh: DeviceMesh, q: Fab, F, out: Fab, limiter=VAN_LEER, smooth_eps=1e-5, bounded=False,
               stream=None):
    g = mesh.geom.c()
    _lib.check(_lib.lib.gust_convection(int(loc), C.byref(g), *_r(q, *F, mesh.detJ), int(limiter), float(smooth_eps),
                                        int(bool(bounded)), C.byref(out.array4()), _stream(stream)))


def scalar_diffusive_fluxes(mesh: DeviceMesh, q: Fab, gamma: Fab, G, face_mean=0, stream=None):
    g = mesh.geom.c()
    _lib.check(_lib.lib.gust_scalar_diffusive_fluxes(C.byref(g), *_r(q, gamma, mesh.z_nd), int(face_mean), *_r(*G),
                                                     _stream(stream)))


def scalar_diffusion(mesh: DeviceMesh, q: Fab, gamma: Fab, out: Fab, face_mean=0, stream=None):
    g = mesh.geom.c()
    _lib.check(_lib.lib.gust_scalar_diffusion(C.byref(g), *_r(q, gamma, mesh.z_nd, mesh.detJ), int(face_mean),
                                              C.byref(out.array4()), _stream(stream)))


def momentum_diffusive_fluxes(c, mesh: DeviceMesh, vel, nu: Fab, G, face_mean=0, stream=None):
    g = mesh.geom.c()
    _lib.check(_lib.lib.gust_momentum_diffusive_fluxes(int(c), C.byref(g), *_r(*vel, nu, mesh.z_nd), int(face_mean),
                                                       *_r(*G), _stream(stream)))


def momentum_diffusion(mesh: DeviceMesh, vel, nu: Fab, out, face_mean=0, stream=None):
    g = mesh.geom.c()
    _lib.check(_lib.lib.gust_momentum_diffusion(C.byref(g), *_r(*vel, nu, mesh.z_nd, mesh.detJ), int(face_mean),
                                                *_r(*out), _stream(stream)))
