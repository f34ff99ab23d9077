"""Mass flux on the terrain-following C-grid (Source/MassFlux.H:49-76), over
device FABs: the contravariant flux Omega and its inverse, the transport
fluxes, the divergence the Poisson layer inverts, and the pressure gradient
(minus its adjoint). The mesh is a :class:`~ablsteadyrans_amd.poisson.DeviceMesh`.
Each call forwards to the C-ABI in ``libgust_amd.so``; ghost points the
operators read are the caller's to fill, as in the reference."""
from __future__ import annotations

import ctypes as C

from . import _lib
from .fab import Fab
from .poisson import DeviceMesh, _stream


def _r(*fabs):
    return [C.byref(f.array4()) for f in fabs]


def omega_from_w(mesh: DeviceMesh, u: Fab, v: Fab, w: Fab, omega: Fab, ngrow: int = 0, stream=None) -> None:
    """Source/MassFlux.cpp:51-74."""
    g = mesh.geom.c()
    _lib.check(_lib.lib.gust_omega_from_w(C.byref(g), *_r(u, v, w, mesh.z_nd, omega), int(ngrow), _stream(stream)))


def w_from_omega(mesh: DeviceMesh, u: Fab, v: Fab, omega: Fab, w: Fab, stream=None) -> None:
    """Source/MassFlux.cpp:76-99."""
    g = mesh.geom.c()
    _lib.check(_lib.lib.gust_w_from_omega(C.byref(g), *_r(u, v, omega, mesh.z_nd, w), _stream(stream)))


def transport_fluxes(mesh: DeviceMesh, u: Fab, v: Fab, w: Fab, F, stream=None) -> None:
    """Source/MassFlux.cpp:101-127; F = (Fx, Fy, Fz), each with one ghost."""
    g = mesh.geom.c()
    _lib.check(_lib.lib.gust_transport_fluxes(C.byref(g), *_r(u, v, w, mesh.z_nd, mesh.ax, mesh.ay, *F),
                                              _stream(stream)))


def velocity_divergence(mesh: DeviceMesh, u: Fab, v: Fab, w: Fab, div: Fab, stream=None) -> None:
    """Source/MassFlux.cpp:129-136."""
    g = mesh.geom.c()
    _lib.check(_lib.lib.gust_velocity_divergence(
        C.byref(g), *_r(u, v, w, mesh.z_nd, mesh.ax, mesh.ay, mesh.az, mesh.detJ, div), _stream(stream)))


def pressure_gradient(mesh: DeviceMesh, p: Fab, Gp, stream=None) -> None:
    """Source/MassFlux.cpp:138-179; Gp = (Gx, Gy, Gz) on the valid faces."""
    g = mesh.geom.c()
    _lib.check(_lib.lib.gust_pressure_gradient(C.byref(g), *_r(p, mesh.z_nd, mesh.detJ, mesh.ax, mesh.ay, *Gp),
                                               _stream(stream)))
