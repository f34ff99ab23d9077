"""The first-order implicit operator A1 (Source/ImplicitOperator.H) over
device FABs; each call forwards to ``libgust_amd.so``."""
from __future__ import annotations

import ctypes as C

from . import _lib
from .fab import Fab
from .poisson import DeviceMesh, _stream

ARITHMETIC, LOGARITHMIC = 0, 1   # FaceMean (Source/Diffusion_K.H:22)


def implicit_coefficients(loc: int, mesh: DeviceMesh, F, g: Fab, A: Fab, face_mean: int = ARITHMETIC,
                          stream=None) -> None:
    """Source/ImplicitOperator.cpp:95-121; F = (Fx, Fy, Fz) with one ghost."""
    geo = mesh.geom.c()
    v = [f.array4() for f in (*F, g, mesh.z_nd, mesh.detJ)]
    _lib.check(_lib.lib.gust_implicit_coefficients(int(loc), C.byref(geo), *[C.byref(a) for a in v], int(face_mean),
                                                   C.byref(A.array4()), _stream(stream)))


def apply_implicit(loc: int, geom, A: Fab, q: Fab, out: Fab, stream=None) -> None:
    """Source/ImplicitOperator.cpp:123-135."""
    g = geom.c()
    _lib.check(_lib.lib.gust_apply_implicit(int(loc), C.byref(g), C.byref(A.array4()), C.byref(q.array4()),
                                            C.byref(out.array4()), _stream(stream)))
