"""The local pseudo-time step (Source/PseudoTime.H), its SER law and the face
means the momentum predictor uses (Source/Predictor.cpp:30-77), over device
FABs; each call forwards to the C-ABI in ``libgust_amd.so``."""
from __future__ import annotations

import ctypes as C

from . import _lib
from .fab import Fab
from .poisson import DeviceMesh, _stream

K_MIN_RATE = 1.0e-12   # Source/PseudoTime.H:27


def local_dtau(mesh: DeviceMesh, u: Fab, v: Fab, w: Fab, cfl: float, dtau: Fab, nu: Fab | None = None,
               stream=None) -> None:
    """Source/PseudoTime.cpp:12-51."""
    g = mesh.geom.c()
    nu_ref = C.byref(nu.array4()) if nu is not None else None
    _lib.check(_lib.lib.gust_local_dtau(C.byref(g), C.byref(u.array4()), C.byref(v.array4()), C.byref(w.array4()),
                                        C.byref(mesh.z_nd.array4()), C.byref(mesh.detJ.array4()), nu_ref,
                                        float(cfl), C.byref(dtau.array4()), _stream(stream)))


def ser_cfl(cfl: float, r: float, cfl0: float, cfl_max: float, p: float, growth: float) -> float:
    """Source/PseudoTime.cpp:53-61."""
    out = C.c_double(0)
    _lib.check(_lib.lib.gust_ser_cfl(cfl, r, cfl0, cfl_max, p, growth, C.byref(out)))
    return out.value


def face_dtau(geom, dtau_cc: Fab, faces, stream=None) -> None:
    """Source/Predictor.cpp:30-77; faces = (dtau_x, dtau_y, dtau_z)."""
    g = geom.c()
    _lib.check(_lib.lib.gust_face_dtau(C.byref(g), C.byref(dtau_cc.array4()), *[C.byref(f.array4()) for f in faces],
                                       _stream(stream)))
