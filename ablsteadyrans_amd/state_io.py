"""The state's files (Source/StateIO.H) over the library's C-ABI: AMReX
plotfiles and the reference's checkpoints, so fields can be exchanged with
GUST runs on AMReX machines and a served request can restart bit for bit.

* :func:`write_plotfile` -- HyperCLaw-V1.1 plotfile: u, v, w (face means), p,
  theta, tke, z (cell-mean height) and optionally detJ, plus the node
  heights as Level_0/z_phys_nd (Source/StateIO.cpp:41-85);
* :func:`write_checkpoint` / :func:`read_checkpoint` -- the text Header
  (mesh in hex floats) and every field as a VisMF MultiFab with its ghosts
  (Source/StateIO.cpp:87-211); reading refuses another mesh, naming
  io.restart.
"""
from __future__ import annotations

import ctypes as C

from . import _lib


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _corners(mesh):
    t = mesh.tile
    return (C.c_double * 3)(*[float(x) for x in t.prob_lo]), (C.c_double * 3)(*[float(x) for x in t.prob_hi])


def write_plotfile(prefix: str, step: int, mesh, q, detJ: bool = False, max_grid_size: int = 0, stream=None) -> str:
    """Returns the plotfile's directory (prefix + step as five digits)."""
    g = mesh.geom.c()
    lo, hi = _corners(mesh)
    qc = q.c()
    z = mesh.z_nd.array4()
    J = mesh.detJ.array4() if detJ else None
    name = C.create_string_buffer(4096)
    _lib.check(_lib.lib.gust_write_plotfile(prefix.encode(), int(step), C.byref(g), lo, hi, C.byref(qc), C.byref(z),
                                            C.byref(J) if J is not None else None, int(max_grid_size), name,
                                            len(name), _stream(stream)))
    return name.value.decode()


def write_checkpoint(prefix: str, step: int, mesh, q, max_grid_size: int = 0, stream=None) -> str:
    """Returns the checkpoint's directory."""
    g = mesh.geom.c()
    lo, hi = _corners(mesh)
    qc = q.c()
    z = mesh.z_nd.array4()
    name = C.create_string_buffer(4096)
    _lib.check(_lib.lib.gust_write_checkpoint(prefix.encode(), int(step), C.byref(g), lo, hi, C.byref(qc),
                                              C.byref(z), int(max_grid_size), name, len(name), _stream(stream)))
    return name.value.decode()


def read_checkpoint(path: str, mesh, q, z_nd=None, stream=None) -> int:
    """Reads the checkpoint into q (and the node heights into z_nd, a Fab,
    when given); returns its step."""
    g = mesh.geom.c()
    lo, hi = _corners(mesh)
    qc = q.c()
    z = z_nd.array4() if z_nd is not None else None
    step = C.c_int32(0)
    _lib.check(_lib.lib.gust_read_checkpoint(path.encode(), C.byref(g), lo, hi, C.byref(qc),
                                             C.byref(z) if z is not None else None, C.byref(step), _stream(stream)))
    return step.value
