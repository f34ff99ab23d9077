"""Host-side FAB and geometry types mirroring AMReX's FArrayBox/Array4 and Geometry.

A :class:`Fab` holds one box's data (valid points plus ghosts) as a contiguous
``(ncomp, nz, ny, nx)`` tensor, i fastest -- the AMReX Array4 layout, so the
C-ABI view is the tensor's pointer and strides with no copy. Device FABs are
torch tensors on ``cuda``; numpy arrays are accepted for host-side use.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ._lib import Array4, GeometryC

# Field locations (Source/CVFluxes.H:22).
CELL, XFACE, YFACE, ZFACE = 0, 1, 2, 4
LOCS = (CELL, XFACE, YFACE, ZFACE)
LOC_NAMES = {CELL: "cell", XFACE: "xface", YFACE: "yface", ZFACE: "zface"}

# Coefficient components (Source/ImplicitOperator.H:57).
COEF_P, COEF_XM, COEF_XP, COEF_YM, COEF_YP, COEF_ZM, COEF_ZP, N_COEFS = range(8)


@dataclass
class Geometry:
    """The cell-centred domain box (inclusive), x/y periodicity and spacing."""

    lo: tuple = (0, 0, 0)
    hi: tuple = (0, 0, 0)
    periodic: tuple = (0, 0, 0)
    dx: tuple = (1.0, 1.0, 1.0)

    @staticmethod
    def of_size(nx: int, ny: int, nz: int, periodic=(0, 0), dx=(1.0, 1.0, 1.0)) -> "Geometry":
        return Geometry((0, 0, 0), (nx - 1, ny - 1, nz - 1), (int(periodic[0]), int(periodic[1]), 0), tuple(dx))

    @property
    def n(self) -> tuple:
        return tuple(self.hi[d] - self.lo[d] + 1 for d in range(3))

    def loc_box(self, loc: int) -> tuple:
        """(lo, hi) inclusive of the domain converted to ``loc``."""
        lo = tuple(self.lo)
        hi = tuple(self.hi[d] + ((loc >> d) & 1) for d in range(3))
        return lo, hi

    def c(self) -> GeometryC:
        g = GeometryC()
        for d in range(3):
            g.lo[d] = int(self.lo[d])
            g.hi[d] = int(self.hi[d])
            g.periodic[d] = int(self.periodic[d]) if d < len(self.periodic) else 0
            g.dx[d] = float(self.dx[d])
        return g


class Fab:
    """One box of data: ``data`` is (ncomp, nz, ny, nx), ``begin`` its (i, j, k) origin."""

    def __init__(self, data, begin):
        if data.ndim != 4:
            raise ValueError("Fab data must be (ncomp, nz, ny, nx)")
        self.data = data
        self.begin = tuple(int(b) for b in begin)

    # -- construction -------------------------------------------------------
    @staticmethod
    def alloc(lo, hi, ncomp=1, ngrow=(0, 0, 0), *, dtype=None, device=None, backend="torch") -> "Fab":
        if isinstance(ngrow, int):
            ngrow = (ngrow, ngrow, ngrow)
        b = tuple(lo[d] - ngrow[d] for d in range(3))
        e = tuple(hi[d] + ngrow[d] + 1 for d in range(3))
        shape = (ncomp, e[2] - b[2], e[1] - b[1], e[0] - b[0])
        if backend == "numpy":
            return Fab(np.zeros(shape, dtype=dtype or np.float64), b)
        import torch
        return Fab(torch.zeros(shape, dtype=dtype or torch.float64, device=device or "cuda"), b)

    @staticmethod
    def at_loc(geom: Geometry, loc: int, ncomp=1, ngrow=0, **kw) -> "Fab":
        lo, hi = geom.loc_box(loc)
        return Fab.alloc(lo, hi, ncomp, ngrow, **kw)

    # -- geometry of the box ------------------------------------------------
    @property
    def end(self) -> tuple:
        return (self.begin[0] + self.data.shape[3], self.begin[1] + self.data.shape[2],
                self.begin[2] + self.data.shape[1])

    @property
    def ncomp(self) -> int:
        return int(self.data.shape[0])

    def region(self, lo, hi, comp=None):
        """View of the data over the inclusive box [lo, hi] (all comps or one)."""
        b = self.begin
        sl = (slice(lo[2] - b[2], hi[2] - b[2] + 1), slice(lo[1] - b[1], hi[1] - b[1] + 1),
              slice(lo[0] - b[0], hi[0] - b[0] + 1))
        if comp is None:
            return self.data[(slice(None),) + sl]
        return self.data[(comp,) + sl]

    # -- the C view ---------------------------------------------------------
    def array4(self) -> Array4:
        d = self.data
        if hasattr(d, "data_ptr"):
            if not d.is_contiguous():
                raise ValueError("Fab tensor must be contiguous")
            ptr = d.data_ptr()
            st = d.stride()
        else:
            if not d.flags["C_CONTIGUOUS"]:
                raise ValueError("Fab array must be C-contiguous")
            ptr = d.ctypes.data
            st = tuple(s // d.itemsize for s in d.strides)
        a = Array4()
        a.p = ptr
        a.nstride, a.kstride, a.jstride = int(st[0]), int(st[1]), int(st[2])
        e = self.end
        for q in range(3):
            a.begin[q] = self.begin[q]
            a.end[q] = e[q]
        a.ncomp = self.ncomp
        return a

    def to(self, *args, **kw) -> "Fab":
        return Fab(self.data.to(*args, **kw).contiguous(), self.begin)

    def numpy(self) -> np.ndarray:
        d = self.data
        if hasattr(d, "detach"):
            return d.detach().cpu().numpy()
        return d


@dataclass
class LineSweepOptions:
    """Source/LineSweeps.H:95-104 (kind, ghosts, overlap, order, alternate)."""

    kind: int = 0        # 0 zjacobi, 1 lusgs
    ghosts: int = 0      # 0 lagged, 1 zero
    overlap: int = 0
    order: int = 0       # 0 sequential, 1 diagonal
    alternate: int = 0

    def c(self):
        from ._lib import LineOptionsC
        o = LineOptionsC()
        o.kind, o.ghosts, o.overlap, o.order, o.alternate = (int(self.kind), int(self.ghosts), int(self.overlap),
                                                            int(self.order), int(self.alternate))
        return o


ZJACOBI, LUSGS = 0, 1
LAGGED, ZERO = 0, 1
SEQUENTIAL, DIAGONAL = 0, 1
MAX_LINE_OVERLAP = 8
