"""MI355X-native SIMPLEC hot path of GUST (hgopalan/ablsteadyrans).

The compute runs in ``libgust_amd.so`` (hand-written HIP for gfx950, C-ABI in
``include/gust_amd.h``); this package is the thin host layer over it.
"""
from . import _lib
from .fab import (CELL, XFACE, YFACE, ZFACE, LOCS, LOC_NAMES, N_COEFS, Fab, Geometry, LineSweepOptions,
                  ZJACOBI, LUSGS, LAGGED, ZERO, SEQUENTIAL, DIAGONAL)
from .line_sweeps import set_jacobi_schedule, JACOBI_AUTO, JACOBI_STREAMING, JACOBI_STAGED
from .line_sweeps import zline_jacobi, zline_lusgs, line_sweeps, check_line_options, LineSystem, \
    zline_jacobi_batched, JacobiPlan
from ._lib import GustError

__all__ = [
    "CELL", "XFACE", "YFACE", "ZFACE", "LOCS", "LOC_NAMES", "N_COEFS", "Fab", "Geometry", "LineSweepOptions",
    "ZJACOBI", "LUSGS", "LAGGED", "ZERO", "SEQUENTIAL", "DIAGONAL", "zline_jacobi", "zline_lusgs", "line_sweeps",
    "check_line_options", "LineSystem", "zline_jacobi_batched", "JacobiPlan", "GustError",
]
