"""ctypes binding of the oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: the FP64 C restatement of the reference's hot-path
arithmetic, used by tests/, ``__graft_entry__.smoke()`` and bench.py's CPU
baseline as the checker. The product (``ablsteadyrans_amd``) never imports it.

Arrays are numpy float64, shape (ncomp, nz, ny, nx) (i fastest), each with a
``begin`` (i, j, k) origin -- the Array4 convention.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OA4(C.Structure):
    _fields_ = [
        ("p", C.c_void_p),
        ("jstride", C.c_int64),
        ("kstride", C.c_int64),
        ("nstride", C.c_int64),
        ("begin", C.c_int32 * 3),
        ("end", C.c_int32 * 3),
        ("ncomp", C.c_int32),
    ]


class OGeom(C.Structure):
    _fields_ = [
        ("lo", C.c_int32 * 3),
        ("hi", C.c_int32 * 3),
        ("periodic", C.c_int32 * 3),
        ("dx", C.c_double * 3),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_last_lusgs_launches.restype = C.c_uint64
        L.oracle_zline_jacobi.argtypes = [C.c_int, P(OA4), P(OA4), P(OA4), C.c_int, P(OGeom), P(OA4),
                                          P(C.c_double)]
        L.oracle_zline_lusgs.argtypes = [C.c_int, P(OA4), P(OA4), P(OA4), C.c_int, P(OGeom), P(OA4), C.c_int,
                                         C.c_int, C.c_int, C.c_int, P(C.c_double)]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def a4(arr: np.ndarray, begin) -> OA4:
    if arr.dtype != np.float64 or not arr.flags["C_CONTIGUOUS"] or arr.ndim != 4:
        raise ValueError("oracle arrays are C-contiguous float64 (ncomp, nz, ny, nx)")
    v = OA4()
    v.p = arr.ctypes.data
    v.nstride, v.kstride, v.jstride = (s // 8 for s in arr.strides[:3])
    for d in range(3):
        v.begin[d] = int(begin[d])
    v.end[0] = int(begin[0]) + arr.shape[3]
    v.end[1] = int(begin[1]) + arr.shape[2]
    v.end[2] = int(begin[2]) + arr.shape[1]
    v.ncomp = arr.shape[0]
    return v


def geom_c(lo, hi, periodic, dx=(1.0, 1.0, 1.0)) -> OGeom:
    g = OGeom()
    for d in range(3):
        g.lo[d], g.hi[d] = int(lo[d]), int(hi[d])
        g.periodic[d] = int(periodic[d]) if d < len(periodic) else 0
        g.dx[d] = float(dx[d])
    return g


def _check(rc):
    if rc != 0:
        raise OracleError(lib().oracle_last_error().decode())


def zline_jacobi(loc, A, A_begin, ap, rhs, v_begin, sweeps, geom, delta, d_begin):
    """Source/Predictor.cpp:87-127. Arrays are modified in place (delta)."""
    ch = C.c_double(0)
    _check(lib().oracle_zline_jacobi(loc, a4(A, A_begin), a4(ap, v_begin), a4(rhs, v_begin), sweeps, geom,
                                     a4(delta, d_begin), C.byref(ch)))
    return ch.value


def zline_lusgs(loc, A, A_begin, ap, rhs, v_begin, sweeps, geom, delta, d_begin, ghosts=0, overlap=0, order=0,
                alternate=0):
    """Source/LineSweeps.cpp:415-428."""
    ch = C.c_double(0)
    _check(lib().oracle_zline_lusgs(loc, a4(A, A_begin), a4(ap, v_begin), a4(rhs, v_begin), sweeps, geom,
                                    a4(delta, d_begin), ghosts, overlap, order, alternate, C.byref(ch)))
    return ch.value


def last_lusgs_launches() -> int:
    return int(lib().oracle_last_lusgs_launches())
