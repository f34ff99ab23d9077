"""Inputs of the reduction tests: the dyadic hash pattern of
Tests/Unit_Tests/TestHelpers.H:42-56 (every value mant 2^expo, exact in
binary) on the 24 x 20 x 16 mesh of test_Reductions.cpp:21-29, cut 8 x 8 in
x/y into 3 x 3 boxes, and the bits that test pins for component 3."""
from __future__ import annotations

import numpy as np

N_CELL = (24, 20, 16)
# x ranges and y ranges of the 3 x 3 boxes (the y boxes are 8, 8, 4)
XS = ((0, 7), (8, 15), (16, 23))
YS = ((0, 7), (8, 15), (16, 19))
# test_Reductions.cpp:37-42 and 68-76
SUM_BITS = 0xc12f9a008f38d09c
EXACT_SUM = float.fromhex("-0x1.f9a008f38d094p+19")


def pattern(i, j, k, c):
    """TestHelpers.H:42-56 on integer arrays (uint32 arithmetic)."""
    M = np.uint64(0xFFFFFFFF)
    i, j, k = (np.asarray(x).astype(np.int64).astype(np.uint64) & M for x in (i, j, k))
    h = ((i * np.uint64(73856093)) & M) ^ ((j * np.uint64(19349663)) & M) ^ ((k * np.uint64(83492791)) & M) \
        ^ np.uint64((c * 2654435761) & 0xFFFFFFFF)
    h = h ^ (h >> np.uint64(13))
    h = (h * np.uint64(0x5bd1e995)) & M
    h = h ^ (h >> np.uint64(15))
    mant = (h & np.uint64(0xFFFFF)).astype(np.float64)
    expo = ((h >> np.uint64(20)) % np.uint64(40)).astype(np.int64) - 40
    v = np.ldexp(mant, expo)
    return np.where((h >> np.uint64(31)) != 0, -v, v)


def pattern_field(c, grow=1, n=N_CELL):
    """(1, nz+2g, ny+2g, nx+2g) pattern values, ghosts included, begin (-g,)*3."""
    K, J, I = np.meshgrid(np.arange(-grow, n[2] + grow), np.arange(-grow, n[1] + grow),
                          np.arange(-grow, n[0] + grow), indexing="ij")
    return pattern(I, J, K, c)[None].copy()
