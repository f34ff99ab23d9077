"""The oracle's Jacobian-vector product (TEST INFRASTRUCTURE): the forward
difference of the oracle's full residual as Source/Residual.cpp:324-396
forms it, composed of pinned pieces (the scaled sums, the ghost fill, the
residual), for the checks of the device's jvp."""
from __future__ import annotations

import math

import numpy as np

from oracle import oracle as orc
from oracle import physics as P

FIELDS = ("u", "v", "w", "p", "theta", "tke")
LOCS = (1, 2, 4, 0, 0, 0)
STEP = math.sqrt(np.finfo(float).eps)
# JvpScales (Source/Residual.H:72-83): velocity, velocity, velocity, pressure, theta, tke
SCALES = {"u": 10.0, "v": 10.0, "w": 10.0, "p": 100.0, "theta": 10.0, "tke": 0.1}


def valid(a):
    return np.ascontiguousarray(a[:, 2:-2, 2:-2, 2:-2])


def jvp_eps(q, v, m, phys, step=STEP):
    """jvp's step (Residual.cpp:340-369): step (1 + mean|q~|) / rms(v~) over
    the six fields' owned points, theta about theta_ref; 0 for v = 0."""
    g = m.geom
    lo = tuple(g.lo)
    tref = np.ascontiguousarray(P.theta_ref_field(m, phys)[:, 1:-1, 1:-1, 1:-1])
    sums = [0.0, 0.0, 0.0]
    for n, loc in zip(FIELDS, LOCS):
        s = orc.scaled_sums(valid(q[n]), valid(v[n]), lo, SCALES[n], loc, g,
                            offset=tref if n == "theta" else None)
        for c in range(3):
            sums[c] += s[c]
    if sums[1] == 0:
        return 0.0
    return step * (1 + sums[0] / sums[2]) / math.sqrt(sums[1] / sums[2])


def jvp(q, v, m, phys, step=STEP, R0=None):
    """(eps, J v): q with its ghosts filled, v with its ghosts filled
    homogeneously; J v = (R(fill(q + eps v)) - R(q)) / eps on the valid
    points."""
    eps = jvp_eps(q, v, m, phys, step)
    if R0 is None:
        R0 = P.residual(q, m, phys)[0]
    if eps == 0:
        return 0.0, {n: np.zeros_like(R0[n]) for n in FIELDS}
    qp = {n: 1.0 * q[n] + eps * v[n] for n in FIELDS}
    P.fill_ghosts(qp, m, phys)
    R1 = P.residual(qp, m, phys)[0]
    inv = 1 / eps
    return eps, {n: (R1[n] - R0[n]) * inv for n in FIELDS}
