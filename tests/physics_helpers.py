"""Meshes and states of the residual tests: the oracle's Mesh from a host
TileMesh, states with two ghost layers (begin -2), the reference tests'
linear and constant theta profiles, and the flat neutral column of the
canonical cases (log-law wind along x, k from the wall law)."""
from __future__ import annotations

import numpy as np

from ablsteadyrans_amd import mesh as M
from oracle import oracle as orc
from oracle import physics as P


def omesh(tile) -> P.Mesh:
    g = tile.geom
    return P.Mesh(tile.z_nd[None].copy(), tile.detJ[None].copy(), tile.ax[None].copy(), tile.ay[None].copy(),
                  tile.az[None].copy(), orc.geom_c(g.lo, g.hi, g.periodic, g.dx))


def flat_mesh(n=(8, 6, 16), size=(800.0, 600.0), top=800.0, dz0=0.0, stretch=0.0):
    geom = M.geometry(n, (0.0, 0.0, 0.0), (size[0], size[1], top), (1, 1))
    zl = M.make_zlevels(n[2], top, dz0=dz0, stretch_ratio=stretch)
    return M.build_mesh(geom, (0.0, 0.0, 0.0), (size[0], size[1], top), zl, M.Terrain(kind="flat"))


def hill_mesh(n=(16, 12, 12), size=(1000.0, 800.0), top=600.0, height=120.0, half=180.0):
    geom = M.geometry(n, (0.0, 0.0, 0.0), (size[0], size[1], top), (1, 1))
    zl = M.make_zlevels(n[2], top, dz0=top / n[2] * 0.5, stretch_ratio=0.0)
    t = M.Terrain(kind="hills", hills=[(size[0] / 2, size[1] / 2, height, half)])
    return M.build_mesh(geom, (0.0, 0.0, 0.0), (size[0], size[1], top), zl, t)


def zeros_state(g):
    nx, ny, nz = (g.hi[d] - g.lo[d] + 1 for d in range(3))
    sh = {"u": (nz + 4, ny + 4, nx + 5), "v": (nz + 4, ny + 5, nx + 4), "w": (nz + 5, ny + 4, nx + 4),
          "p": (nz + 4, ny + 4, nx + 4), "theta": (nz + 4, ny + 4, nx + 4), "tke": (nz + 4, ny + 4, nx + 4)}
    return {k: np.zeros((1,) + s) for k, s in sh.items()}


def positions(tile, loc, grow=2):
    """(X, Y, Z) of every point at loc with `grow` ghosts: node means with the
    lateral node indices clamped and the levels extrapolated linearly."""
    nx, ny, nz = tile.geom.n
    dx, dy = tile.geom.dx[0], tile.geom.dx[1]
    nd = [(loc >> d) & 1 for d in range(3)]
    ii = np.arange(-grow, nx + nd[0] + grow)
    jj = np.arange(-grow, ny + nd[1] + grow)
    kk = np.arange(-grow, nz + nd[2] + grow)
    x = tile.prob_lo[0] + (ii + (0.0 if nd[0] else 0.5)) * dx
    y = tile.prob_lo[1] + (jj + (0.0 if nd[1] else 0.5)) * dy
    Z = tile.z_nd                          # [k+1, j+1, i+1], k = -1..nz+1
    e = grow
    ext = np.zeros((Z.shape[0] + 2 * e,) + Z.shape[1:])
    ext[e:e + Z.shape[0]] = Z
    for m in range(1, e + 1):
        ext[e - m] = Z[0] + m * (Z[0] - Z[1])
        ext[e + Z.shape[0] - 1 + m] = Z[-1] + m * (Z[-1] - Z[-2])
    zc = np.zeros((kk.size, jj.size, ii.size))
    cnt = 0
    for c in range(1 if nd[2] else 2):
        for b in range(1 if nd[1] else 2):
            for a in range(1 if nd[0] else 2):
                zc += ext[np.ix_(kk + c + 1 + e, np.clip(jj + b + 1, 0, Z.shape[1] - 1),
                                 np.clip(ii + a + 1, 0, Z.shape[2] - 1))]
                cnt += 1
    zc /= cnt
    X = np.broadcast_to(x[None, None, :], zc.shape).copy()
    Y = np.broadcast_to(y[None, :, None], zc.shape).copy()
    return X, Y, zc


def log_law_state(tile, ustar=0.4, z0=0.1, kappa=0.41, theta=300.0, k_floor=1e-6):
    """On flat ground at z = 0: the log-law wind along x, u = (u*/kappa)
    ln((z + z0)/z0), v = w = 0, uniform theta, k = u*^2 / Cmu0^2."""
    q = zeros_state(P_geom(tile))
    _, _, zu = positions(tile, 1)
    h = np.maximum(zu, 0.0)
    q["u"][0] = ustar / kappa * np.log((h + z0) / z0)
    q["theta"][:] = theta
    q["tke"][:] = max(ustar * ustar / 0.5562 ** 2, k_floor)
    return q


def P_geom(tile):
    g = tile.geom
    return orc.geom_c(g.lo, g.hi, g.periodic, g.dx)


def random_state(tile, seed, U=8.0, amp=0.5):
    """A smooth mean wind with random perturbations (valid and ghost points),
    theta stably stratified, k positive."""
    g = P_geom(tile)
    rng = np.random.default_rng(seed)
    q = zeros_state(g)
    for n, a in q.items():
        a[:] = rng.uniform(-1.0, 1.0, size=a.shape) * amp
    q["u"] += U
    q["v"] += 0.3 * U
    _, _, zc = positions(tile, 0)
    q["theta"][0] += 300.0 + 0.003 * zc
    q["tke"] = np.abs(q["tke"]) + 0.05
    # a periodic seam face is one unknown: the hi copy is the lo one (OverrideSync)
    nx, ny = g.hi[0] - g.lo[0] + 1, g.hi[1] - g.lo[1] + 1
    q["u"][0, :, :, 2 + nx] = q["u"][0, :, :, 2]
    q["v"][0, :, 2 + ny, :] = q["v"][0, :, 2, :]
    return q
