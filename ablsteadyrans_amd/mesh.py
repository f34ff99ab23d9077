"""Host-side construction of a tile's terrain-following mesh (built once per tile).
Derived in part from ERF (BSD 3-Clause, Copyright (c) 2022 The Regents of the
University of California et al.); the full notice is in THIRD_PARTY_NOTICES.md
at the repository root.

Restates the reference's mesh pipeline operation by operation (float64 numpy,
no contraction, so the node heights and metrics equal the CPU build's):

* z-levels: Source/ZLevels.cpp:66-126 (fixed ratio, or the geometric stretch
  of extern_erf/ERF_GridUtils.H:211-294 when only dz0 is given);
* ground heights: Source/TerrainSurface.cpp:302-417 (flat, Witch of Agnesi,
  ERF grid file with erf_bilinear 220-300) plus the analytic grounds of the
  benchmark configurations (Gaussian hills, a seeded fractal surface);
* basic terrain following (BTF): extern_erf/ERF_TerrainMetrics.cpp:94-242 and
  252-380, with Source/Mesh.cpp:112-150 (OverrideSync, FillBoundary) and the
  ghost-column fill of Source/Mesh.cpp:37-98;
* metrics: make_J / make_areas / make_zcc, extern_erf/ERF_TerrainMetrics.cpp:679-807.

Arrays are (nz, ny, nx) numpy float64 with an explicit ``begin`` (i, j, k)
origin, i fastest -- the Array4 layout. A tile is one box spanning its
domain. The mesh is an input to the device kernels; nothing here runs per
iteration.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .fab import Geometry

BOGUS = 1.0e30   # ERF's bogus_large_value (overwritten before any use)


def geometry(n_cell, prob_lo, prob_hi, periodic=(0, 0)) -> Geometry:
    """amrex::Geometry of Source/State.cpp:43-48: domain (0 .. n-1), cell size
    (prob_hi - prob_lo) / n."""
    dx = tuple((float(prob_hi[d]) - float(prob_lo[d])) / float(n_cell[d]) for d in range(3))
    return Geometry((0, 0, 0), tuple(int(n) - 1 for n in n_cell), (int(periodic[0]), int(periodic[1]), 0), dx)


# ---------------------------------------------------------------------------
# z-levels

def geometric_stretch(dz0: float, z_top: float, nz: int, tol: float):
    """extern_erf/ERF_GridUtils.H:211-260: Newton for the factor r with
    dz0 (r^nz - 1)/(r - 1) = z_top. Returns (dz0, ratio, converged)."""
    Nz = float(nz)
    if nz < 1 or not (z_top > 0.0) or not (dz0 > 0.0):
        return 0.0, 1.0, False
    if nz == 1 or dz0 >= z_top / Nz:
        return z_top / Nz, 1.0, True
    S = 1.03
    F = dz0 * ((S ** Nz - 1.0) / (S - 1.0)) - z_top
    it = 0
    while abs(F) > tol and it < 50:
        dF = dz0 * (Nz * S ** (Nz - 1.0) * (S - 1.0) - S ** Nz + 1.0) / (S - 1.0) ** 2.0
        S -= F / dF
        smin = 1.0 + tol
        S = S if S > smin else smin
        F = dz0 * ((S ** Nz - 1.0) / (S - 1.0)) - z_top
        it += 1
    return dz0, S, abs(F) <= tol


def make_zlevels(nz: int, prob_hi_z: float, dz0: float = 0.0, stretch_ratio: float = 0.0,
                 zlevels=None) -> np.ndarray:
    """Source/ZLevels.cpp:66-126: explicit levels, uniform (dz0 <= 0),
    geometric stretch to prob_hi z (ratio 0) or a fixed ratio from dz0."""
    if zlevels is not None:
        z = np.asarray(zlevels, dtype=np.float64)
        if z.size != nz + 1:
            raise ValueError(f"gust: zlevels has {z.size} values; n_cell z + 1 = {nz + 1} are needed")
        if z[0] != 0:
            raise ValueError("gust: zlevels: the first value must be 0")
        if not np.all(np.isfinite(z)) or not np.all(np.diff(z) > 0):
            raise ValueError("gust: zlevels: values must be finite and increase")
        return z.copy()
    z = np.zeros(nz + 1)
    if not dz0 > 0:
        dz = prob_hi_z / nz
        for k in range(nz + 1):
            z[k] = k * dz
        return z
    if stretch_ratio == 0:
        d0, r, ok = geometric_stretch(dz0, prob_hi_z, nz, 1.0e-8)
        if not ok:
            raise ValueError(f"gust: mesh.dz0 = {dz0}: no geometric stretch factor reaches prob_hi z "
                             "with n_cell z layers")
        dz = d0
        for k in range(1, nz + 1):
            z[k] = z[k - 1] + dz
            dz *= r
        z[nz] = prob_hi_z
        return z
    dz = dz0
    for k in range(1, nz + 1):
        z[k] = z[k - 1] + dz
        dz *= stretch_ratio
    return z


# ---------------------------------------------------------------------------
# Ground surfaces: functions of the node index arrays (ii, jj) already mapped
# to the node each ghost stands for, returning heights.

def node_index(i, lo, hi, periodic):
    """Source/TerrainSurface.cpp:55-60."""
    i = np.asarray(i)
    if not periodic:
        return np.minimum(np.maximum(i, lo), hi)
    n = hi - lo
    return lo + (((i - lo) % n) + n) % n


@dataclass
class ErfGrid:
    """An ERF terrain grid file (Source/TerrainSurface.cpp:187-218): nx, ny,
    x[nx], y[ny], z[nx*ny] contiguous in y."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray

    @staticmethod
    def read(path: str) -> "ErfGrid":
        v = np.loadtxt(path, dtype=np.float64).ravel()
        nx, ny = int(v[0]), int(v[1])
        if nx < 1 or ny < 1 or v.size != 2 + nx + ny + nx * ny:
            raise ValueError(f"gust: terrain.file = {path}: malformed ERF grid")
        return ErfGrid(v[2:2 + nx].copy(), v[2 + nx:2 + nx + ny].copy(), v[2 + nx + ny:].copy())

    def bilinear(self, x: float, y: float) -> float:
        """erf_bilinear, Source/TerrainSurface.cpp:220-300 (same branches and order)."""
        xt, yt, zt = self.x, self.y, self.z
        nx, ny = xt.size, yt.size
        jj = np.nonzero(yt > y)[0]
        jt = (jj[0] - 1) if jj.size else ny - 1
        jt = max(jt, 0)
        ii = np.nonzero(xt > x)[0]
        it = (ii[0] - 1) if ii.size else nx - 1
        it = max(it, 0)
        ip1 = min(it + 1, nx - 1)
        jp1 = min(jt + 1, ny - 1)
        x1, x2, y1, y2 = float(xt[it]), float(xt[ip1]), float(yt[jt]), float(yt[jp1])
        z11, z21 = float(zt[it * ny + jt]), float(zt[ip1 * ny + jt])
        z12, z22 = float(zt[it * ny + jp1]), float(zt[ip1 * ny + jp1])
        ix = (ip1 != it) and (x2 != x1)
        iy = (jp1 != jt) and (y2 != y1)
        if not ix and not iy:
            return z11
        if ix and not iy:
            return ((x2 - x) * z11 + (x - x1) * z21) / (x2 - x1)
        if not ix and iy:
            return ((y2 - y) * z11 + (y - y1) * z12) / (y2 - y1)
        w11 = (x2 - x) * (y2 - y)
        w21 = (x - x1) * (y2 - y)
        w12 = (x2 - x) * (y - y1)
        w22 = (x - x1) * (y - y1)
        den = (x2 - x1) * (y2 - y1)
        return (w11 * z11 + w12 * z12 + w21 * z21 + w22 * z22) / den


@dataclass
class Terrain:
    """A ground description: ``kind`` in flat | agnesi | erf_grid | xyz | hills | fractal."""

    kind: str = "flat"
    # agnesi (Source/TerrainSurface.cpp:329-386)
    hill_height: float = 0.0
    hill_half_width: float = 1.0
    hill_shape: str = "ridge_x"       # ridge_x | ridge_y | round (the reference's default: ridge_x)
    # erf_grid
    grid: ErfGrid | None = None
    # hills: list of (xc, yc, H, L): h = sum H exp(-((x-xc)^2 + (y-yc)^2) / L^2)
    hills: list = field(default_factory=list)
    # fractal: a precomputed (ny+1, nx+1) node height array
    heights: np.ndarray | None = None
    # xyz: scattered points (x, y, z arrays), inverse-distance weighted
    points: tuple | None = None
    idw_neighbors: int = 6
    idw_power: float = 2.0


def read_xyz_points(path: str, key: str = "terrain.file"):
    """Source/TerrainSurface.cpp:64-99: one "x y z" per line, comma or white
    space separated, '#' comments; lines without three numbers are skipped,
    points with a non-finite value are dropped. Returns (x, y, z, nonfinite)."""
    xs, ys, zs, bad = [], [], [], 0
    try:
        lines = open(path).read().split("\n")
    except OSError:
        raise ValueError(f"gust: {key} = {path}: cannot be read") from None
    for line in lines:
        tok = line.split("#", 1)[0].replace(",", " ").split()
        if len(tok) != 3:
            continue
        try:
            x, y, z = (float(t) for t in tok)
        except ValueError:
            continue
        if np.isfinite(x) and np.isfinite(y) and np.isfinite(z):
            xs.append(x)
            ys.append(y)
            zs.append(z)
        else:
            bad += 1
    if not xs:
        raise ValueError(f"gust: {key} = {path}: holds no 'x y z' point")
    return np.array(xs), np.array(ys), np.array(zs), bad


def idw_heights(points, xq: np.ndarray, yq: np.ndarray, k: int = 6, power: float = 2.0) -> np.ndarray:
    """Source/TerrainSurface.cpp:136-187: the inverse-distance-weighted height
    at each query from its k nearest points (squared distances dx^2 + dy^2,
    ties broken by point index), weights d^-power summed nearest first; a
    point closer than 1e-6 m gives its own height."""
    px, py, pz = (np.asarray(a, dtype=np.float64) for a in points[:3])
    n = px.size
    if not 1 <= k <= n:
        raise ValueError(f"gust: terrain.idw_neighbors = {k} must be between 1 and the number of points in "
                         f"terrain.file ({n})")
    xq = np.asarray(xq, dtype=np.float64).ravel()
    yq = np.asarray(yq, dtype=np.float64).ravel()
    out = np.empty(xq.size)
    chunk = max(1, int(4_000_000 // max(n, 1)))
    for s0 in range(0, xq.size, chunk):
        qx, qy = xq[s0:s0 + chunk, None], yq[s0:s0 + chunk, None]
        dx = px[None, :] - qx
        dy = py[None, :] - qy
        d2 = dx * dx + dy * dy
        # the k smallest by (d2, index): a stable sort on d2 keeps index order within ties
        order = np.argsort(d2, axis=1, kind="stable")[:, :k]
        dk = np.take_along_axis(d2, order, axis=1)
        zk = pz[order]
        wsum = np.zeros(dk.shape[0])
        zval = np.zeros(dk.shape[0])
        with np.errstate(divide="ignore", invalid="ignore"):   # exact hits are replaced below
            for i in range(k):
                w = np.power(dk[:, i], -power / 2.0)
                wsum = wsum + w
                zval = zval + w * zk[:, i]
            res = zval / wsum
        hit = dk[:, 0] < 1.0e-12
        res[hit] = zk[hit, 0]
        out[s0:s0 + chunk] = res
    return out


def ground_heights(t: Terrain, geom: Geometry, prob_lo, prob_hi) -> np.ndarray:
    """Ground node heights on the nodal slab grown by one ghost node,
    shape (ny + 3, nx + 3), index (j + 1, i + 1)."""
    nx, ny, _ = geom.n
    dx, dy = geom.dx[0], geom.dx[1]
    i = np.arange(-1, nx + 2)
    j = np.arange(-1, ny + 2)
    perx, pery = bool(geom.periodic[0]), bool(geom.periodic[1])
    out = np.zeros((ny + 3, nx + 3))
    if t.kind == "flat":
        return out
    if t.kind == "agnesi":
        hm, L = float(t.hill_height), float(t.hill_half_width)
        xcen = 0.5 * (prob_lo[0] + prob_hi[0])
        ycen = 0.5 * (prob_lo[1] + prob_hi[1])

        def per_idx(a, n, per):
            if per:
                return np.where(a < 0, a + n, np.where(a > n, a - n, a))
            return np.minimum(np.maximum(a, 0), n)

        ii = per_idx(i, nx, perx)
        jj = per_idx(j, ny, pery)
        X = prob_lo[0] + ii * dx - xcen
        Y = prob_lo[1] + jj * dy - ycen
        if t.hill_shape == "ridge_x":
            xl = X / L
            return np.broadcast_to(hm / (1 + xl * xl), out.shape).copy()
        if t.hill_shape == "ridge_y":
            yl = Y / L
            return np.broadcast_to((hm / (1.0 + yl * yl))[:, None], out.shape).copy()
        if t.hill_shape != "round":
            raise ValueError(f"gust: terrain.hill_shape must be ridge_x, ridge_y or round (got {t.hill_shape})")
        XX, YY = np.meshgrid(X, Y)
        r = np.sqrt(XX * XX + YY * YY)
        rl = r / L
        return hm / (1.0 + rl * rl)
    ii = node_index(i, 0, nx, perx)
    jj = node_index(j, 0, ny, pery)
    if t.kind == "erf_grid":
        for b, jv in enumerate(jj):
            y = float(prob_lo[1]) + int(jv) * float(dy)
            for a, iv in enumerate(ii):
                x = float(prob_lo[0]) + int(iv) * float(dx)
                out[b, a] = t.grid.bilinear(x + 1e-3, y + 1e-3)
        return out
    if t.kind == "hills":
        # along a periodic side each hill repeats with the period (its
        # images one period either way), so the ground is continuous across
        # the seam instead of stepping where a hill meets it
        X = prob_lo[0] + ii * dx
        Y = prob_lo[1] + jj * dy
        XX, YY = np.meshgrid(X, Y)
        h = np.zeros_like(XX)
        sx = (-nx * dx, 0.0, nx * dx) if perx else (0.0,)
        sy = (-ny * dy, 0.0, ny * dy) if pery else (0.0,)
        for (xc, yc, H, L) in t.hills:
            for ox in sx:
                for oy in sy:
                    h = h + H * np.exp(-((XX - xc - ox) ** 2 + (YY - yc - oy) ** 2) / (L * L))
        return h
    if t.kind == "xyz":
        if t.points is None:
            raise ValueError("gust: terrain.type = xyz needs the points (read_xyz_points)")
        X = float(prob_lo[0]) + ii * float(dx)
        Y = float(prob_lo[1]) + jj * float(dy)
        XX, YY = np.meshgrid(X, Y)
        return idw_heights(t.points, XX, YY, t.idw_neighbors, t.idw_power).reshape(XX.shape)
    if t.kind == "fractal":
        hv = np.asarray(t.heights, dtype=np.float64)
        if hv.shape != (ny + 1, nx + 1):
            raise ValueError("fractal heights must be (ny + 1, nx + 1) nodes")
        return hv[np.ix_(jj, ii)]
    raise ValueError(f"unknown terrain kind {t.kind}")


def fractal_heights(nx: int, ny: int, seed: int, rms: float, slope: float = -2.0, clip=(0.0, 400.0)):
    """A seeded isotropic Gaussian random field on the (ny+1, nx+1) nodes with
    power spectrum |k|^slope (amplitude |k|^(slope/2)), zero mean, the given
    rms, clipped (the fractal-terrain configuration)."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((ny + 1, nx + 1))
    W = np.fft.fft2(w)
    ky = np.fft.fftfreq(ny + 1)[:, None]
    kx = np.fft.fftfreq(nx + 1)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    amp = np.zeros_like(k)
    amp[k > 0] = k[k > 0] ** (slope / 2.0)
    h = np.real(np.fft.ifft2(W * amp))
    h -= h.mean()
    h *= rms / np.sqrt(np.mean(h * h))
    return np.clip(h, clip[0], clip[1])


# ---------------------------------------------------------------------------
# The mesh.

@dataclass
class TileMesh:
    geom: Geometry
    prob_lo: tuple
    prob_hi: tuple
    zlevels: np.ndarray
    z_nd: np.ndarray        # (nz+3, ny+3, nx+3), begin (-1, -1, -1)
    detJ: np.ndarray        # (nz+2, ny+2, nx+2), begin (-1, -1, -1)
    ax: np.ndarray          # (nz+2, ny+2, nx+3), begin (-1, -1, -1)
    ay: np.ndarray          # (nz+2, ny+3, nx+2)
    az: np.ndarray          # (nz+3, ny+2, nx+2)
    z_cc: np.ndarray        # (nz+2, ny+2, nx+2)

    BEGIN = (-1, -1, -1)

    def valid(self, name: str) -> np.ndarray:
        """The valid points of a metric (no ghosts), (nz', ny', nx')."""
        a = getattr(self, name)
        if name == "z_nd":
            return a[1:-1, 1:-1, 1:-1]
        return a[1:-1, 1:-1, 1:-1]


@dataclass
class TileFrame:
    """What build_mesh takes, without the host arrays: DeviceMesh builds the
    mesh of a TileFrame on the device (gust_build_mesh)."""
    geom: Geometry
    prob_lo: tuple
    prob_hi: tuple
    zlevels: np.ndarray
    terrain: Terrain
    z_nd = None   # no host arrays


def tile_frame(geom: Geometry, prob_lo, prob_hi, zlevels: np.ndarray, terrain: Terrain) -> TileFrame:
    if zlevels.size != geom.n[2] + 1:
        raise ValueError("zlevels must have nz + 1 values")
    return TileFrame(geom, tuple(prob_lo), tuple(prob_hi), np.asarray(zlevels, dtype=np.float64).copy(), terrain)


def _stf_levels(Z, zlevels, nx, ny, nz, perx, pery, gamma_m):
    """The smoothed terrain-following levels (STF, Klemp 2011), k = 1 .. nz,
    as extern_erf/ERF_TerrainMetrics.cpp:384-573 computes them on the node
    slab grown by one ghost node (Z[1] holds the ground there): the
    smoothed surface h_s of level k from level k-1's by up to 50 Laplacian
    passes with beta_k = 0.2 min(z_k / (2 h_m), 1), each pass accepted only
    while the minimum layer thickness ratio stays >= gamma_m (a rejected
    first pass carries level k-1's surface up unsmoothed); z = z_k + A_k h_s
    with A = cos^6(pi z / (2 z_H)) below z_H = 2.44 h_m / (1 - gamma_m).
    Returns the number of levels whose smoothing was capped."""
    # the ground of the grown slab: ghost columns take the clamped node
    ic = np.minimum(np.maximum(np.arange(-1, nx + 2), 0), nx) + 1
    jc = np.minimum(np.maximum(np.arange(-1, ny + 2), 0), ny) + 1
    zs = Z[1][np.ix_(jc, ic)].copy()
    Z[1] = zs
    h_m = float(zs.max())
    if h_m < np.finfo(np.float64).eps:
        h_m = 1e-16

    def fill(h):
        # FillBoundary of a nodal slab with one ghost node
        if perx:
            h[:, 0] = h[:, nx]
            h[:, nx + 2] = h[:, 2]
        if pery:
            h[0, :] = h[ny, :]
            h[ny + 2, :] = h[2, :]

    H = np.zeros((nz + 1, ny + 3, nx + 3))   # h_s by level
    H[0] = zs
    fill(H[0])
    Hold = H.copy()
    z_H = 2.44 * h_m / (1 - gamma_m)
    capped = 0
    # clipped neighbour indices of every slab node (ii +- 1, jj +- 1 read ghosts)
    ii = np.minimum(np.maximum(np.arange(-1, nx + 2), 0), nx) + 1
    jj = np.minimum(np.maximum(np.arange(-1, ny + 2), 0), ny) + 1
    I, J = np.meshgrid(ii, jj)
    for k in range(1, nz + 1):
        zz, zzm = float(zlevels[k]), float(zlevels[k - 1])
        foo = np.cos((np.pi / 2) * (zz / z_H))
        A = foo * foo * foo * foo * foo * foo if zz < z_H else 0.0
        foom = np.cos((np.pi / 2) * (zzm / z_H))
        Am = foom * foom * foom * foom * foom * foom if zzm < z_H else 0.0
        beta = 0.2 * min(zz / (2 * h_m), 1.0)
        it = 0
        while it < 50:
            src = Hold[k - 1] if it == 0 else Hold[k]
            lap = src[J, I + 1] + src[J, I - 1] + src[J + 1, I] + src[J - 1, I] - 4 * src[J, I]
            H[k] = src + beta * lap
            diff = float(np.min((zz + A * H[k] - (zzm + Am * H[k - 1])) / (zz - zzm)))
            if diff < gamma_m:
                capped += 1
                if it == 0:
                    H[k] = H[k - 1]
                    Hold[:] = H
                    fill(Hold[k])
                else:
                    H[:] = Hold
                break
            Hold[:] = H
            it += 1
            fill(Hold[k])
        Z[k + 1] = zz + A * Hold[k]
        if k == 1:
            Z[0] = 2.0 * Z[1] - Z[2]
    return capped


def _sullivan_levels(Z, zlevels, nx, ny, nz):
    """Sullivan et al. (2014) levels, z = z_k + (1 - z_k / z_top)^3 h
    (extern_erf/ERF_TerrainMetrics.cpp:596-630) on the grown slab, h the
    clamped column's ground."""
    ic = np.minimum(np.maximum(np.arange(-1, nx + 2), 0), nx) + 1
    jc = np.minimum(np.maximum(np.arange(-1, ny + 2), 0), ny) + 1
    zs = Z[1][np.ix_(jc, ic)].copy()
    z_top = float(zlevels[nz])
    for k in range(0, nz + 1):
        z = float(zlevels[k])
        Z[k + 1] = z + np.power(1.0 - (z / z_top), 3.0) * zs
    Z[1] = zs
    Z[0] = 2.0 * zs - Z[2]


def build_mesh(geom: Geometry, prob_lo, prob_hi, zlevels: np.ndarray, terrain: Terrain, smoothing: int = 0,
               gamma_m: float = 0.5, ground: np.ndarray | None = None) -> TileMesh:
    """The terrain-following mesh of one tile (Source/Mesh.cpp:107-140):
    terrain.smoothing = 0 BTF, 1 STF (gamma_m: the minimum layer thickness
    ratio), 2 Sullivan. The STF capped-level count is mesh.stf_capped_levels.
    ground: the node heights of the ground slab grown by one ghost node
    ((ny + 3, nx + 3), index (j + 1, i + 1)) in place of the terrain's
    (Mesh::build_from_ground, Source/Mesh.cpp:118-151)."""
    if smoothing not in (0, 1, 2):
        raise ValueError("gust: terrain.smoothing must be 0 (BTF), 1 (STF) or 2 (Sullivan)")
    if not (np.isfinite(gamma_m) and 0 < gamma_m < 1):
        raise ValueError("gust: terrain.gamma_m must be in (0, 1)")
    nx, ny, nz = geom.n
    if zlevels.size != nz + 1:
        raise ValueError("zlevels must have nz + 1 values")
    perx, pery = bool(geom.periodic[0]), bool(geom.periodic[1])
    dzeta = geom.dx[2]
    # Z[k+1, j+1, i+1] = z_nd(i, j, k), i in [-1, nx+1], k in [-1, nz+1]
    Z = np.zeros((nz + 3, ny + 3, nx + 3))
    if ground is not None:
        if ground.shape != (ny + 3, nx + 3):
            raise ValueError("the ground slab must be (ny + 3, nx + 3) nodes")
        Z[1] = ground
    else:
        Z[1] = ground_heights(terrain, geom, prob_lo, prob_hi)
    # setDomainBndry(bogus): nodes outside the domain along non-periodic axes
    Z[0] = BOGUS
    Z[nz + 2] = BOGUS
    if not perx:
        Z[2:nz + 2, :, 0] = BOGUS
        Z[2:nz + 2, :, nx + 2] = BOGUS
    if not pery:
        Z[2:nz + 2, 0, :] = BOGUS
        Z[2:nz + 2, ny + 2, :] = BOGUS
    z_top = float(zlevels[nz])
    capped = 0
    if smoothing == 1:
        capped = _stf_levels(Z, zlevels, nx, ny, nz, perx, pery, gamma_m)
        _fill_boundary_nodal(Z, nx, ny, perx, pery)
    elif smoothing == 2:
        _sullivan_levels(Z, zlevels, nx, ny, nz)
        _fill_boundary_nodal(Z, nx, ny, perx, pery)
    else:
        # BTF (extern_erf/ERF_TerrainMetrics.cpp:330-356): the clamped column's
        # ground for every column of the grown box, k = 0 .. nz.
        ic = np.minimum(np.maximum(np.arange(-1, nx + 2), 0), nx) + 1
        jc = np.minimum(np.maximum(np.arange(-1, ny + 2), 0), ny) + 1
        zs = Z[1][np.ix_(jc, ic)].copy()
        z_lev_sfc = float(zlevels[0])
        Z[1] = zs
        for k in range(1, nz + 1):
            z = float(zlevels[k])
            Z[k + 1] = ((zs - z_lev_sfc) * z_top + (z_top - zs) * z) / (z_top - z_lev_sfc)
        _fill_boundary_nodal(Z, nx, ny, perx, pery)
        Z[0] = 2.0 * Z[1] - Z[2]
        _fill_boundary_nodal(Z, nx, ny, perx, pery)
    Z[nz + 2] = Z[nz + 1] + 1 * (Z[nz + 1] - Z[nz])
    # Source/Mesh.cpp:129-133: OverrideSync, FillBoundary, ghost columns.
    if perx:
        Z[:, :, nx + 1] = Z[:, :, 1]
    if pery:
        Z[:, ny + 1, :] = Z[:, 1, :]
    _fill_boundary_nodal(Z, nx, ny, perx, pery)
    _fill_ghost_columns(Z, nx, ny, nz, perx, pery)
    # metrics (extern_erf/ERF_TerrainMetrics.cpp:679-807)
    dzInv = 1.0 / dzeta
    detJ = np.zeros((nz + 2, ny + 2, nx + 2))
    # cells i in [-1, nx], k in [0, nz]: Z index offsets +1
    a = Z[2:nz + 3, 0:ny + 2, 0:nx + 2]
    b = Z[2:nz + 3, 0:ny + 2, 1:nx + 3]
    c = Z[2:nz + 3, 1:ny + 3, 0:nx + 2]
    d = Z[2:nz + 3, 1:ny + 3, 1:nx + 3]
    e = Z[1:nz + 2, 0:ny + 2, 0:nx + 2]
    f = Z[1:nz + 2, 0:ny + 2, 1:nx + 3]
    g = Z[1:nz + 2, 1:ny + 3, 0:nx + 2]
    h = Z[1:nz + 2, 1:ny + 3, 1:nx + 3]
    detJ[1:] = 0.25 * dzInv * (a + b + c + d - e - f - g - h)
    _fill_boundary_cell(detJ, nx, ny, perx, pery, nodal=(0, 0))
    ax = np.zeros((nz + 2, ny + 2, nx + 3))
    ax[1:] = 0.5 * dzInv * (Z[2:nz + 3, 0:ny + 2, :] + Z[2:nz + 3, 1:ny + 3, :]
                            - Z[1:nz + 2, 0:ny + 2, :] - Z[1:nz + 2, 1:ny + 3, :])
    ay = np.zeros((nz + 2, ny + 3, nx + 2))
    ay[1:] = 0.5 * dzInv * (Z[2:nz + 3, :, 0:nx + 2] + Z[2:nz + 3, :, 1:nx + 3]
                            - Z[1:nz + 2, :, 0:nx + 2] - Z[1:nz + 2, :, 1:nx + 3])
    _fill_boundary_cell(ax, nx, ny, perx, pery, nodal=(1, 0))
    _fill_boundary_cell(ay, nx, ny, perx, pery, nodal=(0, 1))
    az = np.ones((nz + 3, ny + 2, nx + 2))
    zcc = np.zeros((nz + 2, ny + 2, nx + 2))
    zcc[:] = 0.125 * (Z[0:nz + 2, 0:ny + 2, 0:nx + 2] + Z[0:nz + 2, 0:ny + 2, 1:nx + 3]
                      + Z[0:nz + 2, 1:ny + 3, 0:nx + 2] + Z[0:nz + 2, 1:ny + 3, 1:nx + 3]
                      + Z[1:nz + 3, 0:ny + 2, 0:nx + 2] + Z[1:nz + 3, 0:ny + 2, 1:nx + 3]
                      + Z[1:nz + 3, 1:ny + 3, 0:nx + 2] + Z[1:nz + 3, 1:ny + 3, 1:nx + 3])
    _fill_boundary_cell(zcc, nx, ny, perx, pery, nodal=(0, 0))
    mesh = TileMesh(geom, tuple(prob_lo), tuple(prob_hi), zlevels.copy(), Z, detJ, ax, ay, az, zcc)
    mesh.stf_capped_levels = capped
    ground = Z[1, 1:ny + 2, 1:nx + 2]
    if not float(ground.max()) < z_top:
        raise ValueError(f"gust: the terrain reaches {float(ground.max())} m, at or above the top level {z_top} m")
    th = np.diff(Z[1:nz + 2, 1:ny + 2, 1:nx + 2], axis=0)
    if not np.all(th > 0):
        raise ValueError("gust: the terrain-following mesh has inverted or empty cells")
    return mesh


def _fill_boundary_nodal(Z, nx, ny, perx, pery):
    """FillBoundary of a nodal array with one ghost node (one box): a ghost
    node along a periodic axis takes its image (node -1 <- n-1, n+1 <- 1)."""
    if perx:
        Z[:, :, 0] = Z[:, :, nx]
        Z[:, :, nx + 2] = Z[:, :, 2]
    if pery:
        Z[:, 0, :] = Z[:, ny, :]
        Z[:, ny + 2, :] = Z[:, 2, :]


def _fill_boundary_cell(a, nx, ny, perx, pery, nodal=(0, 0)):
    """FillBoundary of a one-ghost array at a cell or face location (one box)."""
    if perx:
        if nodal[0]:
            a[..., 0] = a[..., nx]
            a[..., nx + 2] = a[..., 2]
        else:
            a[..., 0] = a[..., nx]
            a[..., nx + 1] = a[..., 1]
    if pery:
        if nodal[1]:
            a[..., 0, :] = a[..., ny, :]
            a[..., ny + 2, :] = a[..., 2, :]
        else:
            a[..., 0, :] = a[..., ny, :]
            a[..., ny + 1, :] = a[..., 1, :]


def _fill_ghost_columns(Z, nx, ny, nz, perx, pery):
    """Source/Mesh.cpp:37-98: a node outside a non-periodic face takes its
    row's edge node (k = 0..nz); every column's rows below the ground and
    above the top are recomputed from it."""
    if not perx:
        Z[1:nz + 2, :, 0] = Z[1:nz + 2, :, 1]
        Z[1:nz + 2, :, nx + 2] = Z[1:nz + 2, :, nx + 1]
    if not pery:
        Z[1:nz + 2, 0, :] = Z[1:nz + 2, 1, :]
        Z[1:nz + 2, ny + 2, :] = Z[1:nz + 2, ny + 1, :]
    Z[0] = 2.0 * Z[1] - Z[2]
    Z[nz + 2] = Z[nz + 1] + 1 * (Z[nz + 1] - Z[nz])


# The horizontal grid hierarchy (Source/MeshHierarchy.H/.cpp,
# Source/SmoothSurface.H/.cpp): levels l = 0 .. L-1, level l has 2^l times
# fewer cells in x and y and the same z-levels; its ground is level 0's at
# its nodes, smoothed on its own grid by passes (2^l - 1) passes of
# h += beta (sum of the four neighbours - 4 h) with beta = 1/8; its mesh is
# built from that ground with the run's smoothing.

MAX_HIERARCHY_LEVELS = 10
MAX_HIERARCHY_PASSES = 64
HIERARCHY_BETA = 0.125


def hierarchy_ratio(level: int) -> int:
    """2^level (Source/MeshHierarchy.cpp:56-63)."""
    if level < 0 or level >= MAX_HIERARCHY_LEVELS:
        raise ValueError(f"gust: hierarchy level {level} outside [0, {MAX_HIERARCHY_LEVELS - 1}] "
                         "(mesh.hierarchy_levels)")
    return 1 << level


def ground_passes(level: int, passes_per_level: int) -> int:
    """passes_per_level (2^l - 1) (Source/MeshHierarchy.cpp:65-68)."""
    return passes_per_level * (hierarchy_ratio(level) - 1)


def check_hierarchy_inputs(n_cell, levels: int, passes: int = 1, passes_given: bool = False) -> list:
    """Source/MeshHierarchy.cpp:70-111 for one box: each problem names its key."""
    err = []
    L = int(levels)
    nx, ny = int(n_cell[0]), int(n_cell[1])
    if L < 1 or L > MAX_HIERARCHY_LEVELS:
        err.append(f"mesh.hierarchy_levels must be an integer from 1 to {MAX_HIERARCHY_LEVELS} (got {L})")
    elif L > 1:
        r = 1 << (L - 1)
        if nx >= 1 and ny >= 1 and (nx % r != 0 or ny % r != 0):
            err.append(f"mesh.hierarchy_levels = {L} needs mesh.n_cell x and y divisible by 2^{L - 1} = {r} "
                       f"(got {nx} {ny})")
        if nx >= 1 and ny >= 1 and nx % r == 0 and ny % r == 0 and nx // r == 1 and ny // r == 1:
            err.append(f"mesh.hierarchy_levels = {L} leaves the coarsest level 1 x 1 cells (mesh.n_cell {nx} {ny}); "
                       "use fewer levels")
    if passes < 0 or passes > MAX_HIERARCHY_PASSES:
        err.append(f"terrain.hierarchy_passes must be an integer from 0 to {MAX_HIERARCHY_PASSES} (got {passes})")
    if passes_given and L == 1:
        err.append("terrain.hierarchy_passes has no effect unless mesh.hierarchy_levels > 1")
    return err


def fill_surface_ghosts(h: np.ndarray, nx: int, ny: int, perx: bool, pery: bool) -> None:
    """The ghost nodes of a ground slab ((ny + 3, nx + 3), index (j + 1, i +
    1)): periodic images, else the nearest domain node
    (Source/SmoothSurface.cpp:50-77)."""
    _fill_slab_images(h, nx, ny, perx, pery)
    if not perx:
        h[:, 0] = h[:, 1]
        h[:, nx + 2] = h[:, nx + 1]
    if not pery:
        h[0, :] = h[1, :]
        h[ny + 2, :] = h[ny + 1, :]


def _fill_slab_images(h, nx, ny, perx, pery):
    if perx:
        h[:, 0] = h[:, nx]
        h[:, nx + 2] = h[:, 2]
    if pery:
        h[0, :] = h[ny, :]
        h[ny + 2, :] = h[2, :]


def smooth_surface(h: np.ndarray, nx: int, ny: int, perx: bool, pery: bool, passes: int, beta: float) -> None:
    """passes passes of h += beta (h(i+1) + h(i-1) + h(j+1) + h(j-1) - 4 h) on
    the valid nodes of a ground slab, in place; across a non-periodic face
    the neighbour is the edge node itself; the ghosts refilled after
    (Source/SmoothSurface.cpp:9-48)."""
    if passes < 0 or not np.isfinite(beta) or not (0 <= beta < 0.25):
        raise ValueError("gust: smooth_surface needs passes >= 0 and 0 <= beta < 1/4")
    if passes == 0:
        return
    i = np.arange(0, nx + 1)
    j = np.arange(0, ny + 1)
    im = np.where((~np.asarray(perx)) & (i == 0), i, i - 1) + 1
    ip = np.where((~np.asarray(perx)) & (i == nx), i, i + 1) + 1
    jm = np.where((~np.asarray(pery)) & (j == 0), j, j - 1) + 1
    jp = np.where((~np.asarray(pery)) & (j == ny), j, j + 1) + 1
    I, J = np.meshgrid(i + 1, j + 1)
    IM, _ = np.meshgrid(im, j)
    IP, _ = np.meshgrid(ip, j)
    _, JM = np.meshgrid(i, jm)
    _, JP = np.meshgrid(i, jp)
    for _ in range(passes):
        _fill_slab_images(h, nx, ny, perx, pery)
        ho = h.copy()
        h[J, I] = ho[J, I] + beta * (ho[J, IP] + ho[J, IM] + ho[JP, I] + ho[JM, I] - 4.0 * ho[J, I])
    fill_surface_ghosts(h, nx, ny, perx, pery)


def inject_ground(z_fine: np.ndarray, r: int, nxc: int, nyc: int, perx: bool, pery: bool) -> np.ndarray:
    """The coarse ground slab: node (I, J) takes fine node (r I, r J), the
    ghosts filled (Source/MeshHierarchy.cpp:128-150). z_fine: a mesh's z_nd
    (its k = 0 slab is read)."""
    h = np.zeros((nyc + 3, nxc + 3))
    h[1:nyc + 2, 1:nxc + 2] = z_fine[1, 1:r * nyc + 2:r, 1:r * nxc + 2:r]
    fill_surface_ghosts(h, nxc, nyc, perx, pery)
    return h


def hierarchy(tile: TileMesh, levels: int, passes: int = 1, smoothing: int = 0, gamma_m: float = 0.5,
              passes_given: bool = False) -> list:
    """The meshes of levels 0 .. levels-1 (level 0 is tile), each level's
    ground level 0's at its nodes smoothed by ground_passes(l, passes) passes,
    its mesh built from it with the run's smoothing and gamma_m
    (MeshHierarchy::MeshHierarchy, Source/MeshHierarchy.cpp:152-242)."""
    g0 = tile.geom
    err = check_hierarchy_inputs(g0.n, levels, passes, passes_given)
    if err:
        raise ValueError("gust: invalid grid hierarchy:\n  " + "\n  ".join(err))
    out = [tile]
    perx, pery = bool(g0.periodic[0]), bool(g0.periodic[1])
    for lev in range(1, levels):
        r = hierarchy_ratio(lev)
        nxc, nyc = g0.n[0] // r, g0.n[1] // r
        geom = geometry((nxc, nyc, g0.n[2]), tile.prob_lo, tile.prob_hi, g0.periodic[:2])
        h = inject_ground(tile.z_nd, r, nxc, nyc, perx, pery)
        smooth_surface(h, nxc, nyc, perx, pery, ground_passes(lev, passes), HIERARCHY_BETA)
        m = build_mesh(geom, tile.prob_lo, tile.prob_hi, tile.zlevels, Terrain(kind="flat"), smoothing=smoothing,
                       gamma_m=gamma_m, ground=h)
        m.ground_passes = ground_passes(lev, passes)
        out.append(m)
    return out
