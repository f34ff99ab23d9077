"""The pressure correction of Source/Simplec.cpp:134-220 composed from the
oracle's operators (TEST INFRASTRUCTURE): face_coefficients (87-112), the
divergence, the Poisson solve done exactly (a sparse direct solve of the
oracle operator with beta = 1/d; the singular case pinned at one cell and the
J-weighted means removed as Source/Poisson.cpp:559-567 and 603-611 do), the
fluxes, and the updates of p, u, v, Omega and w."""
from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from oracle import oracle as orc
from tests import poisson_helpers as PH

B = (-1, -1, -1)


def beta_keep(d):
    on = d > 0
    return np.where(on, 1.0 / np.where(on, d, 1.0), 1.0), on.astype(np.float64)


def _embed(mesh_arr, valid):
    out = np.ones_like(mesh_arr)
    out[1:-1, 1:-1, 1:-1] = valid
    return out


def periodic_images(a, per, nodal):
    """FillBoundary of a one-ghost (1, nz', ny', nx') array: every ghost point
    whose periodic image (x and/or y) is a valid point takes its value."""
    a = a.copy()
    _, ez, ey, ex = a.shape
    nvx, nvy = ex - 2, ey - 2              # valid points along x, y (faces included)
    nx, ny = nvx - nodal[0], nvy - nodal[1]   # the period
    K, Jj, I = np.meshgrid(np.arange(ez) - 1, np.arange(ey) - 1, np.arange(ex) - 1, indexing="ij")
    inside = (I >= 0) & (I < nvx) & (Jj >= 0) & (Jj < nvy) & (K >= 0) & (K < ez - 2)
    A, Cc = I.copy(), Jj.copy()
    if per[0]:
        A = np.where(A < 0, A + nx, np.where(A >= nvx, A - nx, A))
    if per[1]:
        Cc = np.where(Cc < 0, Cc + ny, np.where(Cc >= nvy, Cc - ny, Cc))
    ok = (~inside) & (A >= 0) & (A < nvx) & (Cc >= 0) & (Cc < nvy) & (K >= 0) & (K < ez - 2) & ((A != I) | (Cc != Jj))
    src = a[0, K[ok] + 1, Cc[ok] + 1, A[ok] + 1]
    a[0, K[ok] + 1, Jj[ok] + 1, I[ok] + 1] = src
    return a


def pressure_correction(mesh, u, v, w, p, d, bc, alpha_p, source=None, solve=None):
    """u, v, w, p: one-ghost (1, ...) arrays (begin -1); d = (d_u, d_v, d_w)
    valid-face arrays. Returns (u, v, w, p, phi, F, div_before, div_after).
    solve(rhs, coeffs, bc) -> phi, when given, replaces the sparse direct
    solve (for tiles too large to factor); it must return the solution of
    the oracle operator's system with the J-weighted mean removed in the
    singular case (tests/config_helpers.py: refinement on the oracle
    operator)."""
    g = PH.ogeom(mesh)
    per = mesh.geom.periodic
    betas, keeps = zip(*(beta_keep(x) for x in d))
    div0 = orc.velocity_divergence(u, v, w, B, mesh.z_nd, B, mesh.ax, mesh.ay, mesh.az, B, mesh.detJ, B, g)
    rhs = div0 - (source if source is not None else 0.0)
    coeffs = (_embed(mesh.ax, betas[0]), _embed(mesh.ay, betas[1]), _embed(mesh.az, betas[2]))
    if solve is not None:
        phi = solve(rhs, coeffs, bc)
        return _finish_correction(mesh, u, v, w, p, bc, alpha_p, phi, coeffs, keeps, div0, g, per)
    A = PH.sparse_operator(mesh, bc, coeffs=coeffs).tocsr()
    J = mesh.detJ[1:-1, 1:-1, 1:-1]
    singular = all(b != orc.DIRICHLET for b in bc)
    r = rhs.ravel().copy()
    if singular:
        r -= np.sum(J.ravel() * r) / np.sum(J)
        A = A.tolil()
        A[0, :] = 0
        A[0, 0] = 1.0
        r[0] = 0.0
        A = A.tocsr()
    phi = spla.spsolve(A.tocsc(), r).reshape(rhs.shape)
    if singular:
        phi -= np.sum(J * phi) / np.sum(J)
    return _finish_correction(mesh, u, v, w, p, bc, alpha_p, phi, coeffs, keeps, div0, g, per)


def _finish_correction(mesh, u, v, w, p, bc, alpha_p, phi, coeffs, keeps, div0, g, per):
    """The fluxes of phi and the updates of p, u, v, Omega and w (Source/Simplec.cpp:180-215)."""
    F = orc.poisson_fluxes(np.pad(phi, 1)[None].copy(), B, mesh.z_nd, B, g, bc, coeffs=coeffs, coeff_begin=B)
    p = p.copy()
    p[0, 1:-1, 1:-1, 1:-1] += alpha_p * phi
    om, ob = orc.omega_from_w(u, v, w, B, mesh.z_nd, B, g, 0)
    u, v = u.copy(), v.copy()
    u[0, 1:-1, 1:-1, 1:-1] += F[0] * keeps[0]
    v[0, 1:-1, 1:-1, 1:-1] += F[1] * keeps[1]
    om = om + F[2] * keeps[2]
    nx, ny, nz = mesh.geom.n
    if per[0]:   # OverrideSync: the high seam copy takes the low one
        u[0, 1:-1, 1:-1, nx + 1] = u[0, 1:-1, 1:-1, 1]
    if per[1]:
        v[0, 1:-1, ny + 1, 1:-1] = v[0, 1:-1, 1, 1:-1]
    u = periodic_images(u, per, (1, 0))
    v = periodic_images(v, per, (0, 1))
    wn = orc.w_from_omega(u, v, om[None].copy(), B, ob, mesh.z_nd, B, g)
    w = w.copy()
    w[0, 1:-1, 1:-1, 1:-1] = wn
    w = periodic_images(w, per, (0, 0))
    div1 = orc.velocity_divergence(u, v, w, B, mesh.z_nd, B, mesh.ax, mesh.ay, mesh.az, B, mesh.detJ, B, g)
    return u, v, w, p, phi, F, div0, div1


def random_state(mesh, seed, open_top=False):
    """A predicted state as test_Simplec.cpp:176-189 draws it: u* in [5, 10),
    v* in [-2, 2), w* in [-0.5, 0.5), p in [-5, 5), d in [20, 60) on the faces
    with d_w = 0 on the ground and top rows (their Dirichlet rows; an open
    top, the pressure's Dirichlet face, keeps its d_w and w); periodic
    images, even z ghosts. w on the Dirichlet rows is the no-flux w
    (Omega = 0; 0 on a flat mesh) so no net source enters."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = mesh.geom.n
    per = mesh.geom.periodic

    def field(loc, lo, hi):
        nd = [(loc >> q) & 1 for q in range(3)]
        a = np.zeros((1, nz + nd[2] + 2, ny + nd[1] + 2, nx + nd[0] + 2))
        a[0, 1:-1, 1:-1, 1:-1] = rng.uniform(lo, hi, size=(nz + nd[2], ny + nd[1], nx + nd[0]))
        if nd[0] and per[0]:
            a[0, 1:-1, 1:-1, nx + 1] = a[0, 1:-1, 1:-1, 1]
        if nd[1] and per[1]:
            a[0, 1:-1, ny + 1, 1:-1] = a[0, 1:-1, 1, 1:-1]
        a[0, 0] = a[0, 1]
        a[0, -1] = a[0, -2]
        return periodic_images(a, per, nd[:2])

    u = field(1, 5, 10)
    v = field(2, -2, 2)
    w = field(4, -0.5, 0.5)
    p = field(0, -5, 5)
    d = [rng.uniform(20, 60, size=(nz + (q == 2), ny + (q == 1), nx + (q == 0))) for q in range(3)]
    d[2][0] = 0.0
    if not open_top:
        d[2][-1] = 0.0
    if per[0]:   # a periodic seam face is one unknown: its two copies agree
        d[0][:, :, nx] = d[0][:, :, 0]
    if per[1]:
        d[1][:, ny, :] = d[1][:, 0, :]
    g = PH.ogeom(mesh)
    om0 = np.zeros((1, nz + 1, ny, nx))
    wnf = orc.w_from_omega(u, v, om0, B, tuple(mesh.geom.lo), mesh.z_nd, B, g)
    w[0, 1, 1:-1, 1:-1] = wnf[0]
    if not open_top:
        w[0, nz + 1, 1:-1, 1:-1] = wnf[nz]
    w[0, 0] = w[0, 1]
    w[0, -1] = w[0, -2]
    w = periodic_images(w, per, (0, 0))
    return u, v, w, p, d


# --- momentum predictor and scalar sweep (Source/Predictor.cpp:129-405, Source/Simplec.cpp:222-281) ---------

# The line-sweep kind of the composites: None for z-line Jacobi, or
# (ghosts, overlap, order, alternate) for line LU-SGS (set by lusgs_sweeps).
_LUSGS = None


class lusgs_sweeps:
    """Within this context the composites sweep with the oracle's line
    LU-SGS (Source/LineSweeps.cpp:415-428) instead of z-line Jacobi."""

    def __init__(self, ghosts=0, overlap=0, order=0, alternate=0):
        self.opt = (ghosts, overlap, order, alternate)

    def __enter__(self):
        global _LUSGS
        self.prev, _LUSGS = _LUSGS, self.opt
        return self

    def __exit__(self, *exc):
        global _LUSGS
        _LUSGS = self.prev


def _zline(loc, A, ap, rhs, sweeps, geom):
    """The oracle's z-line Jacobi (or LU-SGS, see lusgs_sweeps) on valid
    arrays: returns (delta valid, change)."""
    sh = rhs.shape
    d = np.zeros((1, sh[0] + 2, sh[1] + 2, sh[2] + 2))
    lo = tuple(geom.lo)
    db = (lo[0] - 1, lo[1] - 1, lo[2] - 1)
    if _LUSGS is not None:
        ch = orc.zline_lusgs(loc, A, lo, ap[None].copy(), rhs[None].copy(), lo, sweeps, geom, d, db, *_LUSGS)
    else:
        ch = orc.zline_jacobi(loc, A, lo, ap[None].copy(), rhs[None].copy(), lo, sweeps, geom, d, db)
    return d[0, 1:-1, 1:-1, 1:-1].copy(), ch


def predictor_fields(A, dt, mask=None):
    """a_P = keep/dtau + A_P and d = keep (a_P - sum a_nb) (Predictor.cpp:186-201)."""
    keep = 1.0 - mask if mask is not None else 1.0
    ap = keep / dt + A[0]
    nb = A[1] + A[2] + A[3] + A[4] + A[5] + A[6]
    return ap, keep * (ap - nb)


def momentum_predictor(mesh, vel, R, A, dtau, sweeps, mask=None, w_vnn_cap=0.0, margin=0.0, coriolis_f=0.0,
                       relax=0.3, gravity=0.0, theta0=300.0, theta=None, R_theta=None, A_theta=None):
    """vel: one-ghost (1, ...) arrays; R, A, dtau, mask: valid arrays per component.
    Returns (new vel, ap, d, max |delta| per component, last sweep change per component)."""
    g = PH.ogeom(mesh)
    nx, ny, nz = mesh.geom.n
    per = mesh.geom.periodic
    dxi = (1.0 / mesh.geom.dx[0], 1.0 / mesh.geom.dx[1])
    vel = [v.copy() for v in vel]
    aps, ds, dmax, last = [], [], [], []
    for c in range(3):
        loc = 1 << c
        m = mask[c] if mask is not None else None
        keep = 1.0 - m if m is not None else np.ones_like(dtau[c])
        ap, d = predictor_fields(A[c], dtau[c], m)
        if c == 2 and w_vnn_cap > 0:
            term = (A[c][5] + A[c][6]) / w_vnn_cap
            if margin > 0:
                Z = mesh.z_nd   # [k+1, j+1, i+1]
                z = lambda di, dj: Z[1:nz + 2, 1 + dj:ny + 1 + dj, 1 + di:nx + 1 + di]
                hx = 0.5 * (z(1, 1) + z(1, 0) - z(0, 1) - z(0, 0)) * dxi[0]
                hy = 0.5 * (z(1, 1) + z(0, 1) - z(1, 0) - z(0, 0)) * dxi[1]
                s2 = hx * hx + hy * hy
                root = np.sqrt(1.0 + s2) + 1.0
                with np.errstate(divide="ignore", invalid="ignore"):
                    lim = np.maximum(0.0, 1.0 / (margin * s2 / (root * root)) - 1.0) / dtau[c]
                term = np.where(s2 > 0, np.minimum(term, lim), term)
            ap = ap + keep * term
            d = d + keep * term
        rhs = -R[c]
        if c < 2 and coriolis_f != 0:
            co = 1 - c
            other = R[co]
            # per-level mean over every face once
            sl = [slice(None)] * 3
            if per[co]:
                sl[2 - co] = slice(0, other.shape[2 - co] - 1)
            mean = other[tuple(sl)].sum(axis=(1, 2)) / (
                ((nx if co == 0 else ny) + (0 if per[co] else 1)) * (nx if c == 0 else ny))
            og = np.zeros((other.shape[0] + 2, other.shape[1] + 2, other.shape[2] + 2))
            og[1:-1, 1:-1, 1:-1] = other
            ogp = periodic_images(og[None].copy(), per, (co == 0, co == 1))[0][1:-1]   # drop the z ghosts
            nb = A[c][1] + A[c][2] + A[c][3] + A[c][4] + A[c][5] + A[c][6]
            net = np.maximum(A[c][0] - nb, 0.0)
            tau_c = 1.0 / np.maximum(abs(coriolis_f) / relax, net)
            tau_own = 1.0 / (1.0 / dtau[c] + net)
            stiff = coriolis_f ** 2 * np.maximum(tau_c, tau_own)
            K, Jj, I = np.meshgrid(*(np.arange(s) for s in ap.shape), indexing="ij")
            if c == 0:   # R.v at the x-face (i, j): y-faces (i-1, j), (i, j), (i-1, j+1), (i, j+1)
                avg = 0.25 * (ogp[K, Jj + 1, I] + ogp[K, Jj + 1, I + 1] + ogp[K, Jj + 2, I] + ogp[K, Jj + 2, I + 1])
                n, nn = I, nx
            else:        # R.u at the y-face (i, j): x-faces (i, j-1), (i+1, j-1), (i, j), (i+1, j)
                avg = 0.25 * (ogp[K, Jj, I + 1] + ogp[K, Jj, I + 2] + ogp[K, Jj + 1, I + 1] + ogp[K, Jj + 1, I + 2])
                n, nn = Jj, ny
            on = keep != 0
            if not per[c]:
                on = on & (n != 0) & (n != nn)
            mk = mean[K]
            ap = np.where(on, ap + stiff, ap)
            d = np.where(on, d + stiff, d)
            sg = -coriolis_f if c == 0 else coriolis_f
            rhs = np.where(on, rhs + sg * (tau_c * mk + tau_own * (avg - mk)), rhs)
        if c == 2 and gravity > 0:
            gt = gravity / theta0
            Z = mesh.z_nd
            zc = 0.125 * (Z[0:nz + 2, 1:ny + 1, 1:nx + 1] + Z[0:nz + 2, 1:ny + 1, 2:nx + 2] + Z[0:nz + 2, 2:ny + 2, 1:nx + 1]
                          + Z[0:nz + 2, 2:ny + 2, 2:nx + 2] + Z[1:nz + 3, 1:ny + 1, 1:nx + 1] + Z[1:nz + 3, 1:ny + 1, 2:nx + 2]
                          + Z[1:nz + 3, 2:ny + 2, 1:nx + 1] + Z[1:nz + 3, 2:ny + 2, 2:nx + 2])   # cells k = -1 .. nz
            th = theta[0, :, 1:-1, 1:-1]   # k = -1 .. nz
            kk = np.arange(1, nz)          # interior faces
            gamma = (th[kk + 1] - th[kk]) / (zc[kk + 1] - zc[kk])
            ap_t = 0.5 * (A_theta[0][kk - 1] + A_theta[0][kk])
            tau = 1.0 / (1.0 / dtau[c][kk] + ap_t)
            stiff = gt * np.maximum(gamma, 0.0) * tau
            on = keep[kk] != 0
            ap[kk] = np.where(on, ap[kk] + stiff, ap[kk])
            d[kk] = np.where(on, d[kk] + stiff, d[kk])
            rhs[kk] = np.where(on, rhs[kk] - gt * tau * 0.5 * (R_theta[kk - 1] + R_theta[kk]), rhs[kk])
        delta, ch = _zline(loc, A[c], ap, rhs, sweeps, g)
        vel[c][0, 1:-1, 1:-1, 1:-1] += delta
        aps.append(ap)
        ds.append(d)
        dmax.append(np.max(np.abs(delta)))
        last.append(ch)
    return vel, aps, ds, dmax, last


def scalar_sweep(mesh, field, R, A, dtau, sweeps, floor=0.0, clip=False):
    """field: one-ghost (1, ...) array; returns (field, max |dq|, last change, clipped)."""
    g = PH.ogeom(mesh)
    ap = 1.0 / dtau + A[0]
    delta, ch = _zline(0, A, ap, -R, sweeps, g)
    f = field.copy()
    f[0, 1:-1, 1:-1, 1:-1] += delta
    clipped = 0
    if clip:
        v = f[0, 1:-1, 1:-1, 1:-1]
        low = v < floor
        clipped = int(low.sum())
        v[low] = floor
    return f, np.max(np.abs(delta)), ch, clipped
