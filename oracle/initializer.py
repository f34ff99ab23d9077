"""The mass-consistent initializer's first guess (Source/Initializer.cpp:
25-53, 197-335) and the inflow profile it evaluates (Source/InflowProfile.H:
139-200, Source/InflowProfile.cpp:178-205), restated in numpy over the
oracle's C mass-flux and divergence pieces.

TEST INFRASTRUCTURE ONLY (see ``oracle.py``): used by tests/ and bench.py's
CPU leg as the checker of the HIP library.

Arrays are float64 (1, nz', ny', nx'): the state with two ghosts (begin -2),
the mesh's z_nd, detJ, ax, ay, az with one (begin -1). Every point value is
formed with the reference's operations in its order; numpy's log, exp and
pow may differ from the device's in the last bit.

Pinned by: the reference's Askervein initializer (Tests/Initialization/
Mass_Consistent_Askervein: its divergence before the projection to 1e-9 and
its state plotfile to 1e-8, through the composite of tests/init_helpers.py),
closed forms of the profiles (the log law's k = u*^2/Cmu0^2 at every height
when unbounded, the power law's slope), the tabulated kind against the
analytic one it tabulates, and the O'Brien pass's invariants (zero top
residual, the column continuity of its Omega).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import oracle as orc
from . import physics as P

B1 = (-1, -1, -1)
B2 = (-2, -2, -2)
LOG_LAW, POWER_LAW, TABLE = 0, 1, 2
CC, XF, YF, ZF = 0, 1, 2, 4


@dataclass
class InflowProfile:
    """Source/InflowProfile.H:42-66 (InflowProfileOptions; the defaults are
    the reference's) with the table in memory (z u v theta [k]) and the
    sounding theta_ref(z)."""
    kind: int = LOG_LAW
    ustar: float = 0.4
    z0: float = 0.1
    speed_ref: float = 10.0
    z_ref: float = 10.0
    alpha: float = 0.14
    direction: float = 0.0
    Cmu0: float = 0.5562
    kappa: float = 0.41
    unbounded: bool = False
    l_max: float = 30.0
    table: tuple | None = None          # (z, u, v, theta, k or None)
    theta_ref_z: tuple = (0.0,)
    theta_ref: tuple = (300.0,)


@dataclass
class InitOptions:
    """Source/Initializer.H:63-98 (InitOptions), the profile first guess."""
    first_guess: int = 1                # 0 none, 1 profile
    projection: bool = True
    z_agl_min: float = 0.0
    alpha_h: float = 1.0
    alpha_v: float = 1.0
    k_decay_height: float = 0.0
    obrien: bool = False
    column_theta: bool = False


def c_profile(p: InflowProfile):
    """The C oracle's view of the profile (Source/InflowProfile.cpp:178-205):
    (OProfile, the arrays it points into)."""
    import ctypes as C
    o = orc.OProfile()
    o.kind = int(p.kind)
    o.ustar, o.z0, o.speed_ref, o.z_ref, o.alpha = p.ustar, p.z0, p.speed_ref, p.z_ref, p.alpha
    a = p.direction * 3.14159265358979323846 / 180
    o.cosd, o.sind = math.cos(a), math.sin(a)
    o.Cmu0, o.kappa = p.Cmu0, p.kappa
    o.inv_lmax = 0.0 if p.unbounded else 1.0 / p.l_max
    keep = []

    def arr(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        keep.append(x)
        return x.ctypes.data_as(C.POINTER(C.c_double))

    if p.kind == TABLE:
        tz, tu, tv, tt, tk = p.table
        o.tz, o.tu, o.tv, o.tt = arr(tz), arr(tu), arr(tv), arr(tt)
        o.n = len(tz)
        o.has_k = 1 if tk is not None else 0
        o.tk = arr(tk if tk is not None else np.zeros(len(tz)))
    o.rz, o.rt = arr(p.theta_ref_z), arr(p.theta_ref)
    o.rn = len(p.theta_ref_z)
    return o, keep


def profile_eval_c(p: InflowProfile, z):
    """The C oracle's evaluation at the heights z: (u, v, theta, k, nu_t)."""
    import ctypes as C
    o, keep = c_profile(p)
    z = np.asarray(z, dtype=np.float64)
    out = np.zeros((5,) + z.shape)
    buf = (C.c_double * 5)()
    for idx in np.ndindex(z.shape):
        orc.lib().oracle_profile_eval(C.byref(o), float(z[idx]), buf)
        for n in range(5):
            out[(n,) + idx] = buf[n]
    del keep
    return tuple(out)


def table_interp(x, y, xq):
    """Source/InflowProfile.H:137-147: linear in the table, constant beyond
    its ends."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    xq = np.asarray(xq, dtype=np.float64)
    n = len(x)
    if n == 1:
        return np.full(xq.shape, y[0])
    m = np.clip(np.searchsorted(x, xq, side="right"), 1, n - 1)
    w = (xq - x[m - 1]) / (x[m] - x[m - 1])
    out = y[m - 1] + w * (y[m] - y[m - 1])
    out = np.where(xq >= x[n - 1], y[n - 1], out)
    return np.where(xq <= x[0], y[0], out)


def _search(x, xq):
    n = len(x)
    return np.clip(np.searchsorted(np.asarray(x), xq, side="right"), 1, n - 1)


def inflow_eval(p: InflowProfile, z):
    """Source/InflowProfile.H:149-200 (View::eval) and the view's constants
    (Source/InflowProfile.cpp:178-205): (u, v, theta, k, nu_t) at heights z
    above the ground."""
    z = np.asarray(z, dtype=np.float64)
    z = np.where(z > 0.0, z, 0.0)
    inv_lmax = 0.0 if p.unbounded else 1.0 / p.l_max
    c = 1.0 / (1.0 + p.kappa * (z + p.z0) * inv_lmax)
    if p.kind == TABLE:
        tz, tu, tv, tt, tk = p.table
        tz, tu, tv, tt = (np.asarray(a, dtype=np.float64) for a in (tz, tu, tv, tt))
        n = len(tz)
        u = table_interp(tz, tu, z)
        v = table_interp(tz, tv, z)
        theta = table_interp(tz, tt, z)
        m = _search(tz, z)
        s0 = np.sqrt(tu[m - 1] * tu[m - 1] + tv[m - 1] * tv[m - 1])
        s1 = np.sqrt(tu[m] * tu[m] + tv[m] * tv[m])
        dsdz = (s1 - s0) / (tz[m] - tz[m - 1]) if n > 1 else np.zeros_like(z)
        ustar_l = p.kappa * (z + p.z0) * np.where(dsdz > 0.0, dsdz, 0.0) * c
        if tk is not None:
            k = table_interp(tz, np.asarray(tk, dtype=np.float64), z)
            ustar_l = p.Cmu0 * np.sqrt(np.where(k > 0.0, k, 0.0))
        else:
            k = ustar_l * ustar_l / (p.Cmu0 * p.Cmu0)
        nut = p.kappa * ustar_l * (z + p.z0) * c
        return u, v, theta, k, nut
    if p.kind == LOG_LAW:
        S = p.ustar / p.kappa * np.log((z + p.z0) / p.z0)
        ustar_l = p.ustar * c
    else:
        S = p.speed_ref * np.power(z / p.z_ref, p.alpha)
        dsdz = p.alpha * S / (0.5 * (z + p.z0 + np.abs(z - p.z0)))
        ustar_l = p.kappa * (z + p.z0) * dsdz * c
    a = p.direction * 3.14159265358979323846 / 180
    u = S * math.cos(a)
    v = S * math.sin(a)
    theta = table_interp(p.theta_ref_z, p.theta_ref, z)
    k = ustar_l * ustar_l / (p.Cmu0 * p.Cmu0)
    nut = p.kappa * ustar_l * (z + p.z0) * c
    return u, v, theta, k, nut


def node_z(z, zb, I, J, K):
    """Source/Heights_K.H:20-33 (NodeZ): node heights with the lateral
    indices clamped to the stored nodes, the levels extrapolated linearly."""
    a = z[0]
    lo = zb
    hi = (zb[0] + a.shape[2] - 1, zb[1] + a.shape[1] - 1, zb[2] + a.shape[0] - 1)
    i = np.clip(I, lo[0], hi[0]) - lo[0]
    j = np.clip(J, lo[1], hi[1]) - lo[1]
    kc = np.clip(K, lo[2], hi[2]) - lo[2]
    val = a[kc, j, i]
    below = K < lo[2]
    above = K > hi[2]
    z0, z1 = a[0, j, i], a[1, j, i]
    zt, zt1 = a[-1, j, i], a[-2, j, i]
    val = np.where(below, z0 + (lo[2] - K) * (z0 - z1), val)
    return np.where(above, zt + (K - hi[2]) * (zt - zt1), val)


def height_above_ground(loc, z, zb, I, J, K, klo):
    """Source/Heights_K.H:35-53: the point's node-mean height minus the
    node-mean ground under its footprint."""
    nx = 1 if loc & 1 else 2
    ny = 1 if loc & 2 else 2
    nz = 1 if loc & 4 else 2
    zp = np.zeros(np.shape(I))
    zg = np.zeros(np.shape(I))
    for a in range(nx):
        for b in range(ny):
            zg = zg + node_z(z, zb, I + a, J + b, np.full(np.shape(K), klo))
            for c in range(nz):
                zp = zp + node_z(z, zb, I + a, J + b, K + c)
    return zp / (nx * ny * nz) - zg / (nx * ny)


def _grid(a, begin):
    _, nk, nj, ni = a.shape
    K, J, I = np.meshgrid(np.arange(nk) + begin[2], np.arange(nj) + begin[1], np.arange(ni) + begin[0],
                          indexing="ij")
    return I, J, K


def fill_from_profile(a, begin, loc, comp, prof: InflowProfile, h_min, zi, z, zb, klo):
    """Source/Initializer.cpp:25-53 in place at every point of a: the
    profile's component comp (0 u, 1 v, 2 theta, 3 k) at the point's height
    above the ground floored at h_min; k decays as exp(-(h - z_i)/z_i) above
    z_i > 0."""
    I, J, K = _grid(a, begin)
    h = height_above_ground(loc, z, zb, I, J, K, klo)
    h = np.where(h < h_min, h_min, h)
    vals = inflow_eval(prof, h)
    val = vals[comp]
    if comp == 3 and zi > 0:
        zinv = 1.0 / max(zi, np.finfo(np.float64).tiny)
        d = h - zi
        val = val * np.exp(-(np.where(d > 0.0, d, 0.0) * zinv))
    a[0] = val


def obrien_omega(u, v, vel_begin, mesh: P.Mesh):
    """Source/Initializer.cpp:197-267: Omega on the valid z-faces from the
    column continuity of (a_x u, a_y v) from the ground up, the top residual
    removed with (height above the ground / column depth)^2; returns
    (omega (nz+1, ny, nx), max |Omega_top| before the removal)."""
    g = mesh.geom
    lo, hi = tuple(g.lo), tuple(g.hi)
    nx, ny, nz = (hi[d] - lo[d] + 1 for d in range(3))
    o = (lo[0] - vel_begin[0], lo[1] - vel_begin[1], lo[2] - vel_begin[2])
    uv = np.ascontiguousarray(u[:, o[2]:o[2] + nz, o[1]:o[1] + ny, o[0]:o[0] + nx + 1])
    vv = np.ascontiguousarray(v[:, o[2]:o[2] + nz, o[1]:o[1] + ny + 1, o[0]:o[0] + nx])
    zero = np.zeros((1, nz + 1, ny, nx))
    dh = orc.face_divergence([uv, vv, zero], lo, (mesh.ax, mesh.ay, mesh.az), B1, mesh.J, B1, g)
    J = mesh.J[0, 1:-1, 1:-1, 1:-1]
    az = mesh.az[0, 1:-1, 1:-1, 1:-1]
    zn = mesh.z[0, 1:-1, 1:-1, 1:-1]       # nodes lo .. hi+1
    dzeta = g.dx[2]
    om = np.zeros((nz + 1, ny, nx))
    for k in range(nz):
        om[k + 1] = (az[k] * om[k] - J[k] * dh[k] * dzeta) / az[k + 1]
    E = om[nz].copy()
    worst = float(np.max(np.abs(E))) if E.size else 0.0

    def zf(k):
        return 0.25 * (zn[k, :-1, :-1] + zn[k, :-1, 1:] + zn[k, 1:, :-1] + zn[k, 1:, 1:])

    zg = zf(0)
    depth = zf(nz) - zg
    for k in range(1, nz + 1):
        frac = (zf(k) - zg) / depth
        om[k] = om[k] - frac * frac * E
    om[nz] = 0.0
    return om, worst


def first_guess_state(q, mesh: P.Mesh, prof: InflowProfile, opt: InitOptions):
    """Source/Initializer.cpp:269-335 on the two-ghost state q (in place):
    u, v from the profile at every point, w = 0 but for the ground faces'
    kinematic value, theta the sounding (absolute height, one ghost cell) or
    the table's (column_theta), k the profile's with the decay, p' = 0, then
    the optional O'Brien pass. Returns the O'Brien top residual."""
    if opt.first_guess == 0:
        return 0.0
    g = mesh.geom
    klo = g.lo[2]
    z = mesh.z
    fill_from_profile(q["u"], B2, XF, 0, prof, opt.z_agl_min, 0.0, z, B1, klo)
    fill_from_profile(q["v"], B2, YF, 1, prof, opt.z_agl_min, 0.0, z, B1, klo)
    q["w"][:] = 0.0
    om, ob = orc.omega_from_w(q["u"], q["v"], q["w"], B2, z, B1, g, 0)
    om[0] = 0.0
    _put_w(q, orc.w_from_omega(q["u"], q["v"], om[None], B2, ob, z, B1, g))
    fill_from_profile(q["tke"], B2, CC, 3, prof, opt.z_agl_min, opt.k_decay_height, z, B1, klo)
    if opt.column_theta and prof.kind == TABLE:
        fill_from_profile(q["theta"], B2, CC, 2, prof, opt.z_agl_min, 0.0, z, B1, klo)
    else:
        ph = P.Physics(theta_ref_z=tuple(prof.theta_ref_z), theta_ref=tuple(prof.theta_ref))
        tr = P.theta_ref_field(mesh, ph)
        q["theta"][:] = 0.0
        q["theta"][:, 1:-1, 1:-1, 1:-1] = tr
    q["p"][:] = 0.0
    residual = 0.0
    if opt.obrien:
        om, residual = obrien_omega(q["u"], q["v"], B2, mesh)
        _put_w(q, orc.w_from_omega(q["u"], q["v"], om[None], B2, tuple(g.lo), z, B1, g))
    return residual


def _put_w(q, w_valid):
    q["w"][0, 2:-2, 2:-2, 2:-2] = w_valid


def poisson_bc(faces, periodic):
    """Source/Initializer.cpp:181-195: Dirichlet (lambda = 0) on inflow and
    outflow faces, Neumann elsewhere, periodic where the geometry is
    (faces and the result in the order xlo, ylo, zlo, xhi, yhi, zhi)."""
    out = []
    for o in range(6):
        d = o % 3
        if d < 2 and periodic[d]:
            out.append(orc.PERIODIC)
        elif faces[o] in (P.INFLOW, P.OUTFLOW):
            out.append(orc.DIRICHLET)
        else:
            out.append(orc.NEUMANN)
    return tuple(out)


def periodic_fill(a, begin, loc, geom):
    """amrex::MultiFab::FillBoundary(periodicity) on one tile: every ghost
    point whose periodic image is a valid point takes its value."""
    lo = [geom.lo[d] for d in range(3)]
    n = [geom.hi[d] - geom.lo[d] + 1 for d in range(3)]
    vhi = [geom.hi[d] + ((loc >> d) & 1) for d in range(3)]
    I, J, K = _grid(a, begin)
    src = [I.copy(), J.copy(), K.copy()]
    for d, X in ((0, I), (1, J)):
        if not geom.periodic[d]:
            continue
        s = src[d]
        s[X < lo[d]] += n[d]
        s[X > vhi[d]] -= n[d]
    inside = np.ones(I.shape, dtype=bool)
    for d in range(3):
        inside &= (src[d] >= lo[d]) & (src[d] <= vhi[d])
    ghost = np.zeros(I.shape, dtype=bool)
    for d, X in enumerate((I, J, K)):
        ghost |= (X < lo[d]) | (X > vhi[d])
    sel = inside & ghost
    a[0][sel] = a[0][src[2][sel] - begin[2], src[1][sel] - begin[1], src[0][sel] - begin[0]]
