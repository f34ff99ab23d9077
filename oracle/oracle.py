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


class OClosure(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("Cmu0", "Cb", "Rt_crit", "Rt_min", "max_geom_lscale", "tke_floor",
                                          "sigma_k", "nu_molecular", "alpha_molecular", "kappa")]


class OSurface(C.Structure):
    _fields_ = [("z0", C.c_double), ("Cmu0", C.c_double), ("kappa", C.c_double), ("wall_k_dirichlet", C.c_int),
                ("floor_one", C.c_int)]


class OForcing(C.Structure):
    _fields_ = [("gravity", C.c_double), ("theta0", C.c_double), ("coriolis", C.c_int), ("three_d", C.c_int),
                ("geostrophic", C.c_int), ("latitude", C.c_double), ("rotation_period", C.c_double),
                ("ug", C.c_double), ("vg", C.c_double), ("pgrad", C.c_double * 3)]


class OProfile(C.Structure):
    _fields_ = [("kind", C.c_int), ("ustar", C.c_double), ("z0", C.c_double), ("speed_ref", C.c_double),
                ("z_ref", C.c_double), ("alpha", C.c_double), ("cosd", C.c_double), ("sind", C.c_double),
                ("Cmu0", C.c_double), ("kappa", C.c_double), ("inv_lmax", C.c_double),
                ("tz", C.POINTER(C.c_double)), ("tu", C.POINTER(C.c_double)), ("tv", C.POINTER(C.c_double)),
                ("tt", C.POINTER(C.c_double)), ("tk", C.POINTER(C.c_double)), ("n", C.c_int), ("has_k", C.c_int),
                ("rz", C.POINTER(C.c_double)), ("rt", C.POINTER(C.c_double)), ("rn", C.c_int)]


class ODamping(C.Structure):
    _fields_ = [("rayleigh_zdamp", C.c_double), ("rayleigh_dampcoef", C.c_double), ("rayleigh_fields", C.c_int * 5),
                ("rayleigh_above_ground", C.c_int), ("sponge_width", C.c_double), ("sponge_strength", C.c_double),
                ("ztop", C.c_double), ("prob_lo", C.c_double * 2), ("prob_hi", C.c_double * 2)]


class OBc(C.Structure):
    _fields_ = [("face", C.c_int * 6), ("zhi_theta_grad", C.c_double), ("wall_velocity", (C.c_double * 3) * 6),
                ("ug", C.c_double), ("vg", C.c_double), ("profile", C.POINTER(OProfile)), ("damping", ODamping)]


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
        L.oracle_zline_jacobi_threads.argtypes = [C.c_int, P(OA4), P(OA4), P(OA4), C.c_int, P(OGeom), P(OA4),
                                                  P(C.c_double), C.c_int]
        L.oracle_zline_lusgs.argtypes = [C.c_int, P(OA4), P(OA4), P(OA4), C.c_int, P(OGeom), P(OA4), C.c_int,
                                         C.c_int, C.c_int, C.c_int, P(C.c_double)]
        L.oracle_poisson_error.restype = C.c_char_p
        I6 = C.c_int * 6
        L.oracle_poisson_zero_flux_mask.argtypes = [I6]
        L.oracle_poisson_fill_ghosts.argtypes = [P(OA4), P(OGeom), I6]
        L.oracle_poisson_apply.argtypes = [P(OA4), P(OA4), P(OA4), P(OA4), P(OA4), P(OA4), P(OA4), P(OA4), P(OA4),
                                           P(OGeom), I6]
        L.oracle_poisson_fluxes.argtypes = [P(OA4), P(OA4), P(OA4), P(OA4), P(OA4), P(OA4), P(OA4), P(OA4), P(OA4),
                                            P(OGeom), I6]
        L.oracle_face_divergence.argtypes = [P(OA4), P(OA4), P(OA4), P(OA4), P(OA4), P(OA4), P(OA4), P(OA4),
                                             P(OGeom)]
        L.oracle_face_divergence.restype = None
        V = P(OA4)
        L.oracle_omega_from_w.argtypes = [V, V, V, V, V, P(OGeom), C.c_int]
        L.oracle_w_from_omega.argtypes = [V, V, V, V, V, P(OGeom)]
        L.oracle_transport_fluxes.argtypes = [V, V, V, V, V, V, V, V, V, P(OGeom)]
        L.oracle_velocity_divergence.argtypes = [V, V, V, V, V, V, V, V, V, P(OGeom), V]
        L.oracle_pressure_gradient.argtypes = [V, V, V, V, V, V, V, V, P(OGeom)]
        L.oracle_local_dtau.argtypes = [V, V, V, V, V, V, V, C.c_double, P(OGeom), V]
        L.oracle_ser_cfl.argtypes = [C.c_double] * 6
        L.oracle_ser_cfl.restype = C.c_double
        L.oracle_face_dtau.argtypes = [V, V, V, V, P(OGeom)]
        L.oracle_implicit_coefficients.argtypes = [C.c_int, V, V, V, V, V, V, P(OGeom), C.c_int, V]
        L.oracle_apply_implicit.argtypes = [C.c_int, V, V, P(OGeom), V]
        L.oracle_convective_fluxes.argtypes = [C.c_int, V, V, V, V, C.c_int, C.c_double, P(OGeom), V, V, V]
        L.oracle_flux_divergence.argtypes = [C.c_int, V, V, V, V, P(OGeom), V]
        L.oracle_subtract_transport_divergence.argtypes = [C.c_int, V, V, V, V, V, P(OGeom), V]
        L.oracle_scalar_diffusive_fluxes.argtypes = [V, V, V, P(OGeom), C.c_int, V, V, V]
        L.oracle_momentum_diffusive_fluxes.argtypes = [C.c_int, V, V, V, V, V, P(OGeom), C.c_int, V, V, V]
        I3 = C.c_int * 3
        D = C.c_double
        L.oracle_box_sum.argtypes = [V, I3, I3]
        L.oracle_box_sum.restype = D
        L.oracle_global_max.argtypes = [V, I3, I3]
        L.oracle_global_max.restype = D
        L.oracle_scaled_sums.argtypes = [V, V, V, D, C.c_int, P(OGeom), P(D)]
        L.oracle_scaled_sums.restype = None
        L.oracle_residual_scales.argtypes = [D, D, D, D, P(D)]
        L.oracle_residual_scales.restype = None
        L.oracle_stage_norms.argtypes = [P(OA4) * 6, V, V, V, V, V, D, D, D, D, P(OGeom), P(D), P(D)]
        L.oracle_confidence.argtypes = [P(OA4) * 6, V, D, D, D, D, P(D), P(OGeom), P(D), P(D)]
        L.oracle_confidence.restype = None
        L.oracle_height_above_ground.argtypes = [V, P(OGeom), V]
        L.oracle_height_above_ground.restype = None
        L.oracle_closure_evaluate.argtypes = [V, V, V, V, V, V, V, D, P(OClosure), D, D, P(OGeom), V, V, V, V, V, V,
                                              V]
        L.oracle_surface_state.argtypes = [V, V, V, P(OSurface), P(OGeom), V]
        L.oracle_fill_ghosts.argtypes = [P(OA4) * 6, V, P(OBc), P(OSurface), P(OGeom), V]
        L.oracle_fill_ghosts_homogeneous.argtypes = [P(OA4) * 6, V, P(OBc), P(OGeom)]
        L.oracle_coriolis_parameter.argtypes = [C.c_int, D, D]
        L.oracle_coriolis_parameter.restype = D
        L.oracle_fill_theta_ref.argtypes = [C.c_int, P(D), P(D), V, V]
        L.oracle_fill_theta_ref.restype = None
        L.oracle_forcing.argtypes = [V, V, V, V, V, P(OForcing), P(OGeom), V, V, V]
        L.oracle_forcing.restype = None
        L.oracle_set_ground_fluxes.argtypes = [V, P(OGeom), V, V, V, V]
        L.oracle_set_ground_fluxes.restype = None
        L.oracle_apply_wall_k.argtypes = [V, V, C.c_int, P(OGeom), V]
        L.oracle_apply_wall_k.restype = None
        L.oracle_apply_rows.argtypes = [V, V, V, V, P(OBc), P(OGeom), V, V, V]
        L.oracle_apply_rows.restype = None
        L.oracle_dirichlet_rows.argtypes = [C.c_int, P(OBc), P(OGeom), V]
        L.oracle_dirichlet_rows.restype = None
        L.oracle_profile_eval.argtypes = [P(OProfile), D, P(D)]
        L.oracle_profile_eval.restype = None
        L.oracle_implicit_folds.argtypes = [C.c_int, V, V, V, V, P(OBc), P(OSurface), P(OGeom), V]
        L.oracle_add_damping.argtypes = [P(OA4) * 6, V, P(OBc), P(OGeom), P(OA4) * 6]
        L.oracle_add_damping.restype = None
        L.oracle_implicit_folds.restype = None
        for f in ("oracle_omega_from_w", "oracle_w_from_omega", "oracle_transport_fluxes",
                  "oracle_velocity_divergence", "oracle_pressure_gradient"):
            getattr(L, f).restype = None
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


def zline_jacobi(loc, A, A_begin, ap, rhs, v_begin, sweeps, geom, delta, d_begin, threads=1):
    """Source/Predictor.cpp:87-127. Arrays are modified in place (delta).
    threads > 1: the columns split over that many OpenMP threads (the same
    bits)."""
    ch = C.c_double(0)
    if threads > 1:
        _check(lib().oracle_zline_jacobi_threads(loc, a4(A, A_begin), a4(ap, v_begin), a4(rhs, v_begin), sweeps,
                                                 geom, a4(delta, d_begin), C.byref(ch), int(threads)))
    else:
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


# ---------------------------------------------------------------------------
# Terrain Poisson operator. Fields are (nz', ny', nx') numpy arrays with a
# begin origin; helpers below take (array, begin) pairs.

PERIODIC, NEUMANN, DIRICHLET = 0, 1, 2


def _bc(bc):
    return (C.c_int * 6)(*[int(b) for b in bc])


def _v(a, begin):
    if a is None:
        return None
    a = a if a.ndim == 4 else a[None]
    return C.byref(a4(a, begin))


def _pcheck(rc):
    if rc != 0:
        raise OracleError(lib().oracle_poisson_error().decode())


def poisson_fill_ghosts(phi, phi_begin, geom, bc):
    """Source/Poisson.cpp:84-125 in place (phi with one ghost cell)."""
    _pcheck(lib().oracle_poisson_fill_ghosts(_v(phi, phi_begin), geom, _bc(bc)))


def poisson_apply(phi, phi_begin, areas, areas_begin, J, J_begin, zp, zp_begin, geom, bc, dcross=None,
                  az_mesh=None):
    """A phi on the valid cells (returns (nz, ny, nx)); phi's ghosts are filled.
    areas = (ax, ay, az) premultiplied, each with the same begin."""
    n = (geom.hi[2] - geom.lo[2] + 1, geom.hi[1] - geom.lo[1] + 1, geom.hi[0] - geom.lo[0] + 1)
    y = np.zeros((1,) + n)
    ax, ay, az = areas
    _pcheck(lib().oracle_poisson_apply(_v(y, tuple(geom.lo)), _v(phi, phi_begin), _v(ax, areas_begin),
                                       _v(ay, areas_begin), _v(az, areas_begin), _v(J, J_begin), _v(zp, zp_begin),
                                       _v(dcross, areas_begin), _v(az_mesh, areas_begin), geom, _bc(bc)))
    return y[0]


def poisson_fluxes(phi, phi_begin, zp, zp_begin, geom, bc, coeffs=(None, None, None), coeff_begin=(0, 0, 0),
                   dcross=None):
    """Fluxes on the valid faces (x: (nz, ny, nx+1), ...), begin = geom.lo."""
    nx, ny, nz = (geom.hi[0] - geom.lo[0] + 1, geom.hi[1] - geom.lo[1] + 1, geom.hi[2] - geom.lo[2] + 1)
    fx, fy, fz = np.zeros((1, nz, ny, nx + 1)), np.zeros((1, nz, ny + 1, nx)), np.zeros((1, nz + 1, ny, nx))
    lo = tuple(geom.lo)
    bx, by, bz = coeffs
    _pcheck(lib().oracle_poisson_fluxes(_v(fx, lo), _v(fy, lo), _v(fz, lo), _v(phi, phi_begin), _v(zp, zp_begin),
                                        _v(bx, coeff_begin), _v(by, coeff_begin), _v(bz, coeff_begin),
                                        _v(dcross, coeff_begin), geom, _bc(bc)))
    return fx[0], fy[0], fz[0]


def face_divergence(F, F_begin, areas, areas_begin, J, J_begin, geom):
    n = (geom.hi[2] - geom.lo[2] + 1, geom.hi[1] - geom.lo[1] + 1, geom.hi[0] - geom.lo[0] + 1)
    div = np.zeros((1,) + n)
    lib().oracle_face_divergence(_v(div, tuple(geom.lo)), _v(F[0], F_begin), _v(F[1], F_begin), _v(F[2], F_begin),
                                 _v(areas[0], areas_begin), _v(areas[1], areas_begin), _v(areas[2], areas_begin),
                                 _v(J, J_begin), geom)
    return div[0]


# --- mass flux (oracle/massflux.c) -------------------------------------------------------------------------

def _shape_at(geom, loc, grow=(0, 0, 0)):
    n = [geom.hi[d] - geom.lo[d] + 1 + ((loc >> d) & 1) + 2 * grow[d] for d in range(3)]
    return (1, n[2], n[1], n[0])


def _begin_at(geom, grow=(0, 0, 0)):
    return tuple(geom.lo[d] - grow[d] for d in range(3))


def omega_from_w(u, v, w, vel_begin, z, z_begin, geom, ngrow=0):
    """Source/MassFlux.cpp:51-74: Omega on the valid z-faces grown by ngrow;
    returns (array, begin)."""
    g3 = (ngrow, ngrow, ngrow)
    om = np.zeros(_shape_at(geom, 4, g3))
    b = _begin_at(geom, g3)
    lib().oracle_omega_from_w(_v(om, b), _v(u, vel_begin), _v(v, vel_begin), _v(w, vel_begin), _v(z, z_begin),
                              geom, ngrow)
    return om[0], b


def w_from_omega(u, v, omega, vel_begin, omega_begin, z, z_begin, geom):
    """Source/MassFlux.cpp:76-99: w on the valid z-faces (nz+1, ny, nx)."""
    w = np.zeros(_shape_at(geom, 4))
    lib().oracle_w_from_omega(_v(w, tuple(geom.lo)), _v(u, vel_begin), _v(v, vel_begin), _v(omega, omega_begin),
                              _v(z, z_begin), geom)
    return w[0]


def transport_fluxes(u, v, w, vel_begin, z, z_begin, ax, ay, area_begin, geom):
    """Source/MassFlux.cpp:101-127: (Fx, Fy, Fz) on the valid faces and one
    ghost face, each with begin = lo - 1."""
    g1 = (1, 1, 1)
    F = [np.zeros(_shape_at(geom, 1 << d, g1)) for d in range(3)]
    b = _begin_at(geom, g1)
    lib().oracle_transport_fluxes(_v(F[0], b), _v(F[1], b), _v(F[2], b), _v(u, vel_begin), _v(v, vel_begin),
                                  _v(w, vel_begin), _v(z, z_begin), _v(ax, area_begin), _v(ay, area_begin), geom)
    return F[0][0], F[1][0], F[2][0], b


def velocity_divergence(u, v, w, vel_begin, z, z_begin, ax, ay, az, area_begin, J, J_begin, geom):
    """Source/MassFlux.cpp:129-136: D(U) on the valid cells (nz, ny, nx)."""
    div = np.zeros(_shape_at(geom, 0))
    om = np.zeros(_shape_at(geom, 4))
    lo = tuple(geom.lo)
    lib().oracle_velocity_divergence(_v(div, lo), _v(u, vel_begin), _v(v, vel_begin), _v(w, vel_begin),
                                     _v(z, z_begin), _v(ax, area_begin), _v(ay, area_begin), _v(az, area_begin),
                                     _v(J, J_begin), geom, _v(om, lo))
    return div[0]


def pressure_gradient(p, p_begin, z, z_begin, J, J_begin, ax, ay, area_begin, geom):
    """Source/MassFlux.cpp:138-179: (Gx, Gy, Gz) on the valid faces."""
    G = [np.zeros(_shape_at(geom, 1 << d)) for d in range(3)]
    lo = tuple(geom.lo)
    lib().oracle_pressure_gradient(_v(G[0], lo), _v(G[1], lo), _v(G[2], lo), _v(p, p_begin), _v(z, z_begin),
                                   _v(J, J_begin), _v(ax, area_begin), _v(ay, area_begin), geom)
    return G[0][0], G[1][0], G[2][0]


# --- pseudo-time (oracle/pseudotime.c) ---------------------------------------------------------------------

K_MIN_RATE = 1.0e-12


def local_dtau(u, v, w, vel_begin, z, z_begin, J, J_begin, geom, cfl, nu=None, nu_begin=None):
    """Source/PseudoTime.cpp:12-51: dtau on the valid cells (nz, ny, nx)."""
    lo = tuple(geom.lo)
    d = np.zeros(_shape_at(geom, 0))
    om = np.zeros(_shape_at(geom, 4))
    rc = lib().oracle_local_dtau(_v(d, lo), _v(u, vel_begin), _v(v, vel_begin), _v(w, vel_begin), _v(z, z_begin),
                                 _v(J, J_begin), _v(nu, nu_begin if nu_begin is not None else lo), float(cfl), geom,
                                 _v(om, lo))
    if rc != 0:
        raise OracleError("local_dtau: cfl must be finite and > 0")
    return d[0]


def ser_cfl(cfl, r, cfl0, cfl_max, p, growth):
    """Source/PseudoTime.cpp:53-61."""
    return lib().oracle_ser_cfl(cfl, r, cfl0, cfl_max, p, growth)


def face_dtau(dtau_cc, geom):
    """Source/Predictor.cpp:30-77: (dx, dy, dz) face means of valid dtau (nz, ny, nx)."""
    lo = tuple(geom.lo)
    d = np.ascontiguousarray(dtau_cc, dtype=np.float64)[None]
    F = [np.zeros(_shape_at(geom, 1 << q)) for q in range(3)]
    if lib().oracle_face_dtau(_v(F[0], lo), _v(F[1], lo), _v(F[2], lo), _v(d, lo), geom) != 0:
        raise OracleError("face_dtau: dtau_cc must be finite and > 0 everywhere")
    return F[0][0], F[1][0], F[2][0]


# --- implicit operator A1 (oracle/implicit.c) --------------------------------------------------------------

def implicit_coefficients(loc, F, F_begin, g, g_begin, z, z_begin, J, J_begin, geom, logz=False):
    """Source/ImplicitOperator.cpp:95-121: A (7, nz', ny', nx') at loc (valid points, begin lo)."""
    A = np.zeros((7,) + _shape_at(geom, loc)[1:])
    rc = lib().oracle_implicit_coefficients(loc, _v(F[0], F_begin), _v(F[1], F_begin), _v(F[2], F_begin),
                                            _v(g, g_begin), _v(z, z_begin), _v(J, J_begin), geom, int(logz),
                                            _v(A, tuple(geom.lo)))
    if rc != 0:
        raise OracleError("implicit_coefficients: unknown location")
    return A


def apply_implicit(loc, A, q, q_begin, geom):
    """Source/ImplicitOperator.cpp:123-135: A1 q on the valid points at loc."""
    out = np.zeros(_shape_at(geom, loc))
    lib().oracle_apply_implicit(loc, _v(A, tuple(geom.lo)), _v(q, q_begin), geom, _v(out, tuple(geom.lo)))
    return out[0]


# --- convection and diffusion (oracle/transport.c) ---------------------------------------------------------

VAN_LEER, VAN_ALBADA, UPWIND = 0, 1, 2


def cv_flux_box(loc, geom, d):
    """(lo, hi) of cv_flux_box(loc, valid, d) (Source/CVFluxes.cpp:54-67)."""
    lo, hi = [], []
    for a in range(3):
        nod = (loc >> a) & 1
        l, h = geom.lo[a], geom.hi[a] + nod
        if a == d:
            l, h = (geom.lo[a] - 1, geom.hi[a] + 1) if nod else (geom.lo[a], geom.hi[a] + 1)
        lo.append(l)
        hi.append(h)
    return tuple(lo), tuple(hi)


def _cv_arrays(loc, geom):
    G, begins = [], []
    for d in range(3):
        lo, hi = cv_flux_box(loc, geom, d)
        G.append(np.zeros((1, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)))
        begins.append(lo)
    return G, begins


def convective_fluxes(loc, q, q_begin, F, F_begin, geom, limiter=VAN_LEER, eps=1e-5):
    """Source/Convection.cpp:117-132: ([G0, G1, G2], [begins])."""
    G, b = _cv_arrays(loc, geom)
    rc = lib().oracle_convective_fluxes(loc, _v(q, q_begin), _v(F[0], F_begin), _v(F[1], F_begin),
                                        _v(F[2], F_begin), limiter, float(eps), geom, _v(G[0], b[0]), _v(G[1], b[1]),
                                        _v(G[2], b[2]))
    if rc != 0:
        raise OracleError("convective_fluxes: bad location, limiter or eps")
    return G, b


def flux_divergence(loc, G, begins, J, J_begin, geom):
    """Source/CVFluxes.cpp:69-80: out on the valid points of loc."""
    out = np.zeros(_shape_at(geom, loc))
    lib().oracle_flux_divergence(loc, _v(G[0], begins[0]), _v(G[1], begins[1]), _v(G[2], begins[2]), _v(J, J_begin),
                                 geom, _v(out, tuple(geom.lo)))
    return out[0]


def convection(loc, q, q_begin, F, F_begin, J, J_begin, geom, limiter=VAN_LEER, eps=1e-5, bounded=False):
    """Source/Convection.cpp:134-141."""
    G, b = convective_fluxes(loc, q, q_begin, F, F_begin, geom, limiter, eps)
    out = flux_divergence(loc, G, b, J, J_begin, geom)[None].copy()
    if bounded:
        lib().oracle_subtract_transport_divergence(loc, _v(q, q_begin), _v(F[0], F_begin), _v(F[1], F_begin),
                                                   _v(F[2], F_begin), _v(J, J_begin), geom,
                                                   _v(out, tuple(geom.lo)))
    return out[0]


def scalar_diffusive_fluxes(q, q_begin, gamma, g_begin, z, z_begin, geom, logz=False):
    G, b = _cv_arrays(0, geom)
    lib().oracle_scalar_diffusive_fluxes(_v(q, q_begin), _v(gamma, g_begin), _v(z, z_begin), geom, int(logz),
                                         _v(G[0], b[0]), _v(G[1], b[1]), _v(G[2], b[2]))
    return G, b


def scalar_diffusion(q, q_begin, gamma, g_begin, z, z_begin, J, J_begin, geom, logz=False):
    """Source/Diffusion.cpp:79-85."""
    G, b = scalar_diffusive_fluxes(q, q_begin, gamma, g_begin, z, z_begin, geom, logz)
    return flux_divergence(0, G, b, J, J_begin, geom)


def momentum_diffusive_fluxes(c, vel, vel_begin, nu, nu_begin, z, z_begin, geom, logz=False):
    G, b = _cv_arrays(1 << c, geom)
    rc = lib().oracle_momentum_diffusive_fluxes(c, _v(vel[0], vel_begin), _v(vel[1], vel_begin),
                                                _v(vel[2], vel_begin), _v(nu, nu_begin), _v(z, z_begin), geom,
                                                int(logz), _v(G[0], b[0]), _v(G[1], b[1]), _v(G[2], b[2]))
    if rc != 0:
        raise OracleError("momentum_diffusive_fluxes: bad component")
    return G, b


def momentum_diffusion(vel, vel_begin, nu, nu_begin, z, z_begin, J, J_begin, geom, logz=False):
    """Source/Diffusion.cpp:111-119: [D_u, D_v, D_w] on the valid faces."""
    out = []
    for c in range(3):
        G, b = momentum_diffusive_fluxes(c, vel, vel_begin, nu, nu_begin, z, z_begin, geom, logz)
        out.append(flux_divergence(1 << c, G, b, J, J_begin, geom))
    return out


# --- reductions (oracle/reductions.c) ----------------------------------------------------------------------

def box_sum(a, begin, lo, hi):
    """Source/Reductions.cpp:11-53 over one box: column sums in k order, added in storage order."""
    return lib().oracle_box_sum(_v(a, begin), (C.c_int * 3)(*lo), (C.c_int * 3)(*hi))


def global_max(a, begin, lo, hi):
    """Source/Reductions.cpp:55-66."""
    return lib().oracle_global_max(_v(a, begin), (C.c_int * 3)(*lo), (C.c_int * 3)(*hi))


def scaled_sums(q, v, begin, s, loc, geom, offset=None):
    """Source/Residual.cpp:101-150 on one box: (sum |q~|, sum v~^2, count)."""
    out = (C.c_double * 3)()
    lib().oracle_scaled_sums(_v(q, begin), _v(v, begin), _v(offset, begin), float(s), int(loc), geom, out)
    return out[0], out[1], out[2]


def residual_scales(U, theta_s, k_s, H):
    """Source/Simplec.cpp:374-382."""
    out = (C.c_double * 6)()
    lib().oracle_residual_scales(U, theta_s, k_s, H, out)
    return list(out)


def _rviews(R, begin):
    keep = [a4(r if r.ndim == 4 else r[None], begin) for r in R]
    arr = (C.POINTER(OA4) * 6)(*[C.pointer(k) for k in keep])
    return arr, keep


def stage_norms(R, R_begin, u, v, w, vel_begin, tke, tke_begin, U, theta_s, k_s, H, geom, ustar=None,
                ustar_begin=None):
    """Source/Simplec.cpp:385-409: (res[6], monitors)."""
    arr, _keep = _rviews(R, R_begin)
    res, mon = (C.c_double * 6)(), (C.c_double * 3)()
    n = lib().oracle_stage_norms(arr, _v(u, vel_begin), _v(v, vel_begin), _v(w, vel_begin), _v(tke, tke_begin),
                                 _v(ustar, ustar_begin), U, theta_s, k_s, H, geom, res, mon)
    return list(res), list(mon[:n])


def confidence(R, R_begin, z, z_begin, U, theta_s, k_s, H, thresholds, geom):
    """Source/Simplec.cpp:411-445: (share, top height)."""
    arr, _keep = _rviews(R, R_begin)
    share, top = C.c_double(0), C.c_double(0)
    lib().oracle_confidence(arr, _v(z, z_begin), U, theta_s, k_s, H, (C.c_double * 6)(*thresholds), geom,
                            C.byref(share), C.byref(top))
    return share.value, top.value
