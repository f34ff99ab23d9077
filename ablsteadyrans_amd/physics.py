"""The residual of the steady RANS equations and its physics over device
FABs (Source/Residual.cpp, Source/Closure.cpp, Source/SurfaceLayer.cpp,
Source/BoundaryConditions.cpp, Source/Forcing.cpp), forwarding to
``libgust_amd.so``. Option names and defaults are the reference's
(turb.*, bc.*, forcing.*); the sides of a tile are periodic.

A :class:`DeviceState` holds u, v, w, p, theta, k with two ghost layers; the
mesh is a :class:`~ablsteadyrans_amd.poisson.DeviceMesh`."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

from . import _lib
from .fab import CELL, XFACE, YFACE, ZFACE, Fab, Geometry, LineSweepOptions
from .poisson import DeviceMesh, _stream

PERIODIC, INFLOW, OUTFLOW, SLIP_WALL, NO_SLIP_WALL, SURFACE_LAYER, GEOSTROPHIC = range(7)
EQ_U, EQ_V, EQ_W, EQ_THETA, EQ_TKE = 10, 11, 12, 0, 1
SS_USTAR, SS_SPEED, SS_D0, SS_TAU0, SS_TAU1, SS_TAU2, SS_DRAG, SS_KWALL, SS_AREA, SS_INVL, SS_N = range(11)
VAN_LEER, VAN_ALBADA, UPWIND = 0, 1, 2
FIELDS = ("u", "v", "w", "p", "theta", "tke")
LOCS = (XFACE, YFACE, ZFACE, CELL, CELL, CELL)


@dataclass
class ClosureOptions:
    """Source/Closure.H:78-105 (turb.*)."""
    Cmu0: float = 0.5562
    Cb: float = 0.35
    Rt_crit: float = -1.0
    Rt_min: float = -3.0
    max_geom_lscale: float = 30.0
    tke_floor: float = 0.0
    sigma_k: float = 1.0
    nu_molecular: float = 0.0
    alpha_molecular: float = 0.0
    kappa: float = 0.41

    def c(self):
        return _lib.ClosureOptionsC(self.Cmu0, self.Cb, self.Rt_crit, self.Rt_min, self.max_geom_lscale,
                                    self.tke_floor, self.sigma_k, self.nu_molecular, self.alpha_molecular, self.kappa)


@dataclass
class SurfaceLayerOptions:
    """Source/SurfaceLayer.H:100-146 (bc.surface_*): neutral, first centre."""
    z0: float = 0.1
    Cmu0: float = 0.5562
    kappa: float = 0.41
    wall_k_dirichlet: bool = True
    denominator_floor_one: bool = False

    def c(self):
        return _lib.SurfaceOptionsC(self.z0, self.Cmu0, self.kappa, 0 if self.wall_k_dirichlet else 1,
                                    1 if self.denominator_floor_one else 0)


@dataclass
class ForcingOptions:
    """Source/Forcing.H:40-96 (forcing.*)."""
    gravity: float = 9.81
    theta0: float = 300.0
    coriolis: bool = False
    coriolis_3d: bool = True
    geostrophic: bool = False
    latitude: float = 0.0
    rotation_period: float = 86400.0
    ug: float = 0.0
    vg: float = 0.0
    pressure_gradient: tuple = (0.0, 0.0, 0.0)

    def c(self):
        f = _lib.ForcingC(self.gravity, self.theta0, int(self.coriolis), int(self.coriolis_3d),
                          int(self.geostrophic), self.latitude, self.rotation_period, self.ug, self.vg)
        for d in range(3):
            f.pressure_gradient[d] = self.pressure_gradient[d]
        return f


@dataclass
class Damping:
    """bc.rayleigh_* and bc.sponge_* (Source/BoundaryConditions.H:107-121):
    ``fields`` among u, v, w, theta, tke; ``ztop`` the mesh's top level;
    ``prob_lo``/``prob_hi`` the domain's x and y extent (the sponge's distances)."""
    rayleigh_zdamp: float = 0.0
    rayleigh_dampcoef: float = 0.0
    fields: tuple = ("u", "v", "w", "theta", "tke")
    above_ground: bool = False
    sponge_width: float = 0.0
    sponge_strength: float = 0.0
    ztop: float = 0.0
    prob_lo: tuple = (0.0, 0.0)
    prob_hi: tuple = (0.0, 0.0)

    def c(self):
        d = _lib.DampingC()
        d.rayleigh_zdamp, d.rayleigh_dampcoef = float(self.rayleigh_zdamp), float(self.rayleigh_dampcoef)
        for n, name in enumerate(("u", "v", "w", "theta", "tke")):
            d.rayleigh_fields[n] = 1 if name in self.fields else 0
        d.rayleigh_above_ground = 1 if self.above_ground else 0
        d.sponge_width, d.sponge_strength = float(self.sponge_width), float(self.sponge_strength)
        d.ztop = float(self.ztop)
        for k in range(2):
            d.prob_lo[k], d.prob_hi[k] = float(self.prob_lo[k]), float(self.prob_hi[k])
        return d


@dataclass
class ResidualInputs:
    """Source/Residual.H:90-128 on this build's scope. ``closure`` None takes
    the given cell diffusivities (Fabs with one ghost); ``faces`` in
    amrex::Orientation order (xlo, ylo, zlo, xhi, yhi, zhi), periodic exactly
    where the geometry is, None: no boundary conditions; ``inflow`` the
    inflow profile (initializer.InflowProfile) an inflow face takes."""
    closure: ClosureOptions | None = field(default_factory=ClosureOptions)
    nu: Fab | None = None
    gamma_theta: Fab | None = None
    gamma_tke: Fab | None = None
    surface: SurfaceLayerOptions = field(default_factory=SurfaceLayerOptions)
    forcing: ForcingOptions = field(default_factory=ForcingOptions)
    faces: tuple | None = (PERIODIC, PERIODIC, SURFACE_LAYER, PERIODIC, PERIODIC, SLIP_WALL)
    zhi_theta_grad: float = 0.0
    wall_velocity: tuple = ((0.0,) * 3,) * 6
    theta_ref_z: tuple = (0.0,)
    theta_ref: tuple = (300.0,)
    momentum_limiter: int = VAN_LEER
    momentum_eps: float = 1e-5
    theta_limiter: int = VAN_LEER
    theta_eps: float = 1e-5
    log_face_mean: bool = False
    inflow: object = None   # initializer.InflowProfile: the profile an inflow face takes
    damping: object = None  # Damping: the Rayleigh and sponge layers (off when None)

    def c(self):
        p = _lib.PhysicsC()
        p.use_closure = 1 if self.closure is not None else 0
        if self.closure is not None:
            p.closure = self.closure.c()
        self._keep = []
        for name in ("nu", "gamma_theta", "gamma_tke"):
            f = getattr(self, name)
            if f is not None:
                v = f.array4()
                self._keep.append(v)
                setattr(p, name, C.pointer(v))
        p.surface = self.surface.c()
        p.forcing = self.forcing.c()
        p.use_bc = 1 if self.faces is not None else 0
        for n in range(6):
            p.face[n] = int(self.faces[n]) if self.faces is not None else PERIODIC
            for d in range(3):
                p.wall_velocity[n][d] = float(self.wall_velocity[n][d])
        p.zhi_theta_grad = self.zhi_theta_grad
        n = len(self.theta_ref_z)
        if n != len(self.theta_ref) or not 1 <= n <= 64:
            raise _lib.GustError("gust: residual: the reference profile needs 1 to 64 (z, theta) points")
        p.n_theta_ref = n
        for m in range(n):
            p.theta_ref_z[m] = float(self.theta_ref_z[m])
            p.theta_ref[m] = float(self.theta_ref[m])
        p.momentum_limiter, p.theta_limiter = int(self.momentum_limiter), int(self.theta_limiter)
        p.momentum_eps, p.theta_eps = float(self.momentum_eps), float(self.theta_eps)
        p.face_mean = 1 if self.log_face_mean else 0
        _lib.check(_lib.lib.gust_inflow_default_profile(C.byref(p.inflow)))
        if self.inflow is not None:
            ic, keep = self.inflow.c()
            p.inflow = ic
            self._keep.append(keep)
        if self.damping is not None:
            p.damping = self.damping.c()
        return p


class DeviceState:
    """u, v, w, p, theta, k on the device, two ghost layers each (Source/State.H:13)."""

    def __init__(self, geom: Geometry, **kw):
        self.geom = geom
        self.fabs = {n: Fab.at_loc(geom, loc, 1, 2, **kw) for n, loc in zip(FIELDS, LOCS)}

    def __getitem__(self, name) -> Fab:
        return self.fabs[name]

    def c(self):
        s = _lib.StateC()
        self._views = [self.fabs[n].array4() for n in FIELDS]
        for n in range(6):
            s.f[n] = C.pointer(self._views[n])
        return s


def _mesh_c(mesh: DeviceMesh):
    views = [mesh.z_nd.array4(), mesh.detJ.array4(), mesh.ax.array4(), mesh.ay.array4(), mesh.az.array4()]
    m = _lib.MeshC(*[C.pointer(v) for v in views])
    return m, views


class ClosureFields:
    """Source/Closure.H:112-127: nu, gamma_theta, gamma_tke (one ghost);
    Ps, Pb, eps, dSdk (valid cells)."""

    def __init__(self, geom: Geometry, **kw):
        self.fabs = {n: Fab.at_loc(geom, CELL, 1, 1, **kw) for n in ("nu", "gamma_theta", "gamma_tke")}
        self.fabs.update({n: Fab.at_loc(geom, CELL, 1, 0, **kw) for n in ("Ps", "Pb", "eps", "dSdk")})

    def __getitem__(self, name) -> Fab:
        return self.fabs[name]

    def c(self):
        self._views = {n: f.array4() for n, f in self.fabs.items()}
        return _lib.ClosureFieldsC(*[C.pointer(self._views[n]) for n in
                                     ("nu", "gamma_theta", "gamma_tke", "Ps", "Pb", "eps", "dSdk")])


def surface_slab(geom: Geometry, **kw) -> Fab:
    """The surface state: SS_N components on the ground cells and one lateral ghost column."""
    lo, hi = geom.loc_box(CELL)
    return Fab.alloc((lo[0] - 1, lo[1] - 1, lo[2]), (hi[0] + 1, hi[1] + 1, lo[2]), SS_N, 0, **kw)


def height_above_ground(mesh: DeviceMesh, dist: Fab, stream=None) -> None:
    """Source/Closure.cpp:261-293 (dist: cells with one ghost)."""
    g = mesh.geom.c()
    _lib.check(_lib.lib.gust_height_above_ground(C.byref(g), C.byref(mesh.z_nd.array4()), C.byref(dist.array4()),
                                                 _stream(stream)))


def closure_evaluate(mesh: DeviceMesh, q: DeviceState, dist: Fab, z0: float, opt: ClosureOptions,
                     out: ClosureFields, gravity=9.81, theta0=300.0, stream=None) -> None:
    """Source/Closure.cpp:366-466."""
    g = mesh.geom.c()
    qc, oc, cf = q.c(), opt.c(), out.c()
    _lib.check(_lib.lib.gust_closure_evaluate(C.byref(g), C.byref(qc), C.byref(mesh.z_nd.array4()),
                                              C.byref(dist.array4()), float(z0), C.byref(oc), float(gravity),
                                              float(theta0), C.byref(cf), _stream(stream)))


def surface_state(mesh: DeviceMesh, q: DeviceState, opt: SurfaceLayerOptions, ss: Fab, stream=None) -> None:
    """Source/SurfaceLayer.cpp:355-428."""
    g = mesh.geom.c()
    oc = opt.c()
    _lib.check(_lib.lib.gust_surface_state(C.byref(g), C.byref(q["u"].array4()), C.byref(q["v"].array4()),
                                           C.byref(mesh.z_nd.array4()), C.byref(oc), C.byref(ss.array4()),
                                           _stream(stream)))


def fill_ghosts(mesh: DeviceMesh, q: DeviceState, inputs: ResidualInputs, ss: Fab | None = None,
                stream=None) -> None:
    """BoundaryConditions::fill_ghosts (Source/BoundaryConditions.cpp:597-666)."""
    g = mesh.geom.c()
    qc, pc = q.c(), inputs.c()
    mc, _keep = _mesh_c(mesh)
    _lib.check(_lib.lib.gust_fill_ghosts(C.byref(g), C.byref(qc), C.byref(mc), C.byref(pc),
                                         C.byref(ss.array4()) if ss is not None else None, _stream(stream)))


def residual(mesh: DeviceMesh, q: DeviceState, inputs: ResidualInputs, R: dict, fields: ClosureFields | None = None,
             ss: Fab | None = None, stream=None) -> None:
    """Source/Residual.cpp:159-256: R maps u, v, w, p, theta, tke to Fabs on
    the valid points."""
    g = mesh.geom.c()
    qc, pc = q.c(), inputs.c()
    mc, _keep = _mesh_c(mesh)
    views = [R[n].array4() for n in FIELDS]
    rc = _lib.ResidualsC()
    for n in range(6):
        rc.R[n] = C.pointer(views[n])
    cf = fields.c() if fields is not None else None
    _lib.check(_lib.lib.gust_residual(C.byref(g), C.byref(qc), C.byref(mc), C.byref(pc), C.byref(rc),
                                      C.byref(cf) if cf is not None else None,
                                      C.byref(ss.array4()) if ss is not None else None, _stream(stream)))


def implicit_rows(eq: int, mesh: DeviceMesh, F, inputs: ResidualInputs, A: Fab, fields: ClosureFields | None = None,
                  ss: Fab | None = None, stream=None) -> None:
    """Source/Residual.cpp:272-304: A (7 components) of eq from the transport fluxes F."""
    g = mesh.geom.c()
    pc = inputs.c()
    mc, _keep = _mesh_c(mesh)
    fv = [f.array4() for f in F]
    fc = _lib.FacesC()
    for d in range(3):
        fc.c[d] = C.pointer(fv[d])
    cf = fields.c() if fields is not None else None
    _lib.check(_lib.lib.gust_implicit_rows(int(eq), C.byref(g), C.byref(fc), C.byref(mc), C.byref(pc),
                                           C.byref(cf) if cf is not None else None,
                                           C.byref(ss.array4()) if ss is not None else None, C.byref(A.array4()),
                                           _stream(stream)))


@dataclass
class SimplecOptions:
    """Source/Simplec.H:71-83 (simplec.*) and the stage's settings:
    steady.zline_sweeps, the sweep options, turb.tke_floor, the residual
    scales, the scale height (residual_scale_height) and the Poisson
    controls (precision 1: FP32 V-cycles under FP64 GMRES; the coarsest
    level's zebra sweeps; the most multigrid levels, 0 = coarsen while x or
    y halves)."""
    pressure_reltol: float = 1.0e-3
    alpha_p: float = 1.0
    rotation_relax: float = 0.3
    w_vnn_cap: float = 0.0
    w_vnn_cap_slope_margin: float = 4.0
    sweeps: int = 2
    tke_floor: float = 0.0
    height: float = 1.0
    poisson_precision: int = 0
    poisson_max_iter: int = 200
    poisson_coarse_sweeps: int = 16
    poisson_max_levels: int = 0
    sweep_precision: int = 0
    sweep: LineSweepOptions | None = None   # steady.sweep and steady.line_* (None: z-line Jacobi)

    @property
    def mixed(self) -> bool:
        """FP32 sweeps and V-cycles inside the FP64 defect correction."""
        return self.poisson_precision == 1 and self.sweep_precision == 1

    def c(self):
        o = _lib.SimplecOptionsC()
        _lib.check(_lib.lib.gust_simplec_default_options(C.byref(o)))
        o.pressure_reltol, o.alpha_p, o.rotation_relax = self.pressure_reltol, self.alpha_p, self.rotation_relax
        o.w_vnn_cap, o.w_vnn_cap_slope_margin = self.w_vnn_cap, self.w_vnn_cap_slope_margin
        o.sweeps, o.tke_floor, o.height = int(self.sweeps), self.tke_floor, self.height
        o.poisson.precision = int(self.poisson_precision)
        o.poisson.max_iter = int(self.poisson_max_iter)
        o.poisson.coarse_sweeps = int(self.poisson_coarse_sweeps)
        o.poisson.max_levels = int(self.poisson_max_levels)
        o.sweep_precision = int(self.sweep_precision)
        if self.sweep is not None:
            o.sweep = self.sweep.c()
        return o


@dataclass
class DriverOptions:
    """Source/Driver.H:58-66 (steady.*) with the criterion (SteadyCriterion.H:13-33)."""
    max_iterations: int = 10000
    cfl0: float = 1.0
    cfl_max: float = 10.0
    ser_exponent: float = 1.0
    ser_max_growth: float = 2.0
    residual_tol: float = 5.0e-5
    residual_drop_orders: float = 4.0
    residual_ref_iterations: int = 5
    monitor_rel_tol: float = 1.0e-3
    monitor_window: int = 50

    def c(self):
        crit = _lib.SteadyParamsC(self.residual_tol, self.residual_drop_orders, self.residual_ref_iterations,
                                  self.monitor_rel_tol, self.monitor_window)
        return _lib.DriverOptionsC(int(self.max_iterations), self.cfl0, self.cfl_max, self.ser_exponent,
                                   self.ser_max_growth, crit)


class SimplecStage:
    """Source/Simplec.H:188-230: one SIMPLEC iteration per call on a tile.
    The mesh and the inputs' fields must outlive the stage."""

    def __init__(self, mesh: DeviceMesh, inputs: ResidualInputs, opt: SimplecOptions | None = None, stream=None):
        self.mesh, self.inputs, self.opt = mesh, inputs, opt or SimplecOptions()
        g = mesh.geom.c()
        self._pc = inputs.c()
        self._mc, self._mkeep = _mesh_c(mesh)
        oc = self.opt.c()
        h = C.c_void_p(0)
        _lib.check(_lib.lib.gust_simplec_create(C.byref(g), C.byref(self._mc), C.byref(self._pc), C.byref(oc),
                                                C.byref(h), _stream(stream)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib.lib is not None:
            _lib.lib.gust_simplec_destroy(h)
            self._h = None

    def init(self, q: DeviceState, stream=None) -> None:
        """Source/Simplec.cpp:338-359."""
        qc = q.c()
        _lib.check(_lib.lib.gust_simplec_init(self._h, C.byref(qc), _stream(stream)))

    def reset(self, stream=None) -> None:
        """Forget the closure's history: the next run starts as on a fresh
        stage."""
        _lib.check(_lib.lib.gust_simplec_reset(self._h, _stream(stream)))

    def iterate(self, q: DeviceState, dtau: Fab, stream=None):
        """Source/Simplec.cpp:452-506; returns the iteration's info."""
        qc = q.c()
        info = _lib.SimplecInfoC()
        _lib.check(_lib.lib.gust_simplec_iterate(self._h, C.byref(qc), C.byref(dtau.array4()), C.byref(info),
                                                 _stream(stream)))
        return info

    def norms(self):
        """(residual_norms, monitors) of the last iteration's R(q)."""
        res, mon, n = (C.c_double * 6)(), (C.c_double * 3)(), C.c_int32(0)
        _lib.check(_lib.lib.gust_simplec_norms(self._h, res, mon, C.byref(n)))
        return list(res), list(mon[:n.value])

    def nu_view(self) -> _lib.Array4:
        v = _lib.Array4()
        _lib.check(_lib.lib.gust_simplec_nu(self._h, C.byref(v)))
        return v

    def set_pressure_steps(self, steps: int) -> None:
        """k >= 1: the pressure solve runs k fixed GMRES steps on the device
        (what a captured iteration needs); 0: GMRES to the tolerance."""
        _lib.check(_lib.lib.gust_simplec_set_pressure_steps(self._h, int(steps)))

    def capture(self, q: DeviceState, dtau: Fab, stream=None) -> None:
        """Records one iteration on (q, dtau) as a HIP graph (not run); run
        one iteration on them first, with the pressure steps fixed."""
        self._capture_views = (q.c(), dtau.array4())
        qc, dv = self._capture_views
        _lib.check(_lib.lib.gust_simplec_capture(self._h, C.byref(qc), C.byref(dv), _stream(stream)))

    def replay(self, stream=None):
        """Runs the captured iteration (one launch, one wait, the checks);
        returns its info."""
        info = _lib.SimplecInfoC()
        _lib.check(_lib.lib.gust_simplec_replay(self._h, C.byref(info), _stream(stream)))
        return info

    def solve_graphed(self, q: DeviceState, opt: DriverOptions | None = None, pressure_steps: int = 0,
                      stream=None) -> _lib.DriverResultC:
        """Driver::run with the iteration captured and replayed
        (gust_steady_solve_graphed; pressure_steps 0: the first iteration's
        GMRES count + 1)."""
        qc = q.c()
        oc = (opt or DriverOptions()).c()
        r = _lib.DriverResultC()
        _lib.check(_lib.lib.gust_steady_solve_graphed(self._h, C.byref(qc), C.byref(oc), int(pressure_steps),
                                                      C.byref(r), _stream(stream)))
        return r

    def solve(self, q: DeviceState, opt: DriverOptions | None = None, stream=None) -> _lib.DriverResultC:
        """Driver::run with this stage (Source/Driver.cpp:403-492)."""
        qc = q.c()
        oc = (opt or DriverOptions()).c()
        r = _lib.DriverResultC()
        _lib.check(_lib.lib.gust_steady_solve(self._h, C.byref(qc), C.byref(oc), C.byref(r), _stream(stream)))
        return r
