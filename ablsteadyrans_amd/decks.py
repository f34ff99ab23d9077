"""The reference's steady test decks as ready configurations of this
library: the mesh, the residual inputs, the initializer's profile and
options, the stage and driver settings of each inputs file, and the
reference's measured iteration count.

* :func:`neutral_column` -- Tests/Canonical_RANS/Neutral_Column/inputs_steady_column
* :func:`neutral_abl_flat` -- Tests/Canonical_RANS/Neutral_ABL_Flat/inputs_steady_neutral_flat
* :func:`neutral_hill_3d` -- Tests/Canonical_RANS/Neutral_Hill_3D/inputs_steady_hill3d
* :func:`askervein` -- Tests/Real_Terrain/Askervein/inputs_steady_askervein
  (the terrain grid, extern_erf/askervein.txt, is the caller's)

:func:`prepare` builds the device mesh, the state at the deck's first guess
(the mass-consistent initializer) and the stage; :meth:`Deck.solve` runs the
steady driver.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import initializer as GI
from . import mesh as M
from . import physics as GP
from . import poisson as PS
from . import reductions as RD
from .fab import LUSGS, LineSweepOptions


@dataclass
class Deck:
    name: str
    tile: object
    inputs: GP.ResidualInputs
    profile: GI.InflowProfile
    init: GI.InitOptions
    stage: GP.SimplecOptions
    driver: GP.DriverOptions
    iterations_reference: int
    seconds_reference: float | None = None
    ranks_reference: int | None = None
    init_reltol: float = 1e-10
    state: object = field(default=None, repr=False)
    mesh: object = field(default=None, repr=False)
    simplec: object = field(default=None, repr=False)

    def solve(self):
        """Driver::run from the prepared state (Source/Driver.cpp:403-492)."""
        return self.simplec.solve(self.state, self.driver)


def _height(tile, prob_lo_z, prob_hi_z):
    return RD.residual_scale_height(tile.zlevels, prob_lo_z, prob_hi_z)


def neutral_column() -> Deck:
    """4 x 4 x 64 over 2560 x 2560 x 1000 m, periodic sides, surface layer
    z0 = 0.1 m, slip top, uniform theta, the force F = 1.225e-4 m/s^2
    (u* = 0.35 m/s), the uncapped closure with the logarithmic face mean, two
    sweeps, CFL 1 -> 50, steady at residual_tol 1e-8 and a 6-order drop; the
    first guess the log law with u* = 0.25 m/s and its equilibrium k (no
    projection: flat). The reference measured 2231 iterations."""
    geom = M.geometry((4, 4, 64), (0.0, 0.0, 0.0), (2560.0, 2560.0, 1000.0), (1, 1))
    tile = M.build_mesh(geom, (0.0, 0.0, 0.0), (2560.0, 2560.0, 1000.0), M.make_zlevels(64, 1000.0),
                        M.Terrain(kind="flat"))
    inputs = GP.ResidualInputs(closure=GP.ClosureOptions(max_geom_lscale=1.0e12),
                               surface=GP.SurfaceLayerOptions(z0=0.1),
                               forcing=GP.ForcingOptions(pressure_gradient=(1.225e-4, 0.0, 0.0)),
                               log_face_mean=True)
    return Deck("Neutral_Column", tile, inputs, GI.InflowProfile(kind=GI.LOG_LAW, ustar=0.25, z0=0.1, unbounded=True),
                GI.InitOptions(first_guess=1, projection=False), GP.SimplecOptions(sweeps=2, height=1000.0),
                GP.DriverOptions(max_iterations=20000, cfl0=1.0, cfl_max=50.0, residual_tol=1.0e-8,
                                 residual_drop_orders=6.0), 2231)


def neutral_hill_3d_lusgs() -> Deck:
    """Neutral_Hill_3D with the reference's recommended CPU line solver
    (steady.sweep = lusgs, lagged ghosts, 2 sweeps, CFL cap 16;
    docs/results.rst:10-23, 253-303): the reference measured 823 iterations
    in 51 s on four ranks (four boxes) and 860 on one box
    (docs/results.rst:357-380)."""
    d = neutral_hill_3d()
    d.name = "Neutral_Hill_3D (LU-SGS)"
    d.stage.sweeps = 2
    d.stage.sweep = LineSweepOptions(kind=LUSGS)
    d.driver.cfl_max = 16.0
    d.iterations_reference = 860
    d.seconds_reference = 51.0
    d.ranks_reference = 4
    return d


def neutral_abl_flat() -> Deck:
    """8 x 8 x 64 over 2560 x 2560 x 1000 m, periodic sides, surface layer
    z0 = 0.1 m, slip top with 0.03 K/m above the sounding (300 K to 700 m,
    309 K at 1000 m), the geostrophic wind (10, 0) m/s at latitude 90
    (horizontal Coriolis terms), the closure's defaults with a k floor of
    1e-6, four sweeps, CFL cap 2, the residual reference over 100
    iterations, a 2-order drop; the first guess the log law (u* = 0.35 m/s,
    z0 = 0.1 m, the wall-law length). The reference measured about 2850
    iterations (13 s on two ranks)."""
    geom = M.geometry((8, 8, 64), (0.0, 0.0, 0.0), (2560.0, 2560.0, 1000.0), (1, 1))
    tile = M.build_mesh(geom, (0.0, 0.0, 0.0), (2560.0, 2560.0, 1000.0), M.make_zlevels(64, 1000.0),
                        M.Terrain(kind="flat"))
    sounding = dict(theta_ref_z=(0.0, 700.0, 1000.0), theta_ref=(300.0, 300.0, 309.0))
    inputs = GP.ResidualInputs(closure=GP.ClosureOptions(tke_floor=1e-6), surface=GP.SurfaceLayerOptions(z0=0.1),
                               forcing=GP.ForcingOptions(coriolis=True, coriolis_3d=False, latitude=90.0,
                                                         rotation_period=125663.7061435917, geostrophic=True,
                                                         ug=10.0, vg=0.0),
                               zhi_theta_grad=0.03, **sounding)
    prof = GI.InflowProfile(kind=GI.LOG_LAW, ustar=0.35, z0=0.1, unbounded=True, **sounding)
    return Deck("Neutral_ABL_Flat", tile, inputs, prof, GI.InitOptions(first_guess=1),
                GP.SimplecOptions(sweeps=4, tke_floor=1e-6, height=1000.0),
                GP.DriverOptions(max_iterations=20000, cfl0=1.0, cfl_max=2.0, residual_ref_iterations=100,
                                 residual_drop_orders=2.0), 2850, 13.0, 2)


def neutral_hill_3d() -> Deck:
    """64 x 64 x 20 over 5120 x 5120 x 800 m, periodic sides, the round
    Witch-of-Agnesi hill (100 m, half width 500 m), the surface layer (z0 =
    0.1 m), a slip top with 0.03 K/m over the sounding, Rayleigh damping of w
    and theta above 550 m (0.2 1/s), the geostrophic wind (10, 0) m/s at
    latitude 90, the closure's defaults with a k floor of 1e-6, sixteen
    sweeps, CFL cap 4, rotation_relax 1, the residual reference over 100
    iterations, a 2-order drop; the first guess the log law (u* = 0.35 m/s,
    z0 = 0.1 m) made divergence-free. The reference measured 2092 iterations
    (148 s on four ranks)."""
    lo, hi = (-2560.0, -2560.0, 0.0), (2560.0, 2560.0, 800.0)
    geom = M.geometry((64, 64, 20), lo, hi, (1, 1))
    zl = M.make_zlevels(20, 800.0)
    tile = M.build_mesh(geom, lo, hi, zl, M.Terrain(kind="agnesi", hill_height=100.0, hill_half_width=500.0,
                                                       hill_shape="round"))
    sounding = dict(theta_ref_z=(0.0, 700.0, 1000.0), theta_ref=(300.0, 300.0, 309.0))
    prof = GI.InflowProfile(kind=GI.LOG_LAW, ustar=0.35, z0=0.1, **sounding)
    inputs = GP.ResidualInputs(closure=GP.ClosureOptions(tke_floor=1e-6), surface=GP.SurfaceLayerOptions(z0=0.1),
                               forcing=GP.ForcingOptions(coriolis=True, coriolis_3d=False, latitude=90.0,
                                                         rotation_period=125663.7061435917, geostrophic=True,
                                                         ug=10.0, vg=0.0),
                               zhi_theta_grad=0.03, inflow=prof,
                               damping=GP.Damping(rayleigh_zdamp=550.0, rayleigh_dampcoef=0.2, fields=("w", "theta"),
                                                  ztop=float(zl[-1]), prob_lo=lo[:2], prob_hi=hi[:2]),
                               **sounding)
    return Deck("Neutral_Hill_3D", tile, inputs, prof, GI.InitOptions(first_guess=1),
                GP.SimplecOptions(sweeps=16, tke_floor=1e-6, height=_height(tile, 0.0, 800.0), rotation_relax=1.0),
                GP.DriverOptions(max_iterations=20000, cfl0=1.0, cfl_max=4.0, residual_ref_iterations=100,
                                 residual_drop_orders=2.0), 2092, 148.0, 4)


ASKERVEIN_FACES = (GP.INFLOW, GP.SLIP_WALL, GP.SURFACE_LAYER, GP.OUTFLOW, GP.SLIP_WALL, GP.SLIP_WALL)


def askervein_tile(grid: M.ErfGrid):
    """30 x 30 cells of 200 m, 18 layers from 20 m growing by 1.08 (the
    computational top 440 m), ERF's Askervein grid, no smoothing."""
    lo, hi = (-3000.0, -3000.0, 0.0), (3000.0, 3000.0, 440.0)
    geom = M.geometry((30, 30, 18), lo, hi)
    zl = M.make_zlevels(18, 440.0, dz0=20.0, stretch_ratio=1.08)
    return M.build_mesh(geom, lo, hi, zl, M.Terrain(kind="erf_grid", grid=grid))


def askervein(grid: M.ErfGrid) -> Deck:
    """The Askervein mesh, inflow at xlo (the log law, u* = 0.6 m/s, z0 =
    0.03 m, the closure's length cap), outflow at xhi, slip walls along y,
    the surface layer (z0 = 0.03 m), a slip top, no gravity, a k floor of
    1e-6, four sweeps, CFL cap 2, the residual reference over 100
    iterations, the default 4-order drop; the mass-consistent first guess
    with alpha_v = 0.5. The reference measured 175 iterations (2 s on two
    ranks)."""
    tile = askervein_tile(grid)
    prof = GI.InflowProfile(kind=GI.LOG_LAW, ustar=0.6, z0=0.03, direction=0.0)
    inputs = GP.ResidualInputs(closure=GP.ClosureOptions(tke_floor=1e-6), surface=GP.SurfaceLayerOptions(z0=0.03),
                               forcing=GP.ForcingOptions(gravity=0.0), faces=ASKERVEIN_FACES, inflow=prof)
    return Deck("Askervein", tile, inputs, prof, GI.InitOptions(first_guess=1, alpha_h=1.0, alpha_v=0.5),
                GP.SimplecOptions(sweeps=4, tke_floor=1e-6, height=_height(tile, 0.0, 440.0)),
                GP.DriverOptions(max_iterations=3000, cfl0=1.0, cfl_max=2.0, residual_ref_iterations=100), 175,
                2.0, 2)


def open_faces(direction: float):
    """The face types of an open channel for a wind from `direction` degrees
    (counter-clockwise from +x): inflow on the side the wind enters and
    outflow opposite, along x when |cos| >= |sin| (y periodic), else along y
    (x periodic); the surface layer at the ground, a slip top. Returns
    (faces in Orientation order, periodicity)."""
    a = math.radians(direction)
    c, sn = math.cos(a), math.sin(a)
    if abs(c) >= abs(sn):
        xlo, xhi = (GP.INFLOW, GP.OUTFLOW) if c > 0 else (GP.OUTFLOW, GP.INFLOW)
        return (xlo, GP.PERIODIC, GP.SURFACE_LAYER, xhi, GP.PERIODIC, GP.SLIP_WALL), (0, 1)
    ylo, yhi = (GP.INFLOW, GP.OUTFLOW) if sn > 0 else (GP.OUTFLOW, GP.INFLOW)
    return (GP.PERIODIC, ylo, GP.SURFACE_LAYER, GP.PERIODIC, yhi, GP.SLIP_WALL), (1, 0)


def open_abl(name: str, n, extent, dz0: float, terrain, sweeps: int = 2, cfl_max: float = 10.0,
             drop_orders: float = 6.0, max_iterations: int = 20000, rotation_relax: float = 0.3,
             damping_fraction: float = 0.0, ref_iterations: int = 5, pressure_reltol: float = 1e-3,
             direction: float = 0.0, zlevels=None) -> Deck:
    """A neutral ABL over a tile of n = (nx, ny, nz) cells and extent (Lx,
    Ly) m: levels from dz0 growing by 1.08 (the top free; or the given
    zlevels), the log-law inflow u* = 0.4 m/s with z0 = 0.03 m from
    `direction` on the side the wind enters, outflow opposite, the other
    pair periodic (open_faces; direction 0: inflow at xlo, outflow at xhi, y
    periodic), the surface layer (z0 = 0.03 m), a slip top, Coriolis with
    f = 1e-4 1/s (latitude 90, horizontal terms only, no geostrophic
    forcing), the k-L closure's defaults, `sweeps` z-line sweeps, CFL 1 ->
    cfl_max, the default residual reference (5 iterations) and
    `drop_orders`; the first guess the inflow profile by height above the
    ground made divergence-free. damping_fraction > 0 adds Rayleigh damping
    of w above that fraction of the top (0.2 1/s)."""
    nx, ny, nz = n
    zl = np.asarray(zlevels, dtype=np.float64) if zlevels is not None else \
        M.make_zlevels(nz, 1.0, dz0=dz0, stretch_ratio=1.08)
    top = float(zl[-1])
    lo, hi = (0.0, 0.0, 0.0), (float(extent[0]), float(extent[1]), top)
    faces, periodic = open_faces(direction)
    geom = M.geometry((nx, ny, nz), lo, hi, periodic)
    tile = M.build_mesh(geom, lo, hi, zl, terrain)
    prof = GI.InflowProfile(kind=GI.LOG_LAW, ustar=0.4, z0=0.03, direction=direction)
    damping = None
    if damping_fraction > 0:
        damping = GP.Damping(rayleigh_zdamp=damping_fraction * top, rayleigh_dampcoef=0.2, fields=("w",),
                             ztop=top, prob_lo=lo[:2], prob_hi=hi[:2])
    inputs = GP.ResidualInputs(closure=GP.ClosureOptions(), surface=GP.SurfaceLayerOptions(z0=0.03),
                               forcing=GP.ForcingOptions(coriolis=True, coriolis_3d=False, latitude=90.0,
                                                         rotation_period=125663.7061435917),
                               faces=faces, inflow=prof, damping=damping)
    return Deck(name, tile, inputs, prof, GI.InitOptions(first_guess=1),
                GP.SimplecOptions(sweeps=sweeps, height=_height(tile, 0.0, top), rotation_relax=rotation_relax,
                                  pressure_reltol=pressure_reltol),
                GP.DriverOptions(max_iterations=max_iterations, cfl0=1.0, cfl_max=cfl_max,
                                 residual_drop_orders=drop_orders, residual_ref_iterations=ref_iterations), 0)


def flat_abl_64(**kw) -> Deck:
    """The flat neutral ABL benchmark: 64 x 64 x 48 cells over 5 x 5 km, levels
    from dz1 = 2 m growing by 1.08 (top 980.26 m), two sweeps, CFL 1 -> 10."""
    return open_abl("flat_abl_64x64x48", (64, 64, 48), (5000.0, 5000.0), 2.0, M.Terrain(kind="flat"), **kw)


def gaussian_hill_200(**kw) -> Deck:
    """The Gaussian-hill benchmark: h = 100 exp(-(r / 250 m)^2) m centred in a
    4 x 4 km tile at dx = dy = 20 m (200 x 200 cells), 64 levels from dz1 =
    1 m growing by 1.08 (top 1709.5 m), the open-channel ABL of open_abl."""
    t = M.Terrain(kind="hills", hills=[(2000.0, 2000.0, 100.0, 250.0)])
    return open_abl("gaussian_hill_200x200x64", (200, 200, 64), (4000.0, 4000.0), 1.0, t, **kw)


def fractal_tile_256():
    """The fractal-terrain tile of the kernel-level configuration: 256 x 256 x
    96 cells at dx = dy = 25 m (6.4 km), an isotropic Gaussian random
    surface with spectral slope k^-2 (seed 20264203, rms 80 m, clipped to
    0..400 m), the first layer 2 m with the ratio solved so the top is
    2000 m, BTF, no periodic sides."""
    nx, ny, nz = 256, 256, 96
    lo, hi = (0.0, 0.0, 0.0), (nx * 25.0, ny * 25.0, 2000.0)
    geom = M.geometry((nx, ny, nz), lo, hi)
    zl = M.make_zlevels(nz, 2000.0, dz0=2.0, stretch_ratio=0.0)
    h = M.fractal_heights(nx, ny, seed=20264203, rms=80.0, slope=-2.0, clip=(0.0, 400.0))
    return M.build_mesh(geom, lo, hi, zl, M.Terrain(kind="fractal", heights=h))


def prepare(deck: Deck, precision: int = 0) -> Deck:
    """The device mesh, the state at the deck's first guess and the stage
    (precision 1: FP32 multigrid V-cycles in every pressure correction;
    2: the mixed-precision form, FP32 V-cycles and FP32 z-line sweeps inside
    the FP64 defect correction)."""
    deck.mesh = PS.DeviceMesh(deck.tile)
    deck.state = GP.DeviceState(deck.tile.geom)
    faces = deck.inputs.faces
    if deck.init.projection:
        solver = PS.PoissonSolver(deck.mesh, PS.PoissonOptions(reltol=deck.init_reltol))
        GI.initialize_state(deck.mesh, deck.profile, deck.init, faces, solver, deck.state)
    else:
        GI.first_guess_state(deck.mesh, deck.profile, deck.init, deck.state)
    deck.stage.poisson_precision = 1 if precision >= 1 else 0
    deck.stage.sweep_precision = 1 if precision >= 2 else 0
    deck.simplec = GP.SimplecStage(deck.mesh, deck.inputs, deck.stage)
    return deck


def reinitialize(deck: Deck) -> None:
    """The state back at the deck's first guess (the stage and mesh kept,
    the closure's history cleared: the next solve is a fresh run's)."""
    if deck.simplec is not None:
        deck.simplec.reset()
    if deck.init.projection:
        solver = PS.PoissonSolver(deck.mesh, PS.PoissonOptions(reltol=deck.init_reltol))
        GI.initialize_state(deck.mesh, deck.profile, deck.init, deck.inputs.faces, solver, deck.state)
    else:
        GI.first_guess_state(deck.mesh, deck.profile, deck.init, deck.state)
