"""The terrain Poisson layer, mirroring the reference's PoissonSolver.

* ``PoissonSolver(mesh, opt)`` -- Source/Poisson.H:135-203 (the reference
  takes Mesh, Geometry and z_phys_nd; ``mesh`` here carries all three);
* ``solve(rhs, phi, coeffs, bc, fluxes)`` -> :class:`PoissonResult`;
* ``apply(Aphi, phi, coeffs, bc)``, ``fill_ghosts(phi, bc)``;
* ``face_divergence(mesh, F, div)`` -- Source/Poisson.H:209-210.

Boundary conditions are 6 codes in amrex::Orientation order (xlo, ylo, zlo,
xhi, yhi, zhi): PERIODIC, NEUMANN, DIRICHLET. Every call runs the HIP
library; refusals raise GustError with the reference's message.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .fab import Fab, Geometry

PERIODIC, NEUMANN, DIRICHLET = 0, 1, 2


def _stream(stream):
    if stream is not None:
        return C.c_void_p(int(stream))
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


@dataclass
class PoissonOptions:
    """Source/Poisson.H:40-69; precision 0 = FP64 V-cycle, 1 = FP32 V-cycle."""

    reltol: float = 1e-10
    abstol: float = 0.0
    max_iter: int = 200
    restart: int = 50
    verbose: int = 0
    precision: int = 0
    pre_sweeps: int = 2
    post_sweeps: int = 2
    coarse_sweeps: int = 16
    max_levels: int = 0
    fixed_steps: int = 0   # k > 0: k GMRES steps on the device, no host reads (graph capture)
    seam_copy: int = 1     # odd periodic sizes: seam neighbours from a copy (0: a separate neighbour pass)

    def c(self) -> _lib.PoissonOptionsC:
        o = _lib.PoissonOptionsC()
        for f, _t in _lib.PoissonOptionsC._fields_:
            setattr(o, f, getattr(self, f))
        return o


@dataclass
class PoissonResult:
    iterations: int
    initial_residual: float
    final_residual: float
    singular: bool
    removed_mean: float


@dataclass
class FaceCoeffs:
    """Source/Poisson.H:96-103: optional face coefficient Fabs (None = 1)."""

    x: Fab | None = None
    y: Fab | None = None
    z: Fab | None = None
    z_cross: Fab | None = None
    precond_ratio: float = 1.0

    def c(self):
        cc = _lib.FaceCoeffsC()
        keep = []
        for name in ("x", "y", "z", "z_cross"):
            f = getattr(self, name)
            if f is None:
                setattr(cc, name, C.POINTER(_lib.Array4)())
            else:
                a = f.array4()
                keep.append(a)
                setattr(cc, name, C.pointer(a))
        cc.precond_ratio = float(self.precond_ratio)
        return cc, keep


class DeviceMesh:
    """A tile's mesh on the device: geometry, node heights (one ghost node)
    and the metrics detJ, ax, ay, az and z_cc (one ghost). From a host
    TileMesh (uploaded), or from a TileFrame (built on the device by
    gust_build_mesh, bitwise the host builder's)."""

    def __init__(self, tile, device="cuda", dtype=None, stream=None):
        import torch
        dtype = dtype or torch.float64
        self.tile = tile
        self.geom: Geometry = tile.geom
        b = (-1, -1, -1)
        if getattr(tile, "z_nd", None) is None:
            self._build(tile, device, stream)
            return

        def dev(a):
            return Fab(torch.from_numpy(np.ascontiguousarray(a)[None]).to(device=device, dtype=dtype), b)

        self.z_nd = dev(tile.z_nd)
        self.detJ = dev(tile.detJ)
        self.ax = dev(tile.ax)
        self.ay = dev(tile.ay)
        self.az = dev(tile.az)
        self.z_cc = dev(tile.z_cc)

    def _build(self, frame, device, stream):
        import torch
        from . import mesh as M
        nx, ny, nz = self.geom.n
        b = (-1, -1, -1)

        def zeros(shape):
            return Fab(torch.zeros((1,) + shape, dtype=torch.float64, device=device), b)

        self.z_nd = zeros((nz + 3, ny + 3, nx + 3))
        self.detJ = zeros((nz + 2, ny + 2, nx + 2))
        self.ax = zeros((nz + 2, ny + 2, nx + 3))
        self.ay = zeros((nz + 2, ny + 3, nx + 2))
        self.az = zeros((nz + 3, ny + 2, nx + 2))
        self.z_cc = zeros((nz + 2, ny + 2, nx + 2))
        ground = np.ascontiguousarray(M.ground_heights(frame.terrain, self.geom, frame.prob_lo, frame.prob_hi),
                                      dtype=np.float64)
        zl = np.ascontiguousarray(frame.zlevels, dtype=np.float64)
        g = self.geom.c()
        views = [self.z_nd.array4(), self.detJ.array4(), self.ax.array4(), self.ay.array4(), self.az.array4(),
                 self.z_cc.array4()]
        dp = C.POINTER(C.c_double)
        _lib.check(_lib.lib.gust_build_mesh(C.byref(g), ground.ctypes.data_as(dp), zl.ctypes.data_as(dp),
                                            *[C.byref(v) for v in views], _stream(stream)))


def _bc(bc):
    arr = (C.c_int32 * 6)(*[int(v) for v in bc])
    return arr


class PoissonSolver:
    def __init__(self, mesh: DeviceMesh, opt: PoissonOptions | None = None, stream=None):
        self.mesh = mesh
        self.opt = opt or PoissonOptions()
        h = C.c_void_p(0)
        g = mesh.geom.c()
        self._keep = [mesh.z_nd.array4(), mesh.detJ.array4(), mesh.ax.array4(), mesh.ay.array4(),
                      mesh.az.array4()]
        o = self.opt.c()
        _lib.check(_lib.lib.gust_poisson_create(C.byref(g), *[C.byref(a) for a in self._keep], C.byref(o),
                                                C.byref(h), _stream(stream)))
        self._h = h
        lv = C.c_int32(0)
        _lib.check(_lib.lib.gust_poisson_levels(self._h, C.byref(lv)))
        self.levels = lv.value

    def solve(self, rhs: Fab, phi: Fab, coeffs: FaceCoeffs | None, bc, fluxes=None, stream=None) -> PoissonResult:
        cc, keep = (coeffs or FaceCoeffs()).c()
        res = _lib.PoissonResultC()
        r, p = rhs.array4(), phi.array4()
        if fluxes is not None:
            fa = [f.array4() for f in fluxes]
            fp = [C.byref(a) for a in fa]
        else:
            fp = [None, None, None]
        _lib.check(_lib.lib.gust_poisson_solve(self._h, C.byref(r), C.byref(p), C.byref(cc), _bc(bc), *fp,
                                               C.byref(res), _stream(stream)))
        del keep
        return PoissonResult(res.iterations, res.initial_residual, res.final_residual, bool(res.singular),
                             res.removed_mean)

    def apply(self, Aphi: Fab, phi: Fab, coeffs: FaceCoeffs | None, bc, stream=None) -> None:
        cc, keep = (coeffs or FaceCoeffs()).c()
        a, p = Aphi.array4(), phi.array4()
        _lib.check(_lib.lib.gust_poisson_apply(self._h, C.byref(a), C.byref(p), C.byref(cc), _bc(bc),
                                               _stream(stream)))
        del keep

    def fill_ghosts(self, phi: Fab, bc, stream=None) -> None:
        p = phi.array4()
        _lib.check(_lib.lib.gust_poisson_fill_ghosts(self._h, C.byref(p), _bc(bc), _stream(stream)))

    def fluxes(self, phi: Fab, coeffs: FaceCoeffs | None, bc, F, stream=None) -> None:
        cc, keep = (coeffs or FaceCoeffs()).c()
        p = phi.array4()
        fa = [f.array4() for f in F]
        _lib.check(_lib.lib.gust_poisson_fluxes(self._h, C.byref(p), C.byref(cc), _bc(bc),
                                                *[C.byref(a) for a in fa], _stream(stream)))
        del keep

    def vcycle(self, b, x, coeffs: FaceCoeffs | None, bc, setup=True, stream=None) -> None:
        """x = M^-1 b (one V-cycle) on compact float64 device tensors."""
        cc, keep = (coeffs or FaceCoeffs()).c()
        _lib.check(_lib.lib.gust_poisson_vcycle(self._h, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()),
                                                C.byref(cc), _bc(bc), int(bool(setup)), _stream(stream)))
        del keep

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib.gust_poisson_destroy(self._h)
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def face_divergence(mesh: DeviceMesh, F, div: Fab, stream=None) -> None:
    """Source/Poisson.cpp:733-764 with the mesh's areas and J."""
    g = mesh.geom.c()
    arrs = [mesh.ax.array4(), mesh.ay.array4(), mesh.az.array4(), mesh.detJ.array4()] + [f.array4() for f in F] + \
        [div.array4()]
    _lib.check(_lib.lib.gust_face_divergence(C.byref(g), *[C.byref(a) for a in arrs], _stream(stream)))
