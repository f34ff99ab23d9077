"""The C-ABI library loads, exports every symbol include/*.h declares, and
refuses bad input with the reference's messages -- all without a GPU (the
refusals happen before any device call)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from tests.conftest import ROOT


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(gust_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    import ablsteadyrans_amd._lib as L
    names = _declared()
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(L.lib, n)]
    assert not missing, missing
    # and the Python binding covers each of them
    assert set(names) <= set(L.exported_symbols()), set(names) - set(L.exported_symbols())


def test_version_and_empty_error():
    import ablsteadyrans_amd._lib as L
    assert b"gfx950" in L.lib.gust_version()


def test_option_refusals_name_their_keys():
    import ablsteadyrans_amd as G
    ok = G.LineSweepOptions(kind=G.LUSGS, ghosts=G.ZERO, overlap=2)
    assert G.check_line_options(ok) == []
    for bad, key in ((G.LineSweepOptions(kind=1, ghosts=1, overlap=-1), "steady.line_overlap"),
                     (G.LineSweepOptions(kind=1, ghosts=1, overlap=9), "steady.line_overlap"),
                     (G.LineSweepOptions(kind=1, ghosts=0, overlap=2), "steady.line_overlap"),
                     (G.LineSweepOptions(kind=1, alternate=2), "steady.line_alternate")):
        err = G.check_line_options(bad)
        assert len(err) == 1 and key in err[0], err


def _host_fab(geom, loc, ncomp, ngrow):
    import ablsteadyrans_amd as G
    return G.Fab.at_loc(geom, loc, ncomp, ngrow, backend="numpy")


def test_system_refusals_before_any_device_call():
    import ablsteadyrans_amd as G
    geom = G.Geometry.of_size(4, 3, 5)
    A = _host_fab(geom, G.CELL, G.N_COEFS, 0)
    ap = _host_fab(geom, G.CELL, 1, 0)
    rhs = _host_fab(geom, G.CELL, 1, 0)
    d = _host_fab(geom, G.CELL, 1, 1)
    with pytest.raises(G.GustError, match="gust: zline_jacobi: sweeps must be >= 1"):
        G.zline_jacobi(G.CELL, A, ap, rhs, 0, geom, d, stream=0)
    d0 = _host_fab(geom, G.CELL, 1, 0)
    with pytest.raises(G.GustError, match="delta needs at least 1 ghost"):
        G.zline_jacobi(G.CELL, A, ap, rhs, 1, geom, d0, stream=0)
    A6 = _host_fab(geom, G.CELL, 6, 0)
    with pytest.raises(G.GustError, match="implicit operator's components"):
        G.zline_jacobi(G.CELL, A6, ap, rhs, 1, geom, d, stream=0)
    # a face field on cell-centred data
    with pytest.raises(G.GustError, match="x-faces"):
        G.zline_jacobi(G.XFACE, A, ap, rhs, 1, geom, d, stream=0)
    with pytest.raises(G.GustError, match="steady.line_overlap"):
        G.zline_lusgs(G.CELL, A, ap, rhs, 1, geom, d, G.LineSweepOptions(kind=1, ghosts=1, overlap=5), stream=0)
    # the message is per call: a later failure replaces it
    with pytest.raises(G.GustError, match="sweeps"):
        G.zline_lusgs(G.CELL, A, ap, rhs, -1, geom, d, stream=0)


def test_array4_view_matches_amrex_indexing():
    import ablsteadyrans_amd as G
    f = G.Fab.alloc((2, -1, 0), (5, 3, 4), ncomp=3, ngrow=1, backend="numpy")
    a = f.array4()
    assert tuple(a.begin) == (1, -2, -1) and tuple(a.end) == (7, 5, 6)
    f.data[...] = np.arange(f.data.size, dtype=np.float64).reshape(f.data.shape)
    base = f.data.ctypes.data_as(C.POINTER(C.c_double))
    for (i, j, k, n) in ((1, -2, -1, 0), (4, 2, 3, 2), (6, 4, 5, 1)):
        off = (i - a.begin[0]) + (j - a.begin[1]) * a.jstride + (k - a.begin[2]) * a.kstride + n * a.nstride
        assert base[off] == f.data[n, k - a.begin[2], j - a.begin[1], i - a.begin[0]]


class _HostMesh:
    """The fields of a DeviceMesh as host FABs (refusal tests only)."""

    def __init__(self, nx=6, ny=5, nz=4):
        import ablsteadyrans_amd as G
        self.geom = G.Geometry.of_size(nx, ny, nz)
        b = (-1, -1, -1)
        f = lambda shape: G.Fab(np.zeros((1,) + shape), b)
        self.z_nd = f((nz + 3, ny + 3, nx + 3))
        self.detJ = f((nz + 2, ny + 2, nx + 2))
        self.ax = f((nz + 2, ny + 2, nx + 3))
        self.ay = f((nz + 2, ny + 3, nx + 2))
        self.az = f((nz + 3, ny + 2, nx + 2))


def test_mass_flux_refusals_before_any_device_call():
    import ablsteadyrans_amd as G
    from ablsteadyrans_amd import massflux as MF
    m = _HostMesh()
    g = m.geom
    u1, v1, w1 = (_host_fab(g, loc, 1, 1) for loc in (G.XFACE, G.YFACE, G.ZFACE))
    u0 = _host_fab(g, G.XFACE, 1, 0)
    om = _host_fab(g, G.ZFACE, 1, 0)
    with pytest.raises(G.GustError, match="gust: omega_from_w: ngrow must be 0 or 1"):
        MF.omega_from_w(m, u1, v1, w1, om, ngrow=2, stream=0)
    with pytest.raises(G.GustError, match="gust: omega_from_w: u needs at least 1 ghost cells"):
        MF.omega_from_w(m, u0, v1, w1, om, stream=0)
    with pytest.raises(G.GustError, match="gust: omega_from_w: omega needs at least 1 ghost cells"):
        MF.omega_from_w(m, _host_fab(g, G.XFACE, 1, 2), _host_fab(g, G.YFACE, 1, 2), w1, om, ngrow=1, stream=0)
    with pytest.raises(G.GustError, match="gust: transport_fluxes: u needs at least 2 ghost cells"):
        MF.transport_fluxes(m, u1, v1, w1, (u1, v1, w1), stream=0)
    with pytest.raises(G.GustError, match="gust: velocity_divergence: div is not on the mesh's BoxArray"):
        MF.velocity_divergence(m, u1, v1, w1, _host_fab(G.Geometry.of_size(3, 3, 3), G.CELL, 1, 0), stream=0)
    p0 = _host_fab(g, G.CELL, 1, 0)
    with pytest.raises(G.GustError, match="gust: pressure_gradient: p needs at least 1 ghost cells"):
        MF.pressure_gradient(m, p0, (u0, _host_fab(g, G.YFACE, 1, 0), om), stream=0)
    m.z_nd = G.Fab(np.zeros((1, 6, 7, 8)), (0, 0, 0))
    with pytest.raises(G.GustError, match="z_nd must be the mesh's node heights"):
        MF.w_from_omega(m, u1, v1, om, om, stream=0)


def test_jacobi_schedule_refuses_unknown_values():
    import ablsteadyrans_amd as LS
    import importlib
    L = importlib.import_module("ablsteadyrans_amd.line_sweeps")
    for bad in (-1, 5):
        with pytest.raises(Exception, match="jacobi_set_schedule: schedule must be 0"):
            LS.set_jacobi_schedule(bad)
    for ok in (LS.JACOBI_STREAMING, LS.JACOBI_STAGED, L.JACOBI_RULE, L.JACOBI_REGISTER, LS.JACOBI_AUTO):
        LS.set_jacobi_schedule(ok)
    L.jacobi_tune_reset()
    assert L.jacobi_tuned() == []


def test_newton_option_refusals_name_their_keys():
    """check_newton_options / check_continuation_options (Source/Newton.cpp:
    161-193): the defaults pass; each bad value is refused with the key it
    came from, before any device call."""
    from ablsteadyrans_amd import newton as NW
    assert NW.check_newton_options(NW.NewtonOptions()) == []
    cases = [("gmres_restart", 4, "steady.newton_gmres_restart"), ("gmres_restart", 201, "steady.newton_gmres_restart"),
             ("krylov_rtol", 1.0, "steady.newton_krylov_rtol"), ("preconditioner", 2, "steady.newton_preconditioner"),
             ("start_cfl", 200.0, "steady.cfl_max"), ("armijo_c", 0.5, "steady.newton_armijo_c"),
             ("fd_eps_scale", 1e-13, "steady.newton_fd_eps_scale"), ("ptc_growth", 0.5, "steady.newton_ptc_growth"),
             ("pc_denominator", 2, "pressure denominator")]
    for name, bad, key in cases:
        opt = NW.NewtonOptions(**{name: bad})
        msgs = NW.check_newton_options(opt)
        assert len(msgs) == 1 and key in msgs[0], (name, msgs)
    both = NW.check_newton_options(NW.NewtonOptions(gmres_restart=1, lambda_min=2.0))
    assert len(both) == 2
