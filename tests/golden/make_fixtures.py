"""Writes the reference-held data fixtures used by the CPU tests into
tests/golden/ (data only: inputs and expected outputs, no source text).

  askervein_grid.npz   extern_erf/askervein.txt (ERF terrain grid: x, y, z)
  askervein_mesh.json  Tests/Discretization/Terrain_Mesh/Gold/askervein_mesh.json
                       (statistics of z_cc and detJ, 60 x 60 x 18 BTF mesh)
  poisson_askervein.npz  Tests/Discretization/Poisson_Askervein/Gold/:
                       phi of poisson_askervein_plt00000 (30 x 30 x 18,
                       i fastest) and the gold JSON numbers (div_before,
                       div_after, iterations, reltol) of both back ends.
  init_askervein.npz   Tests/Initialization/Mass_Consistent_Askervein/Gold/:
                       the cell fields of init_askervein_plt00000 (u, v, w
                       as the plotfile's face means, p, theta, tke, z,
                       detJ; (nz, ny, nx), i fastest) and the gold JSON
                       numbers.
  erf_hill2d.npz       Tests/Unit_Tests/Gold/erf_hill2d_{btf,stf,sullivan}:
                       ERF's z_phys and detJ (cell, (nz, ny, nx)) and the
                       node heights (Level_0/Nu_nd component 2 = z - k dz,
                       plus k dz; (nz+1, ny+1, nx+1)) of the 32 x 1 x 16
                       Agnesi hill for terrain.smoothing 0, 1, 2.
  askervein_hierarchy.npz  Tests/Discretization/Grid_Hierarchy/Gold/:
                       z and detJ (cell, (nz, ny, nx)) of levels 1 and 2 of
                       the three-level STF Askervein hierarchy
                       (askervein_hierarchy_level{1,2}_plt00000) and the
                       gold JSON (askervein_hierarchy_mesh.json) as text.

Run from the repository root with the reference tree at /root/reference.
"""
import json
import os
import re
import sys

import numpy as np

REF = os.environ.get("GUST_REFERENCE", "/root/reference")
OUT = os.path.dirname(os.path.abspath(__file__))


def read_plotfile_cells(plt, nvar_expected=1):
    """AMReX single-level plotfile, cell data: returns (lo, hi, array
    (nvar, nz, ny, nx)) assembled over the FABs listed in Level_0/Cell_H."""
    hdr = open(os.path.join(plt, "Header")).read().split("\n")
    nvar = int(hdr[1])
    names = hdr[2:2 + nvar]
    dom = hdr[2 + nvar + 6]
    m = re.match(r"\(\((\d+),(\d+),(\d+)\) \((\d+),(\d+),(\d+)\)", dom.strip())
    lo = tuple(int(m.group(i)) for i in (1, 2, 3))
    hi = tuple(int(m.group(i)) for i in (4, 5, 6))
    out = np.full((nvar, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1), np.nan)
    ch = open(os.path.join(plt, "Level_0", "Cell_H")).read()
    boxes = re.findall(r"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\)", ch)
    fods = re.findall(r"FabOnDisk: (\S+) (\d+)", ch)
    for (bx, (fname, off)) in zip(boxes, fods):
        with open(os.path.join(plt, "Level_0", fname), "rb") as f:
            f.seek(int(off))
            line = f.readline().decode()
            mm = re.search(r"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) \(\d+,\d+,\d+\)\) (\d+)", line)
            b = [int(mm.group(i)) for i in range(1, 7)]
            nc = int(mm.group(7))
            if "(8 7 6 5 4 3 2 1)" not in line or "(64 11 52 0 1 12 0 1023)" not in line:
                raise SystemExit("unexpected FAB real descriptor: " + line)
            n = (b[3] - b[0] + 1) * (b[4] - b[1] + 1) * (b[5] - b[2] + 1)
            data = np.frombuffer(f.read(8 * n * nc), dtype="<f8").reshape(nc, b[5] - b[2] + 1, b[4] - b[1] + 1,
                                                                            b[3] - b[0] + 1)
            out[:, b[2] - lo[2]:b[5] - lo[2] + 1, b[1] - lo[1]:b[4] - lo[1] + 1, b[0] - lo[0]:b[3] - lo[0] + 1] = data
    if np.isnan(out).any():
        raise SystemExit("plotfile does not cover its domain")
    return names, lo, hi, out


def read_vismf_nodal(mf_base, comp):
    """One component of a single-FAB VisMF MultiFab (Level_0/<name>_H and its
    data file) as (lo, array (nz', ny', nx'))."""
    ch = open(mf_base + "_H").read()
    fods = re.findall(r"FabOnDisk: (\S+) (\d+)", ch)
    if len(fods) != 1:
        raise SystemExit("expected one FAB in " + mf_base)
    fname, off = fods[0]
    with open(os.path.join(os.path.dirname(mf_base), fname), "rb") as f:
        f.seek(int(off))
        line = f.readline().decode()
        mm = re.search(r"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) \(\d+,\d+,\d+\)\) (\d+)", line)
        b = [int(mm.group(i)) for i in range(1, 7)]
        nc = int(mm.group(7))
        if "(8 7 6 5 4 3 2 1)" not in line or "(64 11 52 0 1 12 0 1023)" not in line:
            raise SystemExit("unexpected FAB real descriptor: " + line)
        n = (b[3] - b[0] + 1) * (b[4] - b[1] + 1) * (b[5] - b[2] + 1)
        data = np.frombuffer(f.read(8 * n * nc), dtype="<f8").reshape(nc, b[5] - b[2] + 1, b[4] - b[1] + 1,
                                                                        b[3] - b[0] + 1)
    return tuple(b[:3]), data[comp].copy()


def main():
    v = np.loadtxt(os.path.join(REF, "extern_erf", "askervein.txt"), dtype=np.float64).ravel()
    nx, ny = int(v[0]), int(v[1])
    np.savez_compressed(os.path.join(OUT, "askervein_grid.npz"), x=v[2:2 + nx], y=v[2 + nx:2 + nx + ny],
                        z=v[2 + nx + ny:])
    g = json.load(open(os.path.join(REF, "Tests/Discretization/Terrain_Mesh/Gold/askervein_mesh.json")))
    json.dump(g, open(os.path.join(OUT, "askervein_mesh.json"), "w"), indent=1, sort_keys=True)
    base = os.path.join(REF, "Tests/Discretization/Poisson_Askervein/Gold")
    names, lo, hi, phi = read_plotfile_cells(os.path.join(base, "poisson_askervein_plt00000"))
    assert names == ["phi"], names
    ga = json.load(open(os.path.join(base, "poisson_askervein.json")))
    gb = json.load(open(os.path.join(base, "poisson_askervein_mlmg.json")))
    np.savez_compressed(os.path.join(OUT, "poisson_askervein.npz"), phi=phi[0],
                        div_before=ga["div_before"], div_after_fft=ga["div_after"], iterations_fft=ga["iterations"],
                        div_after_mlmg=gb["div_after"], iterations_mlmg=gb["iterations"], reltol=ga["reltol"])
    base = os.path.join(REF, "Tests/Initialization/Mass_Consistent_Askervein/Gold")
    names, lo, hi, st = read_plotfile_cells(os.path.join(base, "init_askervein_plt00000"))
    assert names == ["u", "v", "w", "p", "theta", "tke", "z", "detJ"], names
    gi = json.load(open(os.path.join(base, "init_askervein.json")))
    np.savez_compressed(os.path.join(OUT, "init_askervein.npz"), **{n: st[q] for q, n in enumerate(names)},
                        **{k: gi[k] for k in ("div_before", "div_after", "iterations", "reltol", "speed_max", "wmax",
                                              "theta_min", "theta_max", "p_absmax", "k_min")})
    hill = {}
    for name in ("btf", "stf", "sullivan"):
        base = os.path.join(REF, "Tests/Unit_Tests/Gold", f"erf_hill2d_{name}")
        names, lo, hi, cells = read_plotfile_cells(base)
        assert names == ["z_phys", "detJ"], names
        nlo, zmk = read_vismf_nodal(os.path.join(base, "Level_0", "Nu_nd"), 2)
        assert nlo == (0, 0, 0)
        dz = 1000.0 / 16
        k = np.arange(zmk.shape[0], dtype=np.float64)[:, None, None]
        hill[f"{name}_z_phys"] = cells[0]
        hill[f"{name}_detJ"] = cells[1]
        hill[f"{name}_z_nd"] = zmk + k * dz
    np.savez_compressed(os.path.join(OUT, "erf_hill2d.npz"), **hill)
    base = os.path.join(REF, "Tests/Discretization/Grid_Hierarchy/Gold")
    hier = {"mesh_json": open(os.path.join(base, "askervein_hierarchy_mesh.json")).read()}
    for lev in (1, 2):
        names, lo, hi, cells = read_plotfile_cells(os.path.join(base, f"askervein_hierarchy_level{lev}_plt00000"),
                                                   nvar_expected=2)
        assert names == ["z", "detJ"], names
        hier[f"level{lev}_z"] = cells[0]
        hier[f"level{lev}_detJ"] = cells[1]
    np.savez_compressed(os.path.join(OUT, "askervein_hierarchy.npz"), **hier)
    print("wrote fixtures to", OUT)


if __name__ == "__main__":
    sys.exit(main())
