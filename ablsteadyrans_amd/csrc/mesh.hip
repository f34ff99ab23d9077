// mesh.hip -- a tile's terrain-following mesh built on the device.
// Derived in part from ERF (BSD 3-Clause, Copyright (c) 2022 The Regents of the University of California
// et al.); the full notice is in THIRD_PARTY_NOTICES.md at the repository root.
//
// Reference behaviour (GUST):
//   Source/Mesh.cpp:107-140        build_from_ground (terrain.smoothing = 0):
//                                  BTF levels, OverrideSync, FillBoundary,
//                                  the ghost columns (Source/Mesh.cpp:37-98)
//   extern_erf/ERF_TerrainMetrics.cpp:330-356   BTF node heights
//   extern_erf/ERF_TerrainMetrics.cpp:679-807   make_J, make_areas, make_zcc
//
// Every node height is a function of its column's ground height alone (the
// ghost handling of the reference maps each node column to a source column:
// the periodic image, or the clamped edge column), so one kernel writes all
// nodes and the metrics follow in one pass each, reading the nodes in the
// reference's order of terms. The ground slab (a host array of (nx + 3)
// (ny + 3) node heights) is the only input besides the levels.
#include "cgrid.hpp"
#include "deferred.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace gust_amd {

namespace {

using namespace cgrid;

struct Col {
    int nx, ny, perx, pery;
    __device__ __forceinline__ int src(int i, int n, int per) const
    {
        if (per) { return i < 0 ? i + n : (i >= n ? i - n : i); }
        return i < 0 ? 0 : (i > n ? n : i);
    }
};

// BTF height of level k (0 <= k <= nz) over ground zs
// (ERF_TerrainMetrics.cpp:346-352).
__device__ __forceinline__ double btf(double zs, double zk, double z0, double ztop)
{
    return ((zs - z0) * ztop + (ztop - zs) * zk) / (ztop - z0);
}

__global__ __launch_bounds__(kThreads) void node_kernel(Box3 b, V z, const double* __restrict__ ground,
                                                        const double* __restrict__ zl, Col c, int nz)
{
    const double z0 = zl[0], ztop = zl[nz];
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const int si = c.src(i, c.nx, c.perx), sj = c.src(j, c.ny, c.pery);
        const double zs = ground[int64_t(sj + 1) * (c.nx + 3) + (si + 1)];
        double v;
        if (k == 0) {
            v = zs;
        } else if (k >= 1 && k <= nz) {
            v = btf(zs, zl[k], z0, ztop);
        } else if (k < 0) {
            const double z1 = nz >= 1 ? btf(zs, zl[1], z0, ztop) : zs;
            v = 2.0 * zs - z1;
        } else {
            const double zn = btf(zs, zl[nz], z0, ztop);
            const double zm = nz - 1 >= 1 ? btf(zs, zl[nz - 1], z0, ztop) : zs;
            v = zn + (zn - zm);
        }
        z(i, j, k) = v;
    }
}

// detJ (cells k >= 0; the row below stays 0) and z_cc (every row).
__global__ __launch_bounds__(kThreads) void cell_metrics_kernel(Box3 b, CV z, V J, V zcc, int has_zcc, double dzinv)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        if (k >= 0) {
            const double s = z(i, j, k + 1) + z(i + 1, j, k + 1) + z(i, j + 1, k + 1) + z(i + 1, j + 1, k + 1) -
                             z(i, j, k) - z(i + 1, j, k) - z(i, j + 1, k) - z(i + 1, j + 1, k);
            J(i, j, k) = 0.25 * dzinv * s;
        } else {
            J(i, j, k) = 0.0;
        }
        if (has_zcc) {
            zcc(i, j, k) = 0.125 * (z(i, j, k) + z(i + 1, j, k) + z(i, j + 1, k) + z(i + 1, j + 1, k) + z(i, j, k + 1) +
                                    z(i + 1, j, k + 1) + z(i, j + 1, k + 1) + z(i + 1, j + 1, k + 1));
        }
    }
}

// ax (d = 0) or ay (d = 1): half the face's vertical extent over dzeta
// (make_areas); the row below the ground stays 0.
template <int D>
__global__ __launch_bounds__(kThreads) void area_kernel(Box3 b, CV z, V a, double dzinv)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        if (k < 0) {
            a(i, j, k) = 0.0;
            continue;
        }
        const int i1 = D == 1 ? i + 1 : i, j1 = D == 0 ? j + 1 : j;
        a(i, j, k) = 0.5 * dzinv * (z(i, j, k + 1) + z(i1, j1, k + 1) - z(i, j, k) - z(i1, j1, k));
    }
}

__global__ __launch_bounds__(kThreads) void ones_kernel(Box3 b, V a)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        a(i, j, k) = 1.0;
    }
}

// A flag where a valid node column's layer is not strictly thicker than 0.
__global__ __launch_bounds__(kThreads) void thickness_kernel(Box3 b, CV z, unsigned int* flag)
{
    bool bad = false;
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        if (!(z(i, j, k + 1) - z(i, j, k) > 0)) { bad = true; }
    }
    if (bad) { atomicOr(flag, 1u); }
}

// The shortest decimal that reads back as x, written as Python writes a
// float (positional for decimal exponents -4 .. 15, else scientific).
std::string shortest(double x)
{
    if (!std::isfinite(x)) { return std::isnan(x) ? "nan" : (x > 0 ? "inf" : "-inf"); }
    char buf[64];
    int p = 1;
    for (; p <= 17; ++p) {
        std::snprintf(buf, sizeof buf, "%.*e", p - 1, x);
        if (std::strtod(buf, nullptr) == x) { break; }
    }
    const char* e = std::strchr(buf, 'e');
    const int ex = e ? std::atoi(e + 1) : 0;
    if (ex < -4 || ex >= 16) { return buf; }
    const int dec = p - 1 - ex > 0 ? p - 1 - ex : 0;
    std::snprintf(buf, sizeof buf, "%.*f", dec, x);
    std::string s(buf);
    if (s.find('.') == std::string::npos) { s += ".0"; }
    return s;
}

IBox box(int lo0, int hi0, int lo1, int hi1, int lo2, int hi2)
{
    IBox b;
    b.lo[0] = lo0; b.hi[0] = hi0;
    b.lo[1] = lo1; b.hi[1] = hi1;
    b.lo[2] = lo2; b.hi[2] = hi2;
    return b;
}

}   // namespace

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_build_mesh(const gust_geometry* geom, const double* ground, const double* zlevels, gust_array4* z_nd,
                    gust_array4* detJ, gust_array4* ax, gust_array4* ay, gust_array4* az, gust_array4* zcc,
                    void* stream)
{
    return guarded([&] {
        const char* op = "build_mesh";
        check_geom(op, geom);
        if (!ground || !zlevels || !z_nd || !detJ || !ax || !ay || !az) { GUST_FAIL("gust: %s: null argument", op); }
        const gust_geometry& g = *geom;
        const int nx = g.hi[0] - g.lo[0] + 1, ny = g.hi[1] - g.lo[1] + 1, nz = g.hi[2] - g.lo[2] + 1;
        if (g.lo[0] != 0 || g.lo[1] != 0 || g.lo[2] != 0) { GUST_FAIL("gust: %s: the domain must start at 0", op); }
        for (int k = 0; k < nz; ++k) {
            if (!(zlevels[k + 1] > zlevels[k])) { GUST_FAIL("gust: %s: the levels must increase", op); }
        }
        const double ztop = zlevels[nz];
        // the valid nodes' ground as the mesh holds it (a periodic axis's
        // last node is its first's image)
        double gmax = -INFINITY;
        for (int j = 0; j <= ny; ++j) {
            for (int i = 0; i <= nx; ++i) {
                const int si = g.periodic[0] && i == nx ? 0 : i, sj = g.periodic[1] && j == ny ? 0 : j;
                const double h = ground[int64_t(sj + 1) * (nx + 3) + (si + 1)];
                if (!std::isfinite(h)) { GUST_FAIL("gust: %s: the ground heights must be finite", op); }
                gmax = h > gmax ? h : gmax;
            }
        }
        if (!(gmax < ztop)) {
            GUST_FAIL("gust: the terrain reaches %s m, at or above the top level %s m", shortest(gmax).c_str(),
                      shortest(ztop).c_str());
        }
        const IBox nodes = box(-1, nx + 1, -1, ny + 1, -1, nz + 1);
        const IBox cells = box(-1, nx, -1, ny, -1, nz);
        const IBox xf = box(-1, nx + 1, -1, ny, -1, nz);
        const IBox yf = box(-1, nx, -1, ny + 1, -1, nz);
        const IBox zf = box(-1, nx, -1, ny, -1, nz + 1);
        if (!covers(z_nd, nodes, 1) || !covers(detJ, cells, 1) || !covers(ax, xf, 1) || !covers(ay, yf, 1) ||
            !covers(az, zf, 1) || (zcc && !covers(zcc, cells, 1))) {
            GUST_FAIL("gust: %s: the outputs must cover the mesh's nodes, cells and faces with one ghost layer", op);
        }
        hipStream_t st = static_cast<hipStream_t>(stream);
        // the ground slab and the levels travel once (a build is a set-up step)
        const size_t ng = size_t(nx + 3) * (ny + 3);
        auto* dg = static_cast<double*>(Workspace::get(st, 125, (ng + nz + 1) * sizeof(double)));
        std::vector<double> h(ng + nz + 1);
        std::copy(ground, ground + ng, h.begin());
        std::copy(zlevels, zlevels + nz + 1, h.begin() + ng);
        GUST_HIP(hipMemcpyAsync(dg, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st));
        const Col c{nx, ny, g.periodic[0] ? 1 : 0, g.periodic[1] ? 1 : 0};
        const Box3 bn = box3(nodes), bc = box3(cells), bx = box3(xf), by = box3(yf), bz = box3(zf);
        node_kernel<<<blocks_for(bn.npts()), kThreads, 0, st>>>(bn, dview<double>(z_nd), dg, dg + ng, c, nz);
        const double dzinv = 1.0 / g.dx[2];
        cell_metrics_kernel<<<blocks_for(bc.npts()), kThreads, 0, st>>>(
            bc, dview<const double>(z_nd), dview<double>(detJ), dview<double>(zcc ? zcc : detJ), zcc ? 1 : 0, dzinv);
        area_kernel<0><<<blocks_for(bx.npts()), kThreads, 0, st>>>(bx, dview<const double>(z_nd), dview<double>(ax),
                                                                   dzinv);
        area_kernel<1><<<blocks_for(by.npts()), kThreads, 0, st>>>(by, dview<const double>(z_nd), dview<double>(ay),
                                                                   dzinv);
        ones_kernel<<<blocks_for(bz.npts()), kThreads, 0, st>>>(bz, dview<double>(az));
        GUST_HIP(hipGetLastError());
        DeferredUse dq(st);
        auto* flag = reinterpret_cast<unsigned int*>(dq->zero_slot([](unsigned long long f) {
            if (f) { GUST_FAIL("gust: the terrain-following mesh has inverted or empty cells"); }
        }));
        const Box3 bt = box3(box(0, nx, 0, ny, 0, nz - 1));
        thickness_kernel<<<blocks_for(bt.npts()), kThreads, 0, st>>>(bt, dview<const double>(z_nd), flag);
        GUST_HIP(hipGetLastError());
        // the host copy of the slab must outlive its upload
        GUST_HIP(hipStreamSynchronize(st));
        dq.finish();
    });
}

}   // extern "C"
