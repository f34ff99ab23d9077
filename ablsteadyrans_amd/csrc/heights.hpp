// heights.hpp -- heights of the points of the terrain-following mesh from
// its nodes, shared by the boundary fills and the initializer
// (Source/Heights_K.H:20-68), and the sounding theta_ref(z) by absolute
// height (Source/Forcing.cpp:84-125). Sums run from 0 in the reference's
// loop order and scale last, so FP64 values are its bits.
#pragma once

#include "cgrid.hpp"
#include "stencil.hpp"

namespace gust_amd::heights {

// Node heights with the lateral indices clamped to the stored nodes and the
// levels extrapolated linearly beyond them (Heights_K.H:20-33).
struct NodeZ {
    cgrid::CV z;
    int lo0, lo1, lo2, hi0, hi1, hi2;
    __device__ __forceinline__ double operator()(int i, int j, int k) const
    {
        i = i < lo0 ? lo0 : (i > hi0 ? hi0 : i);
        j = j < lo1 ? lo1 : (j > hi1 ? hi1 : j);
        if (k < lo2) { return z(i, j, lo2) + (lo2 - k) * (z(i, j, lo2) - z(i, j, lo2 + 1)); }
        if (k > hi2) { return z(i, j, hi2) + (k - hi2) * (z(i, j, hi2) - z(i, j, hi2 - 1)); }
        return z(i, j, k);
    }
};

inline NodeZ node_z(const gust_array4* z)
{
    return NodeZ{dview<const double>(z), z->begin[0], z->begin[1], z->begin[2], z->end[0] - 1, z->end[1] - 1,
                 z->end[2] - 1};
}

// The node-mean height of a point of type T (Heights_K.H:55-68).
template <int T>
__device__ __forceinline__ double point_z(const NodeZ& z, int i, int j, int k)
{
    namespace sc = stencil_cg;
    constexpr int nx = sc::nodal(T, 0) ? 1 : 2, ny = sc::nodal(T, 1) ? 1 : 2, nz = sc::nodal(T, 2) ? 1 : 2;
    double zp = 0;
#pragma unroll
    for (int a = 0; a < nx; ++a)
#pragma unroll
        for (int b = 0; b < ny; ++b)
#pragma unroll
            for (int c = 0; c < nz; ++c) { zp += z(i + a, j + b, k + c); }
    return zp / (nx * ny * nz);
}

// The point's node-mean height minus the node-mean ground under its
// footprint, the ground layer klo (Heights_K.H:35-53).
template <int T>
__device__ __forceinline__ double height_above_ground(const NodeZ& z, int i, int j, int k, int klo)
{
    namespace sc = stencil_cg;
    constexpr int nx = sc::nodal(T, 0) ? 1 : 2, ny = sc::nodal(T, 1) ? 1 : 2, nz = sc::nodal(T, 2) ? 1 : 2;
    double zp = 0, zg = 0;
#pragma unroll
    for (int a = 0; a < nx; ++a)
#pragma unroll
        for (int b = 0; b < ny; ++b) {
            zg += z(i + a, j + b, klo);
#pragma unroll
            for (int c = 0; c < nz; ++c) { zp += z(i + a, j + b, k + c); }
        }
    return zp / (nx * ny * nz) - zg / (nx * ny);
}

// The sounding: up to 64 (height, theta) points, heights increasing.
struct Sounding {
    int n;
    double z[64], t[64];
};

inline Sounding sounding(int n, const double* z, const double* t)
{
    Sounding s{};
    s.n = n;
    for (int m = 0; m < 64; ++m) {
        s.z[m] = m < n ? z[m] : 0.0;
        s.t[m] = m < n ? t[m] : 0.0;
    }
    return s;
}

// theta_ref at the absolute height zc: linear between the points, the end
// values beyond them (Forcing.cpp:104-120).
__device__ __forceinline__ double sounding_at(const Sounding& p, double zc)
{
    double v = p.t[0];
    if (zc >= p.z[p.n - 1]) {
        v = p.t[p.n - 1];
    } else {
        for (int m = 1; m < p.n; ++m) {
            if (zc < p.z[m]) {
                if (zc > p.z[m - 1]) {
                    const double wgt = (zc - p.z[m - 1]) / (p.z[m] - p.z[m - 1]);
                    v = p.t[m - 1] + wgt * (p.t[m] - p.t[m - 1]);
                } else {
                    v = p.t[m - 1];
                }
                break;
            }
        }
    }
    return v;
}

namespace {

// theta_ref on the cells of b from the node heights (Forcing.cpp:98-123);
// one copy per translation unit.
__global__ __launch_bounds__(cgrid::kThreads) void theta_ref_kernel(cgrid::Box3 b, cgrid::CV z, Sounding p,
                                                                    cgrid::V tr)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        tr(i, j, k) = sounding_at(p, stencil_cg::interp<stencil_cg::NODE, stencil_cg::CC>(z, i, j, k));
    }
}

}   // namespace

}   // namespace gust_amd::heights
