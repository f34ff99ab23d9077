// stencil.hpp -- the C-grid's staggered interpolation and index differences
// (Source/Stencil_K.H:23-140) as compile-time templates over location types
// (bit d set: nodal along d). Loop order (z outer, then y, then x),
// accumulation from 0 and the final 1/n scale follow the reference, so FP64
// values are its bits.
#pragma once

#include "common.hpp"

namespace gust_amd::stencil_cg {

constexpr int CC = 0, XF = 1, YF = 2, ZF = 4, NODE = 7, NONE = 3;

__host__ __device__ constexpr bool nodal(int t, int d) { return ((t >> d) & 1) != 0; }
__host__ __device__ constexpr int face(int c) { return 1 << c; }
__host__ __device__ constexpr int lo_off(int tq, int ts, int d)
{
    return (nodal(tq, d) == nodal(ts, d) || nodal(tq, d)) ? 0 : -1;
}
__host__ __device__ constexpr int npts(int tq, int ts, int d) { return (nodal(tq, d) == nodal(ts, d)) ? 1 : 2; }
__host__ __device__ constexpr int flux_lo(int ts, int d) { return nodal(ts, d) ? -1 : 0; }

template <int TQ, int TS, class A>
__device__ __forceinline__ double interp(const A& q, int i, int j, int k)
{
    constexpr int lx = lo_off(TQ, TS, 0), ly = lo_off(TQ, TS, 1), lz = lo_off(TQ, TS, 2);
    constexpr int nx = npts(TQ, TS, 0), ny = npts(TQ, TS, 1), nz = npts(TQ, TS, 2);
    double s = 0;
#pragma unroll
    for (int c = 0; c < nz; ++c)
#pragma unroll
        for (int b = 0; b < ny; ++b)
#pragma unroll
            for (int a = 0; a < nx; ++a) { s += q(i + lx + a, j + ly + b, k + lz + c); }
    return s * (1.0 / double(nx * ny * nz));
}

template <int TQ, int TS, int M, class A>
__device__ __forceinline__ double deriv(const A& q, int i, int j, int k)
{
    constexpr bool same = nodal(TQ, M) == nodal(TS, M);
    constexpr int m0 = same ? -1 : lo_off(TQ, TS, M);
    constexpr int m1 = same ? 1 : m0 + 1;
    constexpr double w = same ? 0.5 : 1.0;
    constexpr int ex = M == 0, ey = M == 1, ez = M == 2;
    constexpr int lx = M == 0 ? 0 : lo_off(TQ, TS, 0), ly = M == 1 ? 0 : lo_off(TQ, TS, 1),
                  lz = M == 2 ? 0 : lo_off(TQ, TS, 2);
    constexpr int nx = M == 0 ? 1 : npts(TQ, TS, 0), ny = M == 1 ? 1 : npts(TQ, TS, 1),
                  nz = M == 2 ? 1 : npts(TQ, TS, 2);
    double s = 0;
#pragma unroll
    for (int c = 0; c < nz; ++c)
#pragma unroll
        for (int b = 0; b < ny; ++b)
#pragma unroll
            for (int a = 0; a < nx; ++a) {
                const int ii = i + lx + a, jj = j + ly + b, kk = k + lz + c;
                s += q(ii + m1 * ex, jj + m1 * ey, kk + m1 * ez) - q(ii + m0 * ex, jj + m0 * ey, kk + m0 * ez);
            }
    return w * (s * (1.0 / double(nx * ny * nz)));
}

struct Metric {
    double zx, zy, zz;
};

template <int TS, class A>
__device__ __forceinline__ Metric metric_at(const A& z, int i, int j, int k, const double (&dxinv)[3])
{
    return Metric{deriv<NODE, TS, 0>(z, i, j, k) * dxinv[0], deriv<NODE, TS, 1>(z, i, j, k) * dxinv[1],
                  deriv<NODE, TS, 2>(z, i, j, k) * dxinv[2]};
}

// std::max / std::min as amrex::max / amrex::min.
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }

}   // namespace gust_amd::stencil_cg
