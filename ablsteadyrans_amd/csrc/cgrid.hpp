// cgrid.hpp -- device stencils of the terrain-following C-grid shared by the
// mass-flux, pseudo-time and SIMPLEC kernels: a z-face's slopes from its four
// nodes (Source/MassFlux_K.H:19-39), the means of u, v around a z-face and
// of a cell's faces, and Omega (Source/MassFlux.cpp:51-74). Each sums from 0
// in the reference's loop order (Source/Stencil_K.H:49-99) and then scales,
// so FP64 values equal the CPU build's.
#pragma once

#include "common.hpp"

#include <algorithm>

namespace gust_amd::cgrid {

constexpr int kThreads = 256;

inline int blocks_for(int64_t n)
{
    return int(std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, 16384)));
}

// The grid of a reduction kernel: grid-stride over the box, one atomic per
// block (2 blocks per CU).
inline int reduce_blocks(int64_t n)
{
    return int(std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, 512)));
}

struct Box3 {
    int lo0, lo1, lo2, n0, n1, n2;
    __host__ __device__ int64_t npts() const { return int64_t(n0) * n1 * n2; }
    __device__ __forceinline__ void at(int64_t t, int& i, int& j, int& k) const
    {
        i = lo0 + int(t % n0);
        const int64_t r = t / n0;
        j = lo1 + int(r % n1);
        k = lo2 + int(r / n1);
    }
};

inline Box3 box3(const IBox& b)
{
    return Box3{b.lo[0], b.lo[1], b.lo[2], b.len(0), b.len(1), b.len(2)};
}

inline void check_geom(const char* op, const gust_geometry* g)
{
    if (!g) { GUST_FAIL("gust: %s: null geometry", op); }
    for (int d = 0; d < 3; ++d) {
        if (g->hi[d] < g->lo[d] || !(g->dx[d] > 0)) { GUST_FAIL("gust: %s: invalid geometry", op); }
    }
}

using V = DView<double>;
using CV = DView<const double>;

__device__ __forceinline__ double slope_x(const CV& z, int i, int j, int k, double dxinv)
{
    double s = 0.0;
    s += z(i + 1, j, k) - z(i, j, k);
    s += z(i + 1, j + 1, k) - z(i, j + 1, k);
    return (1.0 * (s * 0.5)) * dxinv;
}

__device__ __forceinline__ double slope_y(const CV& z, int i, int j, int k, double dyinv)
{
    double s = 0.0;
    s += z(i, j + 1, k) - z(i, j, k);
    s += z(i + 1, j + 1, k) - z(i + 1, j, k);
    return (1.0 * (s * 0.5)) * dyinv;
}

__device__ __forceinline__ double u_at_zface(const CV& u, int i, int j, int k)
{
    double s = 0.0;
    s += u(i, j, k - 1);
    s += u(i + 1, j, k - 1);
    s += u(i, j, k);
    s += u(i + 1, j, k);
    return s * 0.25;
}

__device__ __forceinline__ double v_at_zface(const CV& v, int i, int j, int k)
{
    double s = 0.0;
    s += v(i, j, k - 1);
    s += v(i, j + 1, k - 1);
    s += v(i, j, k);
    s += v(i, j + 1, k);
    return s * 0.25;
}

__device__ __forceinline__ double omega_at(const CV& u, const CV& v, const CV& w, const CV& z, int i, int j, int k,
                                           double dxinv, double dyinv)
{
    return w(i, j, k) - slope_x(z, i, j, k, dxinv) * u_at_zface(u, i, j, k) -
           slope_y(z, i, j, k, dyinv) * v_at_zface(v, i, j, k);
}

// interp<CC, Face(d)>: the mean of the two cells beside a d-face.
__device__ __forceinline__ double J_at_face(const CV& J, int i, int j, int k, int d)
{
    double s = 0.0;
    s += J(i - (d == 0), j - (d == 1), k - (d == 2));
    s += J(i, j, k);
    return s * 0.5;
}

// interp<Face(d), CC>: the mean of a cell's two d-faces.
template <class A>
__device__ __forceinline__ double cell_mean(const A& f, int i, int j, int k, int d)
{
    double s = 0.0;
    s += f(i, j, k);
    s += f(i + (d == 0), j + (d == 1), k + (d == 2));
    return s * 0.5;
}

#define GUST_GRID_LOOP(b)                                                                                        \
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < (b).npts();                            \
         t += int64_t(gridDim.x) * blockDim.x)

}   // namespace gust_amd::cgrid
