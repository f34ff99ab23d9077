// poisson.hip -- the terrain-following pressure-correction solve for gfx950.
// Derived in part from ERF (BSD 3-Clause, Copyright (c) 2022 The Regents of the University of California
// et al.); the full notice is in THIRD_PARTY_NOTICES.md at the repository root.
//
// Reference behaviour (GUST):
//   Source/Poisson.H:119-203, Source/Poisson.cpp:395-731   PoissonSolver
//   extern_erf/ERF_TerrainPoisson_3D_K.H                  the operator
//   extern_erf/ERF_TerrainPoisson.cpp:74-264               apply, BCs, fluxes
//   Source/PoissonCross_K.H                                split z coefficient
//   Source/Poisson.cpp:733-764                             face_divergence
//
// Design (MI355X):
//  * Fields are compact (nx*ny*nz, i fastest); the boundary conditions are
//    applied when a stencil reads one cell outside the domain (periodic
//    image, even or odd mirror), so no ghost-fill pass runs between the
//    smoother's half sweeps. The values read are exactly the ghost values of
//    fill_bc_ghosts, so FP64 operator applies equal the CPU build's bits.
//  * Solver: restarted GMRES in FP64 (deterministic two-stage reductions,
//    classical Gram-Schmidt with one re-orthogonalisation), right-
//    preconditioned by one multigrid V-cycle. The multigrid coarsens in x and
//    y only (columns are never split), re-discretises the same terrain
//    operator on injected node heights with area-averaged coefficients, and
//    smooths with red-black ("zebra") z-line relaxation: each column's exact
//    tridiagonal part, taken from the operator itself by evaluating the
//    stencil on values linear in the column (terrain_stencil.hpp), is solved
//    by Thomas. The V-cycle runs in FP64 or, for the mixed-precision path,
//    in FP32 while GMRES's Arnoldi vectors, the operator and the residuals
//    stay FP64.
//  * The AMReX MLMG/GMRES the reference calls are third party; only the
//    discrete operator and the converged answer are the reference's, so the
//    iteration counts are this solver's own.
#include "common.hpp"
#include "deferred.hpp"
#include "terrain_stencil.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

namespace gust_amd {

namespace {

using stencil::Lin3;
using stencil::OpCtx;

constexpr int kThreads = 256;
constexpr int kRedBlocks = 1024;

// ---------------------------------------------------------------------------
// Accessors.

template <typename R>
struct NodeAcc {   // node heights, one ghost node: index (i+1, j+1, k+1)
    const R* p;
    int64_t sj, sk;
    __device__ __forceinline__ R operator()(int i, int j, int k) const
    {
        return p[(i + 1) + (j + 1) * sj + int64_t(k + 1) * sk];
    }
};

template <typename R>
struct GridAcc {   // compact array of valid points
    const R* p;
    int64_t sj, sk;
    __device__ __forceinline__ R operator()(int i, int j, int k) const { return p[i + j * sj + int64_t(k) * sk]; }
};

// A cell field read with the boundary conditions one layer outside the
// domain: periodic images, even (Neumann) or odd (Dirichlet) mirrors, x
// then y then z (Source/Poisson.cpp:84-125).
template <typename R>
struct MirrorAcc {
    const R* p;
    int nx, ny, nz;
    int64_t sj, sk;
    int perx, pery;
    R sxl, sxh, syl, syh, szl, szh;
    __device__ __forceinline__ void resolve(int& i, int& j, int& k, R& s) const
    {
        s = R(1);
        if (i < 0) {
            if (perx) { i += nx; } else { i = 0; s = s * sxl; }
        } else if (i >= nx) {
            if (perx) { i -= nx; } else { i = nx - 1; s = s * sxh; }
        }
        if (j < 0) {
            if (pery) { j += ny; } else { j = 0; s = s * syl; }
        } else if (j >= ny) {
            if (pery) { j -= ny; } else { j = ny - 1; s = s * syh; }
        }
        if (k < 0) { k = 0; s = s * szl; }
        else if (k >= nz) { k = nz - 1; s = s * szh; }
    }
    __device__ __forceinline__ R operator()(int i, int j, int k) const
    {
        if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) { return p[i + j * sj + int64_t(k) * sk]; }
        R s;
        resolve(i, j, k, s);
        return s * p[i + j * sj + int64_t(k) * sk];
    }
};

// The 15 points of the operator's footprint around (ci, cj, ck) held in
// registers: columns own, x-, x+, y-, y+ at levels ck-1, ck, ck+1. The
// stencil reads it with offsets known at compile time once inlined, so
// every read is a register.
template <typename R>
struct Patch {
    R v[5][3];
    int ci, cj, ck;
    __device__ __forceinline__ R operator()(int i, int j, int k) const
    {
        const int di = i - ci, dj = j - cj;
        const int col = di == -1 ? 1 : di == 1 ? 2 : dj == -1 ? 3 : dj == 1 ? 4 : 0;
        return v[col][k - ck + 1];
    }
};

__device__ __forceinline__ void patch_point(int col, int d, int i, int j, int k, int& a, int& b, int& c)
{
    a = i + (col == 1 ? -1 : col == 2 ? 1 : 0);
    b = j + (col == 3 ? -1 : col == 4 ? 1 : 0);
    c = k + d;
}

template <typename R, class X>
__device__ __forceinline__ Patch<R> load_patch(const X& x, int i, int j, int k)
{
    Patch<R> p;
    p.ci = i;
    p.cj = j;
    p.ck = k;
#pragma unroll
    for (int col = 0; col < 5; ++col)
#pragma unroll
        for (int d = -1; d <= 1; ++d) {
            int a, b, c;
            patch_point(col, d, i, j, k, a, b, c);
            p.v[col][d + 1] = x(a, b, c);
        }
    return p;
}

template <typename R>
using Ctx = OpCtx<R, NodeAcc<R>, GridAcc<R>, stencil::ArrayMetric<R>>;

// Split ("red-black") layout of a multigrid level: the cells of colour
// c = (i + j) % 2 are stored apart, each colour as (H, ny, nz) with
// H = ceil(nx / 2), so a half sweep of one colour reads and writes
// contiguous memory. Cell (i, j, k) is slot c * plane + i / 2 + H * (j + ny * k);
// when nx is odd, one slot per row of one colour is padding.
struct Split {
    int nx, ny, nz, H;
    int64_t plane;   // H * ny * nz
    __host__ __device__ __forceinline__ int64_t at(int i, int j, int k) const
    {
        return int64_t((i + j) & 1) * plane + (i >> 1) + int64_t(H) * (j + int64_t(ny) * k);
    }
    // The cell of slot t (t < 2 * plane); false for a padding slot.
    __host__ __device__ __forceinline__ bool cell(int64_t t, int& i, int& j, int& k) const
    {
        const int c = t >= plane ? 1 : 0;
        const int64_t r = t - c * plane;
        const int ih = int(r % H);
        const int64_t q = r / H;
        j = int(q % ny);
        k = int(q / ny);
        i = 2 * ih + ((j + c) & 1);
        return i < nx;
    }
};

// A split-layout field read with the boundary conditions (as MirrorAcc).
template <typename R>
struct SplitMirror {
    MirrorAcc<R> m;   // resolve only
    Split s;
    const R* p;
    __device__ __forceinline__ R operator()(int i, int j, int k) const
    {
        R sg;
        m.resolve(i, j, k, sg);
        return sg * p[s.at(i, j, k)];
    }
};

// The horizontal neighbour columns (x-, x+, y-, y+) of column (i, j) as the
// operator couples them: periodic images wrapped; a column outside a
// non-periodic side is absent (its mirror folds into the own column); a
// column met twice (a periodic direction of one or two cells) is kept once,
// at its first appearance, the own column counting first.
__device__ __forceinline__ void nb_columns(int i, int j, int nx, int ny, int perx, int pery, int (&ni)[4],
                                           int (&nj)[4], bool (&ok)[4])
{
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        int a = i + (p == 0 ? -1 : p == 1 ? 1 : 0);
        int b = j + (p == 2 ? -1 : p == 3 ? 1 : 0);
        bool v = true;
        if (a < 0) { if (perx) { a += nx; } else { v = false; a = i; } }
        else if (a >= nx) { if (perx) { a -= nx; } else { v = false; a = i; } }
        if (b < 0) { if (pery) { b += ny; } else { v = false; b = j; } }
        else if (b >= ny) { if (pery) { b -= ny; } else { v = false; b = j; } }
        if (a == i && b == j) { v = false; }
#pragma unroll
        for (int q = 0; q < p; ++q) {
            if (ok[q] && ni[q] == a && nj[q] == b) { v = false; }
        }
        ni[p] = a;
        nj[p] = b;
        ok[p] = v;
    }
}

// A column on a seam that joins two columns of one colour: the first or the
// last of a periodic direction of odd length above one.
__host__ __device__ __forceinline__ bool on_seam(int i, int j, int nx, int ny, int perx, int pery)
{
    return (perx && nx > 1 && (nx & 1) && (i == 0 || i == nx - 1)) ||
           (pery && ny > 1 && (ny & 1) && (j == 0 || j == ny - 1));
}

// A level's explicit stencil, one bundle of kBundle reals per split slot:
// the own column's lower entry, 1 / Thomas pivot, (k-1, k, k+1) of the x-,
// x+, y-, y+ neighbour columns, the own diagonal and upper entries. The
// Thomas c' has a plane of its own (read by the back substitution alone).
enum : int { kBLo = 0, kBMinv = 1, kBNb = 2, kBDg = 14, kBUp = 15, kBundle = 16 };

// ---------------------------------------------------------------------------
// Kernels.

__device__ __forceinline__ void cell_of(int64_t t, int nx, int ny, int& i, int& j, int& k)
{
    i = int(t % nx);
    const int64_t r = t / nx;
    j = int(r % ny);
    k = int(r / ny);
}

// y = A x (mode 0) or y = b - A x (mode 1) on the operator level.
template <typename R, int kMode>
__global__ __launch_bounds__(kThreads) void apply_kernel(Ctx<R> c, MirrorAcc<R> x, const R* __restrict__ b,
                                                          R* __restrict__ y)
{
    const int64_t n = int64_t(c.nx) * c.ny * c.nz;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        int i, j, k;
        cell_of(t, c.nx, c.ny, i, j, k);
        const Patch<R> px = load_patch<R>(x, i, j, k);
        const R v = stencil::apply_cell<R, R>(px, c, i, j, k);
        y[t] = kMode == 0 ? v : b[t] - v;
    }
}

// The mesh factors of a level (stencil::ArrayMetric) from its node heights.
template <typename R>
__global__ void metric_kernel(stencil::ZMetric<R, NodeAcc<R>> zm, int nx, int ny, int nz, R* __restrict__ ry,
                              R* __restrict__ rx, R* __restrict__ zf)
{
    const int64_t n_ry = int64_t(nx + 1) * ny * (nz + 1), n_rx = int64_t(nx) * (ny + 1) * (nz + 1),
                  n_zf = int64_t(nx) * ny * (nz + 1);
    const int64_t n = n_ry + n_rx + n_zf;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        int i, j, k;
        if (t < n_ry) {
            cell_of(t, nx + 1, ny, i, j, k);
            ry[t] = zm.ry(i, j, k);
        } else if (t < n_ry + n_rx) {
            const int64_t u = t - n_ry;
            cell_of(u, nx, ny + 1, i, j, k);
            rx[u] = zm.rx(i, j, k);
        } else {
            const int64_t u = t - n_ry - n_rx;
            cell_of(u, nx, ny, i, j, k);
            zf[u] = zm.hzi(i, j, k);
            zf[u + n_zf] = zm.hxz(i, j, k);
            zf[u + 2 * n_zf] = zm.hez(i, j, k);
        }
    }
}

// The mesh factors and face coefficients one cell's operator reads, loaded
// once into registers: the 15 probes of coef_seed_kernel evaluate the same
// expressions on the same values (hence the same bits) without reloading
// them. Reads use offsets known at compile time once inlined.
template <typename R>
struct MetricCache {
    R ryv[2][2];     // ry(i + a, j, k + c)
    R rxv[2][2];     // rx(i, j + b, k + c)
    R zfv[3][2];     // hzi, hxz, hez at (i, j, k + c)
    int ci, cj, ck;
    __device__ __forceinline__ R ry(int i, int, int k) const { return ryv[i - ci][k - ck]; }
    __device__ __forceinline__ R rx(int, int j, int k) const { return rxv[j - cj][k - ck]; }
    __device__ __forceinline__ R hzi(int, int, int k) const { return zfv[0][k - ck]; }
    __device__ __forceinline__ R hxz(int, int, int k) const { return zfv[1][k - ck]; }
    __device__ __forceinline__ R hez(int, int, int k) const { return zfv[2][k - ck]; }
};

// A face or cell array read at the cell's own point and one step up in a
// single direction dir (0: i + 1, 1: j + 1, 2: k + 1).
template <typename R>
struct FaceCache {
    const R* p;   // non-null when the array is present (z_cross)
    R v0, v1;
    int ci, cj, ck;
    __device__ __forceinline__ R operator()(int i, int j, int k) const
    {
        return (i - ci) + (j - cj) + (k - ck) == 0 ? v0 : v1;
    }
};

template <typename R>
__device__ __forceinline__ FaceCache<R> face_cache(const GridAcc<R>& a, int dir, bool two, int i, int j, int k)
{
    FaceCache<R> f;
    f.p = a.p;
    f.ci = i;
    f.cj = j;
    f.ck = k;
    f.v0 = R(0);
    f.v1 = R(0);
    if (a.p) {
        f.v0 = a(i, j, k);
        if (two) { f.v1 = a(i + (dir == 0), j + (dir == 1), k + (dir == 2)); }
    }
    return f;
}

template <typename R>
using CachedCtx = OpCtx<R, NodeAcc<R>, FaceCache<R>, MetricCache<R>>;

template <typename R>
__device__ __forceinline__ CachedCtx<R> cache_ctx(const Ctx<R>& c, int i, int j, int k)
{
    CachedCtx<R> q;
    q.zp = c.zp;
    q.mt.ci = i;
    q.mt.cj = j;
    q.mt.ck = k;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            q.mt.ryv[a][d] = c.mt.ry(i + a, j, k + d);
            q.mt.rxv[a][d] = c.mt.rx(i, j + a, k + d);
        }
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        q.mt.zfv[0][d] = c.mt.hzi(i, j, k + d);
        q.mt.zfv[1][d] = c.mt.hxz(i, j, k + d);
        q.mt.zfv[2][d] = c.mt.hez(i, j, k + d);
    }
    q.ax = face_cache(c.ax, 0, true, i, j, k);
    q.ay = face_cache(c.ay, 1, true, i, j, k);
    q.az = face_cache(c.az, 2, true, i, j, k);
    q.J = face_cache(c.J, 2, false, i, j, k);
    q.dcross = face_cache(c.dcross, 2, true, i, j, k);
    q.az_mesh = face_cache(c.az_mesh, 2, true, i, j, k);
    q.dxinv = c.dxinv;
    q.dyinv = c.dyinv;
    q.dzinv = c.dzinv;
    q.nx = c.nx;
    q.ny = c.ny;
    q.nz = c.nz;
    q.zf = c.zf;
    q.dir_ground = c.dir_ground;
    q.top_open = c.top_open;
    return q;
}

// The explicit 15-point stencil of a level: each coefficient is the
// operator applied to a probe that is 1 at one point (column, level) and 0
// elsewhere, with the boundary mirrors resolved exactly as the operator
// reads them (a mirrored point carries its sign to the point it images).
// Writes the bundle's lo, dg, up and neighbour entries.
template <typename R>
__global__ __launch_bounds__(kThreads) void coef_seed_kernel(Ctx<R> c, MirrorAcc<R> m, Split s, int perx, int pery,
                                                             R* __restrict__ bundle)
{
    const int64_t n = 2 * s.plane;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        int i, j, k;
        if (!s.cell(t, i, j, k)) { continue; }
        int ni[4], nj[4];
        bool ok[4];
        nb_columns(i, j, c.nx, c.ny, perx, pery, ni, nj, ok);
        const CachedCtx<R> cc = cache_ctx(c, i, j, k);
        // the resolved footprint
        int ri[5][3], rj[5][3], rk[5][3];
        R sg[5][3];
#pragma unroll
        for (int col = 0; col < 5; ++col)
#pragma unroll
            for (int d = -1; d <= 1; ++d) {
                int a, b, cc;
                patch_point(col, d, i, j, k, a, b, cc);
                R sv;
                m.resolve(a, b, cc, sv);
                ri[col][d + 1] = a;
                rj[col][d + 1] = b;
                rk[col][d + 1] = cc;
                sg[col][d + 1] = sv;
            }
#pragma unroll 1
        for (int tp = 0; tp < 15; ++tp) {
            const int p = tp / 3, d = tp % 3 - 1;
            const int ti = p == 0 ? i : ni[p - 1], tj = p == 0 ? j : nj[p - 1];
            R v = R(0);
            if (p == 0 || ok[p - 1]) {
                Patch<R> pt;
                pt.ci = i;
                pt.cj = j;
                pt.ck = k;
#pragma unroll
                for (int col = 0; col < 5; ++col)
#pragma unroll
                    for (int e = 0; e < 3; ++e) {
                        const bool hit = ri[col][e] == ti && rj[col][e] == tj && rk[col][e] == k + d;
                        pt.v[col][e] = hit ? sg[col][e] : R(0);
                    }
                v = stencil::apply_cell<R, R>(pt, cc, i, j, k);
            }
            const int e = p == 0 ? (d == -1 ? kBLo : d == 0 ? kBDg : kBUp) : kBNb + 3 * (p - 1) + d + 1;
            bundle[t * kBundle + e] = v;
        }
    }
}

// Thomas factors of every column: 1 / pivot (into the bundle) and c' (its
// own plane), the part of the elimination that does not depend on the
// right-hand side.
template <typename R>
__global__ void coef_factor_kernel(Split s, R* __restrict__ bundle, R* __restrict__ cpl)
{
    const int64_t rowc = int64_t(s.H) * s.ny, ncol = 2 * rowc;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < ncol; t += int64_t(gridDim.x) * blockDim.x) {
        const int c = t >= rowc ? 1 : 0;
        const int64_t r = t - c * rowc;
        const int ih = int(r % s.H), j = int(r / s.H);
        if (2 * ih + ((j + c) & 1) >= s.nx) { continue; }
        const int64_t q0 = c * s.plane + r;
        R cprev = R(0);
        for (int k = 0; k < s.nz; ++k) {
            const int64_t q = q0 + k * rowc;
            R* bq = bundle + q * kBundle;
            const R lo = bq[kBLo], dg = bq[kBDg], up = bq[kBUp];
            const R minv = R(1) / (k == 0 ? dg : dg - lo * cprev);
            const R cp = up * minv;
            bq[kBMinv] = minv;
            cpl[q] = cp;
            cprev = cp;
        }
    }
}

// A cell's bundle in 16-byte loads.
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void load_bundle(float (&c)[kBundle], const float* p)
{
    const v4f* v = reinterpret_cast<const v4f*>(p);
#pragma unroll
    for (int m = 0; m < kBundle / 4; ++m) {
        const v4f w = dev::gld(v + m);
        c[4 * m] = w.x;
        c[4 * m + 1] = w.y;
        c[4 * m + 2] = w.z;
        c[4 * m + 3] = w.w;
    }
}

__device__ __forceinline__ void load_bundle(double (&c)[kBundle], const double* p)
{
    const v2d* v = reinterpret_cast<const v2d*>(p);
#pragma unroll
    for (int m = 0; m < kBundle / 2; ++m) {
        const v2d w = dev::gld(v + m);
        c[2 * m] = w.x;
        c[2 * m + 1] = w.y;
    }
}

// The neighbour-column terms of a cell: sum over the four columns of the
// bundle's (k-1, k, k+1) entries times x there.
template <typename R>
__device__ __forceinline__ R nb_terms(const R (&c)[kBundle], const R (&xm)[4], const R (&x0)[4], const R (&xp)[4])
{
    R v = R(0);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        v = v + (c[kBNb + 3 * p] * xm[p] + c[kBNb + 3 * p + 1] * x0[p] + c[kBNb + 3 * p + 2] * xp[p]);
    }
    return v;
}

// The operands of one level of a column's forward elimination.
template <typename R>
struct ZLevel {
    R c[kBundle];
    R bb;
    R xn[4];   // the neighbour columns at the next level up
};

template <typename R, bool kNb>
__device__ __forceinline__ void zload(ZLevel<R>& o, const R* __restrict__ bundle, int64_t q,
                                      const R* __restrict__ rhs, const R* x, const int64_t (&nb)[4],
                                      const bool (&ok)[4], int64_t xk)
{
    using dev::gld;
    load_bundle(o.c, bundle + q * kBundle);
    o.bb = gld(rhs + q);
    if (kNb) {
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const R v = gld(x + nb[p] + xk);
            o.xn[p] = ok[p] ? v : R(0);
        }
    }
}

// One colour of a zebra sweep as one pass per column: the right-hand side
// b - (neighbour columns) x, the forward elimination with the stored
// factors, the back substitution into x. The neighbours of a column all
// have the other colour, so the columns of one launch are independent.
//   kZero:  x is zero on entry (the first half sweep): no neighbour terms.
//   kFromR: the right-hand side was formed beforehand (rhs = r), for levels
//           where a periodic seam joins two columns of one colour.
//   kLds:   e' stays in LDS (64 x nz reals); else in eg (split layout).
//   kP:     levels of operands prefetched ahead (a register ring indexed at
//           compile time); kP x (loads per level) stays under the 63
//           vector memory operations a wave can have outstanding.
// One wave per 64 columns of the colour; each lane walks its column.
template <typename R, bool kZero, bool kFromR, bool kLds, int kP>
__global__ __launch_bounds__(64) void zebra_kernel(Split s, int color, int perx, int pery,
                                                   const R* __restrict__ bundle, const R* __restrict__ cpl,
                                                   const R* __restrict__ rhs, R* __restrict__ x, R* __restrict__ eg)
{
    using dev::gld;
    using dev::gst;
    extern __shared__ unsigned char lds_raw[];
    R* el = reinterpret_cast<R*>(lds_raw);
    constexpr bool kNb = !kZero && !kFromR;
    const int lane = threadIdx.x;
    const int64_t rowc = int64_t(s.H) * s.ny;
    const int64_t t = int64_t(dev::xcd_remap(blockIdx.x, gridDim.x)) * 64 + lane;
    if (t >= rowc) { return; }
    const int ih = int(t % s.H), j = int(t / s.H);
    const int i = 2 * ih + ((j + color) & 1);
    if (i >= s.nx) { return; }
    const int nz = s.nz;
    const int64_t q0 = color * s.plane + t;
    int64_t nb[4] = {0, 0, 0, 0};
    bool ok[4] = {false, false, false, false};
    R xm1[4] = {R(0), R(0), R(0), R(0)}, x0[4] = {R(0), R(0), R(0), R(0)};
    if (kNb) {
        int ni[4], nj[4];
        nb_columns(i, j, s.nx, s.ny, perx, pery, ni, nj, ok);
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            nb[p] = s.at(ni[p], nj[p], 0);
            const R v = gld(x + nb[p]);
            x0[p] = ok[p] ? v : R(0);
        }
    }
    ZLevel<R> ring[kP];
#pragma unroll
    for (int u = 0; u < kP; ++u) {
        const int k = u < nz ? u : nz - 1;
        const int kx = k + 1 < nz ? k + 1 : nz - 1;
        zload<R, kNb>(ring[u], bundle, q0 + k * rowc, rhs, x, nb, ok, kx * rowc);
    }
    R eprev = R(0);
    auto eliminate = [&](const ZLevel<R>& o, int k) {
        R r = o.bb;
        if (kNb) {
            r = r - nb_terms(o.c, xm1, x0, o.xn);
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                xm1[p] = x0[p];
                x0[p] = o.xn[p];
            }
        }
        const R e = (r - o.c[kBLo] * eprev) * o.c[kBMinv];
        if (kLds) { el[k * 64 + lane] = e; } else { gst(eg + q0 + k * rowc, e); }
        eprev = e;
    };
    const int nfull = (nz / kP) * kP;
    for (int k0 = 0; k0 < nfull; k0 += kP) {
#pragma unroll
        for (int u = 0; u < kP; ++u) {
            const int k = k0 + u;
            const ZLevel<R> o = ring[u];
            const int kn = k + kP < nz ? k + kP : nz - 1;
            const int kx = kn + 1 < nz ? kn + 1 : nz - 1;
            zload<R, kNb>(ring[u], bundle, q0 + kn * rowc, rhs, x, nb, ok, kx * rowc);
            eliminate(o, k);
        }
    }
#pragma unroll
    for (int u = 0; u < kP; ++u) {
        if (nfull + u < nz) { eliminate(ring[u], nfull + u); }
    }
    // Back substitution: x_k = e'_k - c'_k x_{k+1}.
    constexpr int kB = 8;
    R xo = eprev;
    gst(x + q0 + (nz - 1) * rowc, xo);
    const int nbk = nz - 1;
    R cb[kB], eb[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
        const int k = nz - 2 - u >= 0 ? nz - 2 - u : 0;
        cb[u] = gld(cpl + q0 + k * rowc);
        eb[u] = kLds ? R(0) : gld(eg + q0 + k * rowc);
    }
    auto substitute = [&](R ck, R ek, int k) {
        const R e = kLds ? el[k * 64 + lane] : ek;
        xo = e - ck * xo;
        gst(x + q0 + k * rowc, xo);
    };
    const int nbfull = nbk > 0 ? (nbk / kB) * kB : 0;
    for (int m0 = 0; m0 < nbfull; m0 += kB) {
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int k = nz - 2 - (m0 + u);
            const R ck = cb[u], ek = eb[u];
            const int kn = k - kB >= 0 ? k - kB : 0;
            cb[u] = gld(cpl + q0 + kn * rowc);
            eb[u] = kLds ? R(0) : gld(eg + q0 + kn * rowc);
            substitute(ck, ek, k);
        }
    }
#pragma unroll
    for (int u = 0; u < kB; ++u) {
        if (nbfull + u < nbk) { substitute(cb[u], eb[u], nz - 2 - (nbfull + u)); }
    }
}

// The same half sweep with one wave per column, for levels with too few
// columns to fill the chip one lane per column. Lane L holds levels
// L*M .. L*M+M-1; both Thomas recurrences, e'_k = (r_k - lo_k e'_{k-1}) / m_k
// and x_k = e'_k - c'_k x_{k+1}, are affine in the carried value, so each
// lane composes its levels' maps, a shuffle scan across the wave composes
// the lanes', and each lane then replays its levels from its carry-in.
template <typename R, bool kZero, bool kFromR, int M, bool kShadow>
__device__ __forceinline__ void scan_column(const Split& s, int color, int perx, int pery,
                                            const R* __restrict__ bundle, const R* __restrict__ cpl,
                                            const R* __restrict__ rhs, R* __restrict__ x, R* __restrict__ xs,
                                            int64_t t, int lane)
{
    using dev::gld;
    using dev::gst;
    constexpr bool kNb = !kZero && !kFromR;
    const int64_t rowc = int64_t(s.H) * s.ny;
    if (t >= rowc) { return; }
    const int ih = int(t % s.H), j = int(t / s.H);
    if (kShadow) {
        // The other colour's column t, when it sits on a seam: its values as
        // they stand (zero for a first sweep) into the copy its own half
        // sweep reads for its same-colour neighbours.
        const int io = 2 * ih + ((j + color + 1) & 1);
        if (io < s.nx && on_seam(io, j, s.nx, s.ny, perx, pery)) {
            const int64_t qo = int64_t(color ^ 1) * s.plane + t;
            for (int k = lane; k < s.nz; k += 64) { gst(xs + qo + k * rowc, kZero ? R(0) : gld(x + qo + k * rowc)); }
        }
    }
    const int i = 2 * ih + ((j + color) & 1);
    if (i >= s.nx) { return; }   // uniform over the wave
    const int nz = s.nz;
    const int64_t q0 = color * s.plane + t;
    int64_t nb[4] = {0, 0, 0, 0};
    bool ok[4] = {false, false, false, false};
    bool same[4] = {false, false, false, false};
    if (kNb) {
        int ni[4], nj[4];
        nb_columns(i, j, s.nx, s.ny, perx, pery, ni, nj, ok);
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            nb[p] = s.at(ni[p], nj[p], 0);
            same[p] = kShadow && ((ni[p] + nj[p]) & 1) == color;
        }
    }
    // Forward: e_k = bet_k + alp_k e_{k-1} on my levels.
    R alp[M], bet[M], g[M];
    R A = R(1), B = R(0);
#pragma unroll
    for (int u = 0; u < M; ++u) {
        const int k = lane * M + u;
        alp[u] = R(1);
        bet[u] = R(0);
        g[u] = R(1);
        if (k < nz) {
            const int64_t q = q0 + k * rowc;
            R c[kBundle];
            load_bundle(c, bundle + q * kBundle);
            R r = gld(rhs + q);
            if (kNb) {
                const int km = k > 0 ? k - 1 : 0, kp = k + 1 < nz ? k + 1 : nz - 1;
                R xm[4], x0[4], xp[4];
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    const R* src = same[p] ? xs : x;
                    const R a = gld(src + nb[p] + km * rowc), b0 = gld(src + nb[p] + k * rowc),
                            cc = gld(src + nb[p] + kp * rowc);
                    xm[p] = ok[p] ? a : R(0);
                    x0[p] = ok[p] ? b0 : R(0);
                    xp[p] = ok[p] ? cc : R(0);
                }
                r = r - nb_terms(c, xm, x0, xp);
            }
            alp[u] = -c[kBLo] * c[kBMinv];
            bet[u] = r * c[kBMinv];
            g[u] = -gld(cpl + q);
        }
        B = bet[u] + alp[u] * B;
        A = alp[u] * A;
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const R Ao = __shfl_up(A, off, 64), Bo = __shfl_up(B, off, 64);
        if (lane >= off) {
            B = B + A * Bo;
            A = A * Ao;
        }
    }
    R e = __shfl_up(B, 1, 64);
    if (lane == 0) { e = R(0); }
    R ev[M];
#pragma unroll
    for (int u = 0; u < M; ++u) {
        e = bet[u] + alp[u] * e;
        ev[u] = e;
    }
    // Backward: x_k = e_k + g_k x_{k+1} on my levels, top down.
    R C = R(1), D = R(0);
#pragma unroll
    for (int u = M - 1; u >= 0; --u) {
        const int k = lane * M + u;
        const R eu = k < nz ? ev[u] : R(0);
        D = eu + g[u] * D;
        C = g[u] * C;
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const R Co = __shfl_down(C, off, 64), Do = __shfl_down(D, off, 64);
        if (lane + off < 64) {
            D = D + C * Do;
            C = C * Co;
        }
    }
    R xv = __shfl_down(D, 1, 64);
    if (lane == 63) { xv = R(0); }
#pragma unroll
    for (int u = M - 1; u >= 0; --u) {
        const int k = lane * M + u;
        if (k < nz) {
            xv = ev[u] + g[u] * xv;
            gst(x + q0 + k * rowc, xv);
        }
    }
}

// kShadow: a level whose periodic seam joins two columns of one colour.
// Such a column reads its same-colour neighbour from xs, a copy taken before
// this half sweep (by the other colour's half sweep, or the prolongation),
// so every column sees the values the half sweep started from; and the
// launch refreshes xs for the other colour's seam columns.
template <typename R, bool kZero, bool kFromR, int M, bool kShadow>
__global__ __launch_bounds__(256) void zebra_scan_kernel(Split s, int color, int perx, int pery,
                                                         const R* __restrict__ bundle, const R* __restrict__ cpl,
                                                         const R* __restrict__ rhs, R* __restrict__ x,
                                                         R* __restrict__ xs)
{
    const int64_t t = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    scan_column<R, kZero, kFromR, M, kShadow>(s, color, perx, pery, bundle, cpl, rhs, x, xs, t,
                                              int(threadIdx.x & 63));
}

// The neighbour-column part of one colour's right-hand side,
// r = b - (neighbour columns) x, for the levels zebra_kernel<kFromR> serves.
template <typename R>
__device__ __forceinline__ void offcol_cell(const Split& s, int color, int perx, int pery, const R* __restrict__ bundle,
                                            const R* __restrict__ b, const R* __restrict__ x, R* __restrict__ r,
                                            int64_t t)
{
    const int64_t rowc = int64_t(s.H) * s.ny;
    const int ih = int(t % s.H);
    const int64_t rr = t / s.H;
    const int j = int(rr % s.ny), k = int(rr / s.ny);
    const int i = 2 * ih + ((j + color) & 1);
    if (i >= s.nx) { return; }
    int ni[4], nj[4];
    bool ok[4];
    nb_columns(i, j, s.nx, s.ny, perx, pery, ni, nj, ok);
    const int64_t q = color * s.plane + t;
    const int km = k > 0 ? k - 1 : 0, kp = k + 1 < s.nz ? k + 1 : s.nz - 1;
    R c[kBundle];
    load_bundle(c, bundle + q * kBundle);
    R xm[4], x0[4], xp[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int64_t c0 = s.at(ni[p], nj[p], 0);
        xm[p] = ok[p] ? x[c0 + km * rowc] : R(0);
        x0[p] = ok[p] ? x[c0 + k * rowc] : R(0);
        xp[p] = ok[p] ? x[c0 + kp * rowc] : R(0);
    }
    r[q] = b[q] - nb_terms(c, xm, x0, xp);
}

template <typename R>
__global__ __launch_bounds__(kThreads) void offcol_kernel(Split s, int color, int perx, int pery,
                                                          const R* __restrict__ bundle, const R* __restrict__ b,
                                                          const R* __restrict__ x, R* __restrict__ r)
{
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < s.plane; t += int64_t(gridDim.x) * blockDim.x) {
        offcol_cell(s, color, perx, pery, bundle, b, x, r, t);
    }
}

// b_c = sum J_f r_f / sum J_f over the fine cells of each coarse cell, with
// the fine residual r_f = b_f - A_f x_f formed on the fly.
template <typename R>
__global__ __launch_bounds__(kThreads) void resid_restrict_kernel(Split f, int perx, int pery,
                                                                  const R* __restrict__ bundle,
                                                                  const R* __restrict__ b, const R* __restrict__ x,
                                                                  const R* __restrict__ Js, Split sc, int rx, int ry,
                                                                  R* __restrict__ bc)
{
    const int64_t rowc = int64_t(f.H) * f.ny;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < 2 * sc.plane;
         t += int64_t(gridDim.x) * blockDim.x) {
        int I, Jc, k;
        if (!sc.cell(t, I, Jc, k)) { continue; }
        const int km = k > 0 ? k - 1 : 0, kp = k + 1 < f.nz ? k + 1 : f.nz - 1;
        R num = R(0), den = R(0);
        for (int bb = 0; bb < ry; ++bb)
            for (int a = 0; a < rx; ++a) {
                const int i = I * rx + a, j = Jc * ry + bb;
                if (i >= f.nx || j >= f.ny) { continue; }   // the last cell of an odd extent
                const int64_t q = f.at(i, j, k);
                const int64_t c0 = f.at(i, j, 0);
                R c[kBundle];
                load_bundle(c, bundle + q * kBundle);
                R v = b[q] - (c[kBLo] * x[c0 + km * rowc] + c[kBDg] * x[c0 + k * rowc] + c[kBUp] * x[c0 + kp * rowc]);
                int ni[4], nj[4];
                bool ok[4];
                nb_columns(i, j, f.nx, f.ny, perx, pery, ni, nj, ok);
                R xm[4], x0[4], xp[4];
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    const int64_t n0 = f.at(ni[p], nj[p], 0);
                    xm[p] = ok[p] ? x[n0 + km * rowc] : R(0);
                    x0[p] = ok[p] ? x[n0 + k * rowc] : R(0);
                    xp[p] = ok[p] ? x[n0 + kp * rowc] : R(0);
                }
                v = v - nb_terms(c, xm, x0, xp);
                num += Js[q] * v;
                den += Js[q];
            }
        bc[t] = num / den;
    }
}

// x_f += bilinear (in the coarsened directions) interpolation of x_c.
// xs (may be null): the seam columns' copy the shadowed half sweeps read,
// given the prolonged values too.
template <typename R>
__global__ __launch_bounds__(kThreads) void prolong_kernel(Split f, int rx, int ry, SplitMirror<R> xc,
                                                           R* __restrict__ xf, R* __restrict__ xs, int perx,
                                                           int pery)
{
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < 2 * f.plane;
         t += int64_t(gridDim.x) * blockDim.x) {
        int i, j, k;
        if (!f.cell(t, i, j, k)) { continue; }
        const int I = i / rx, Jc = j / ry;
        const int di = rx == 2 ? ((i & 1) ? 1 : -1) : 0;
        const int dj = ry == 2 ? ((j & 1) ? 1 : -1) : 0;
        const R wx = rx == 2 ? R(0.75) : R(1), wy = ry == 2 ? R(0.75) : R(1);
        R v = wx * wy * xc(I, Jc, k);
        if (di) { v += (R(1) - wx) * wy * xc(I + di, Jc, k); }
        if (dj) { v += wx * (R(1) - wy) * xc(I, Jc + dj, k); }
        if (di && dj) { v += (R(1) - wx) * (R(1) - wy) * xc(I + di, Jc + dj, k); }
        const R nv = xf[t] + v;
        xf[t] = nv;
        if (xs && on_seam(i, j, f.nx, f.ny, perx, pery)) { xs[t] = nv; }
    }
}

// Natural (i fastest) <-> split layout, with a precision change.
template <typename A, typename B>
__global__ void to_split_kernel(const A* __restrict__ nat, B* __restrict__ sp, Split s)
{
    const int64_t n = int64_t(s.nx) * s.ny * s.nz;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        int i, j, k;
        cell_of(t, s.nx, s.ny, i, j, k);
        sp[s.at(i, j, k)] = B(nat[t]);
    }
}

template <typename A, typename B>
__global__ void from_split_kernel(const A* __restrict__ sp, B* __restrict__ nat, Split s)
{
    const int64_t n = int64_t(s.nx) * s.ny * s.nz;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        int i, j, k;
        cell_of(t, s.nx, s.ny, i, j, k);
        nat[t] = B(sp[s.at(i, j, k)]);
    }
}

template <typename A, typename B>
__global__ void convert_kernel(const A* __restrict__ a, B* __restrict__ b, int64_t n)
{
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        b[t] = B(a[t]);
    }
}

template <typename R>
__global__ void fill_kernel(R* __restrict__ a, R v, int64_t n)
{
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        a[t] = v;
    }
}

// a = b * c (elementwise), the premultiplied areas.
template <typename R>
__global__ void mul_kernel(const R* __restrict__ b, const R* __restrict__ c, R* __restrict__ a, int64_t n)
{
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        a[t] = c ? b[t] * c[t] : b[t];
    }
}

// Coarse face values: the mean of the fine faces a coarse face covers.
// An odd fine extent coarsens to (n + 1) / 2 cells, the last coarse cell
// covering one fine cell: fine faces and cells past the fine extent are
// left out of the means (the last x-face maps to the domain's face).
template <typename R>
__global__ void coarsen_faces_kernel(int dir, int nxc, int nyc, int nz, int rx, int ry, int nxf, int nyf,
                                     const R* __restrict__ f, R* __restrict__ c)
{
    // face array extents (coarse): x-faces (nxc+1, nyc, nz), y (nxc, nyc+1, nz), z (nxc, nyc, nz+1)
    const int ex = nxc + (dir == 0), ey = nyc + (dir == 1), ez = nz + (dir == 2);
    const int fx = nxf + (dir == 0), fy = nyf + (dir == 1);
    const int64_t n = int64_t(ex) * ey * ez;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        int I, Jc, k;
        cell_of(t, ex, ey, I, Jc, k);
        const int na = dir == 0 ? 1 : rx, nb = dir == 1 ? 1 : ry;
        R s = R(0);
        int cnt = 0;
        for (int b = 0; b < nb; ++b)
            for (int a = 0; a < na; ++a) {
                const int i = dir == 0 ? min(I * rx, nxf) : I * rx + a;
                const int j = dir == 1 ? min(Jc * ry, nyf) : Jc * ry + b;
                if (i >= fx || j >= fy) { continue; }
                s += f[i + int64_t(j) * fx + int64_t(k) * fx * fy];
                ++cnt;
            }
        c[t] = s / R(cnt);
    }
}

// Deterministic reductions: fixed grid, per-block partials, then one block
// sums the partials in a fixed tree.
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double sh[kThreads / 64];
    for (int off = 32; off >= 1; off >>= 1) { v += __shfl_xor(v, off, 64); }
    if ((threadIdx.x & 63) == 0) { sh[threadIdx.x >> 6] = v; }
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < kThreads / 64; ++w) { s += sh[w]; }
    }
    return s;
}

// partial[y * gridDim.x + x] = sum over block x of a_y[p] * b[p] (a_y = a +
// y * lda); weight (may be null) multiplies each term.
__global__ __launch_bounds__(kThreads) void dot_partial_kernel(const double* __restrict__ a, int64_t lda,
                                                               const double* __restrict__ b, int64_t n,
                                                               double* __restrict__ partial)
{
    const double* ay = a + blockIdx.y * lda;
    double v = 0;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        v += ay[t] * b[t];
    }
    const double s = block_sum(v);
    if (threadIdx.x == 0) { partial[blockIdx.y * gridDim.x + blockIdx.x] = s; }
}

__global__ __launch_bounds__(kThreads) void sum_partial_kernel(const double* __restrict__ partial, int nblk,
                                                               double* __restrict__ out)
{
    double v = 0;
    for (int t = threadIdx.x; t < nblk; t += blockDim.x) { v += partial[blockIdx.x * nblk + t]; }
    const double s = block_sum(v);
    if (threadIdx.x == 0) { out[blockIdx.x] = s; }
}

// w -= sum_i h[i] V_i (nv vectors, stride ldv).
__global__ __launch_bounds__(kThreads) void multi_axpy_kernel(double* __restrict__ w, const double* __restrict__ V,
                                                              int64_t ldv, const double* __restrict__ h, int nv,
                                                              double sign, int64_t n)
{
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        double s = 0;
        for (int i = 0; i < nv; ++i) { s += h[i] * V[i * ldv + t]; }
        w[t] += sign * s;
    }
}

__global__ void scale_copy_kernel(const double* __restrict__ a, double s, double* __restrict__ b, int64_t n)
{
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        b[t] = a[t] * s;
    }
}

__global__ void add_scalar_kernel(double* __restrict__ a, const double* __restrict__ s, double sign, int64_t n)
{
    const double v = sign * s[0];
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        a[t] += v;
    }
}

template <typename R>
__global__ void add_scalar_kernel_r(R* __restrict__ a, const double* __restrict__ num, const double* __restrict__ den,
                                    int64_t n)
{
    const R v = R(num[0] / den[0]);
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        a[t] -= v;
    }
}

// *out = the bits of num / den (a reported mean).
__global__ void ratio_bits_kernel(const double* __restrict__ num, const double* __restrict__ den,
                                  unsigned long long* __restrict__ out)
{
    *out = __double_as_longlong(num[0] / den[0]);
}

__global__ void copy_bits_kernel(const double* __restrict__ src, unsigned long long* __restrict__ out)
{
    *out = __double_as_longlong(src[0]);
}

// GMRES(k) with its Hessenberg matrix, Givens rotations and right-hand side
// on the device (a fixed number of Arnoldi steps, no host reads: the form a
// captured graph can hold). g[] layout: [0] r0, [1] target, [2] the scale of
// the next Arnoldi vector, [3] spare, then H ((k+1) x k, column-major), cs
// (k), sn (k), rhs (k+1).
struct GmDev {
    double* g;
    int k;
    __device__ double& H(int i, int j) const { return g[4 + i + j * (k + 1)]; }
    __device__ double& cs(int i) const { return g[4 + (k + 1) * k + i]; }
    __device__ double& sn(int i) const { return g[4 + (k + 1) * k + k + i]; }
    __device__ double& rhs(int i) const { return g[4 + (k + 1) * k + 2 * k + i]; }
};

__global__ void gm_start_kernel(GmDev m, const double* __restrict__ r0sq, double reltol, double abstol)
{
    const double r0 = sqrt(*r0sq);
    m.g[0] = r0;
    m.g[1] = fmax(reltol * r0, abstol);
    m.g[2] = r0 > 0 ? 1.0 / r0 : 0.0;
    for (int i = 0; i <= m.k; ++i) { m.rhs(i) = 0.0; }
    m.rhs(0) = r0;
}

// b = a * (*s) (the Arnoldi vector from w and its scale on the device).
__global__ void scale_copy_dev_kernel(const double* __restrict__ a, const double* __restrict__ s,
                                      double* __restrict__ b, int64_t n)
{
    const double f = *s;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        b[t] = a[t] * f;
    }
}

// Column j of H from the two Gram-Schmidt passes and |w|, the earlier
// rotations applied to it, its own rotation, the rotated right-hand side.
__global__ void gm_arnoldi_kernel(GmDev m, int j, const double* __restrict__ hv, const double* __restrict__ hv2,
                                  const double* __restrict__ hn2)
{
    for (int i = 0; i <= j; ++i) { m.H(i, j) = hv[i] + hv2[i]; }
    const double hn = sqrt(*hn2);
    m.H(j + 1, j) = hn;
    m.g[2] = hn > 0 ? 1.0 / hn : 0.0;
    for (int i = 0; i < j; ++i) {
        const double a = m.H(i, j), c2 = m.H(i + 1, j);
        m.H(i, j) = m.cs(i) * a + m.sn(i) * c2;
        m.H(i + 1, j) = -m.sn(i) * a + m.cs(i) * c2;
    }
    const double a = m.H(j, j), c2 = m.H(j + 1, j);
    const double den = hypot(a, c2);
    m.cs(j) = den > 0 ? a / den : 1.0;
    m.sn(j) = den > 0 ? c2 / den : 0.0;
    m.H(j, j) = den;
    m.H(j + 1, j) = 0.0;
    m.rhs(j + 1) = -m.sn(j) * m.rhs(j);
    m.rhs(j) = m.cs(j) * m.rhs(j);
}

// y = H^-1 g (upper triangular; a zero pivot, an exhausted Krylov space,
// gives 0).
__global__ void gm_backsolve_kernel(GmDev m, double* __restrict__ y)
{
    for (int i = m.k - 1; i >= 0; --i) {
        double s = m.rhs(i);
        for (int q = i + 1; q < m.k; ++q) { s -= m.H(i, q) * y[q]; }
        const double d = m.H(i, i);
        y[i] = d != 0 ? s / d : 0.0;
    }
}

// The true residual's norm against the target: its bits, and 1 when it is
// not within the target.
__global__ void gm_check_kernel(GmDev m, const double* __restrict__ rr2, unsigned long long* __restrict__ final_bits,
                                unsigned long long* __restrict__ miss, unsigned long long* __restrict__ r0_bits)
{
    const double f = sqrt(*rr2);
    *final_bits = static_cast<unsigned long long>(__double_as_longlong(f));
    *r0_bits = static_cast<unsigned long long>(__double_as_longlong(m.g[0]));
    *miss = (f <= m.g[1]) ? 0ull : 1ull;
}

// Copy between a compact field and a view's valid region (dir 0: view ->
// compact, 1: compact -> view).
template <typename R>
__global__ void view_copy_kernel(R* __restrict__ view, int64_t vj, int64_t vk, R* __restrict__ compact, int nx,
                                 int ny, int nz, int dir)
{
    const int64_t n = int64_t(nx) * ny * nz;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        int i, j, k;
        cell_of(t, nx, ny, i, j, k);
        R* v = view + i + j * vj + int64_t(k) * vk;
        if (dir == 0) { compact[t] = *v; } else { *v = compact[t]; }
    }
}

// The three premultiplied areas in one launch (blockIdx.y: the direction):
// out_d[t] = mesh_d[t] * beta_d(view point t), the product mul_kernel forms
// after view_copy_kernel (the same bits, one pass). A null view: the mesh
// areas alone.
struct PremulSet {
    const double* mesh[3];
    const double* view[3];
    int64_t vj[3], vk[3];
    double* out[3];
    int ex[3], ey[3], ez[3];
};

__global__ void premul_kernel(PremulSet p)
{
    const int d = blockIdx.y;
    const int64_t n = int64_t(p.ex[d]) * p.ey[d] * p.ez[d];
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        if (!p.view[d]) {
            p.out[d][t] = p.mesh[d][t];
            continue;
        }
        int i, j, k;
        cell_of(t, p.ex[d], p.ey[d], i, j, k);
        p.out[d][t] = p.mesh[d][t] * p.view[d][i + j * p.vj[d] + int64_t(k) * p.vk[d]];
    }
}

// The coefficient checks of up to four views in one launch (blockIdx.y).
struct MinmaxSet {
    const double* p[4];
    int64_t sj[4], sk[4];
    int nx[4], ny[4], nz[4];
    unsigned long long* flag[4];
};

__global__ void minmax_multi_kernel(MinmaxSet s)
{
    const int q = blockIdx.y;
    const int64_t n = int64_t(s.nx[q]) * s.ny[q] * s.nz[q];
    bool bad = false;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        int i, j, k;
        cell_of(t, s.nx[q], s.ny[q], i, j, k);
        const double v = s.p[q][i + j * s.sj[q] + int64_t(k) * s.sk[q]];
        if (!(v > 0) || !(v < std::numeric_limits<double>::infinity())) { bad = true; }
    }
    if (bad) { atomicOr(s.flag[q], 1ull); }
}

// Fill the one-layer ghost shell of a view from its valid region with the
// boundary conditions (Source/Poisson.cpp:84-125).
template <typename R>
__global__ void fill_view_ghosts_kernel(R* __restrict__ view, int64_t vj, int64_t vk, MirrorAcc<R> m)
{
    const int gx = m.nx + 2, gy = m.ny + 2, gz = m.nz + 2;
    const int64_t n = int64_t(gx) * gy * gz;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        int i, j, k;
        cell_of(t, gx, gy, i, j, k);
        i -= 1;
        j -= 1;
        k -= 1;
        if (i >= 0 && i < m.nx && j >= 0 && j < m.ny && k >= 0 && k < m.nz) { continue; }
        MirrorAcc<R> vm = m;
        vm.p = view;
        vm.sj = vj;
        vm.sk = vk;
        view[i + j * vj + int64_t(k) * vk] = vm(i, j, k);
    }
}

// Fluxes F = -gradient * beta on the valid faces, zero on masked faces,
// minus the z_cross part (extern_erf/ERF_TerrainPoisson.cpp:212-264,
// Source/Poisson.cpp:192-227, 256-273).
template <typename R>
__global__ void flux_kernel(int dir, Ctx<R> c, MirrorAcc<R> x, const R* __restrict__ beta, int64_t bj, int64_t bk,
                            R* __restrict__ f, int64_t fj, int64_t fk)
{
    const int ex = c.nx + (dir == 0), ey = c.ny + (dir == 1), ez = c.nz + (dir == 2);
    const int64_t n = int64_t(ex) * ey * ez;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        int i, j, k;
        cell_of(t, ex, ey, i, j, k);
        R v;
        if (dir == 0) {
            const bool zero = ((c.zf & stencil::kXlo) && i == 0) || ((c.zf & stencil::kXhi) && i == c.nx);
            v = zero ? R(0) : -stencil::grad_x<R, R>(x, c.mt, i, j, k, c.dxinv);
        } else if (dir == 1) {
            const bool zero = ((c.zf & stencil::kYlo) && j == 0) || ((c.zf & stencil::kYhi) && j == c.ny);
            v = zero ? R(0) : -stencil::grad_y<R, R>(x, c.mt, i, j, k, c.dyinv);
        } else {
            if ((c.zf & stencil::kZhi) && k == c.nz) {
                v = R(0);
            } else if (c.dir_ground && k == 0) {
                v = -stencil::grad_z_dir<R, R>(x, c.zp, i, j, k, c.dxinv, c.dyinv);
            } else {
                v = -stencil::grad_z<R, R>(x, c.mt, i, j, k, c.dxinv, c.dyinv);
                if (k == 0) { v = -R(0); }
            }
        }
        if (beta) { v *= beta[i + j * bj + int64_t(k) * bk]; }
        if (dir == 2 && c.dcross.p && !(k == 0 || (k == c.nz && !c.top_open))) {
            v -= c.dcross(i, j, k) * stencil::z_face_cross<R, R>(x, c.mt, i, j, k, c.dxinv, c.dyinv);
        }
        f[i + j * fj + int64_t(k) * fk] = v;
    }
}

// face_divergence (Source/Poisson.cpp:733-764) on views.
__global__ void face_div_kernel(int nx, int ny, int nz, double dxinv, double dyinv, double dzinv, const double* ax,
                                int64_t axj, int64_t axk, const double* ay, int64_t ayj, int64_t ayk, const double* az,
                                int64_t azj, int64_t azk, const double* fx, int64_t fxj, int64_t fxk, const double* fy,
                                int64_t fyj, int64_t fyk, const double* fz, int64_t fzj, int64_t fzk, const double* J,
                                int64_t Jj, int64_t Jk, double* d, int64_t dj, int64_t dk)
{
    const int64_t n = int64_t(nx) * ny * nz;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        int i, j, k;
        cell_of(t, nx, ny, i, j, k);
        auto A = [&](const double* p, int64_t sj, int64_t sk, int a, int b, int c) { return p[a + b * sj + c * sk]; };
        d[i + j * dj + k * dk] =
            ((A(ax, axj, axk, i + 1, j, k) * A(fx, fxj, fxk, i + 1, j, k) - A(ax, axj, axk, i, j, k) * A(fx, fxj, fxk, i, j, k)) * dxinv
             + (A(ay, ayj, ayk, i, j + 1, k) * A(fy, fyj, fyk, i, j + 1, k) - A(ay, ayj, ayk, i, j, k) * A(fy, fyj, fyk, i, j, k)) * dyinv
             + (A(az, azj, azk, i, j, k + 1) * A(fz, fzj, fzk, i, j, k + 1) - A(az, azj, azk, i, j, k) * A(fz, fzj, fzk, i, j, k)) * dzinv)
            / A(J, Jj, Jk, i, j, k);
    }
}

inline int blocks_for(int64_t n, int cap = 8192)
{
    return int(std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, cap)));
}

// ---------------------------------------------------------------------------
// Host side.

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) { (void)hipFree(p); } }
    void ensure(size_t b)
    {
        if (bytes >= b) { return; }
        if (p) { GUST_HIP(hipFree(p)); p = nullptr; }
        GUST_HIP(hipMalloc(&p, std::max<size_t>(b, 8)));
        bytes = std::max<size_t>(b, 8);
    }
};

template <typename T>
T* ptr(const DevBuf& b)
{
    return static_cast<T*>(b.p);
}

enum { PF_PERIODIC = 0, PF_NEUMANN = 1, PF_DIRICHLET = 2 };

struct BcInfo {
    int face[6];
    int zf = 0, dir_ground = 0, top_open = 0, singular = 1;
};

BcInfo bc_info(const int* bc)
{
    BcInfo b;
    for (int f = 0; f < 6; ++f) { b.face[f] = bc[f]; }
    if (bc[0] == PF_NEUMANN) { b.zf |= stencil::kXlo; }
    if (bc[3] == PF_NEUMANN) { b.zf |= stencil::kXhi; }
    if (bc[1] == PF_NEUMANN) { b.zf |= stencil::kYlo; }
    if (bc[4] == PF_NEUMANN) { b.zf |= stencil::kYhi; }
    if (bc[5] == PF_NEUMANN) { b.zf |= stencil::kZhi; }
    b.dir_ground = bc[2] == PF_DIRICHLET;
    b.top_open = bc[5] != PF_NEUMANN;
    for (int f = 0; f < 6; ++f) { if (bc[f] == PF_DIRICHLET) { b.singular = 0; } }
    return b;
}

template <typename R>
MirrorAcc<R> mirror(const R* p, int nx, int ny, int nz, const int* per, const BcInfo& b)
{
    auto sg = [&](int f) { return b.face[f] == PF_DIRICHLET ? R(-1) : R(1); };
    MirrorAcc<R> m;
    m.p = p;
    m.nx = nx; m.ny = ny; m.nz = nz;
    m.sj = nx; m.sk = int64_t(nx) * ny;
    m.perx = per[0]; m.pery = per[1];
    m.sxl = sg(0); m.sxh = sg(3); m.syl = sg(1); m.syh = sg(4);
    m.szl = b.dir_ground ? R(-1) : R(1);   // the ground is even unless Dirichlet
    m.szh = sg(5);
    return m;
}

// One multigrid level (or the FP64 operator level). The mesh, the areas and
// the mesh factors are in natural layout (i fastest); a multigrid level's
// stencil planes and x, b, r are in the split layout.
template <typename R>
struct Level {
    int nx = 0, ny = 0, nz = 0, rx = 1, ry = 1;   // rx/ry: this level's coarsening factor from the finer one
    double dx[3] = {1, 1, 1};
    DevBuf zp, J, ax, ay, az, dcross, azm, mry, mrx, mzf;
    DevBuf coef, Js, x, b, r;
    bool has_cross = false;
    int64_t n() const { return int64_t(nx) * ny * nz; }
    Split split() const
    {
        Split s;
        s.nx = nx; s.ny = ny; s.nz = nz;
        s.H = (nx + 1) / 2;
        s.plane = int64_t(s.H) * ny * nz;
        return s;
    }
    int64_t slots() const { return 2 * split().plane; }
    Ctx<R> ctx(const BcInfo& bc) const
    {
        Ctx<R> c;
        c.zp = NodeAcc<R>{ptr<R>(zp), nx + 3, int64_t(nx + 3) * (ny + 3)};
        c.mt = stencil::ArrayMetric<R>{ptr<R>(mry), ptr<R>(mrx), ptr<R>(mzf), nx, ny, nz};
        c.ax = GridAcc<R>{ptr<R>(ax), nx + 1, int64_t(nx + 1) * ny};
        c.ay = GridAcc<R>{ptr<R>(ay), nx, int64_t(nx) * (ny + 1)};
        c.az = GridAcc<R>{ptr<R>(az), nx, int64_t(nx) * ny};
        c.J = GridAcc<R>{ptr<R>(J), nx, int64_t(nx) * ny};
        c.dcross = GridAcc<R>{has_cross ? ptr<R>(dcross) : nullptr, nx, int64_t(nx) * ny};
        c.az_mesh = GridAcc<R>{has_cross ? ptr<R>(azm) : nullptr, nx, int64_t(nx) * ny};
        c.dxinv = R(1.0 / dx[0]);
        c.dyinv = R(1.0 / dx[1]);
        c.dzinv = R(1.0 / dx[2]);
        c.nx = nx; c.ny = ny; c.nz = nz;
        c.zf = bc.zf;
        c.dir_ground = bc.dir_ground;
        c.top_open = bc.top_open;
        return c;
    }
};

// The mesh factors of a level from its node heights (setup).
template <typename R>
void fill_metric(Level<R>& L, hipStream_t st)
{
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    L.mry.ensure(size_t(nx + 1) * ny * (nz + 1) * sizeof(R));
    L.mrx.ensure(size_t(nx) * (ny + 1) * (nz + 1) * sizeof(R));
    L.mzf.ensure(size_t(3) * nx * ny * (nz + 1) * sizeof(R));
    stencil::ZMetric<R, NodeAcc<R>> zm;
    zm.zp = NodeAcc<R>{ptr<R>(L.zp), nx + 3, int64_t(nx + 3) * (ny + 3)};
    zm.dxinv = R(1.0 / L.dx[0]);
    zm.dyinv = R(1.0 / L.dx[1]);
    const int64_t tot =
        int64_t(nx + 1) * ny * (nz + 1) + int64_t(nx) * (ny + 1) * (nz + 1) + int64_t(nx) * ny * (nz + 1);
    metric_kernel<R><<<blocks_for(tot), kThreads, 0, st>>>(zm, nx, ny, nz, ptr<R>(L.mry), ptr<R>(L.mrx),
                                                           ptr<R>(L.mzf));
    GUST_HIP(hipGetLastError());
}

}   // namespace

// The solver object behind gust_poisson.
struct PoissonImpl {
    gust_geometry geom{};
    gust_poisson_options opt{};
    int nx = 0, ny = 0, nz = 0;
    // The FP64 operator level (mesh copy, current premultiplied areas).
    Level<double> L0;
    // Mesh areas (unmultiplied) on level 0, for a beta and face_divergence.
    DevBuf ax_mesh, ay_mesh, az_mesh;
    // Host copies of the node heights (for the coarse meshes).
    std::vector<double> zp_host;
    std::vector<std::unique_ptr<Level<double>>> mg64;
    std::vector<std::unique_ptr<Level<float>>> mg32;
    DevBuf krylov_v, krylov_z, w, partial, dots, hcoef, scal, gm;
    // sum J over the level-0 cells (the mean removal's denominator), formed
    // once: the mesh does not change under a solver
    DevBuf sum_j;
    bool have_sum_j = false;
    // The fixed-step form's tolerance misses: an error (default), or, for a
    // caller that adapts the step count, counted.
    bool soft_miss = false;
    int64_t misses = 0;
    std::vector<double> hbuf;
    int per[2] = {0, 0};

    int64_t n() const { return int64_t(nx) * ny * nz; }
};

namespace {

// Setup-time upload; synchronous, so h may be a temporary.
template <typename T>
void upload(DevBuf& d, const std::vector<T>& h, hipStream_t st)
{
    d.ensure(h.size() * sizeof(T));
    GUST_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
    GUST_HIP(hipStreamSynchronize(st));
}

// Read a view's region [lo, hi] (inclusive, within the view) to a compact
// host vector (setup only).
std::vector<double> view_to_host(const gust_array4& v, const int lo[3], const int hi[3], hipStream_t st)
{
    const int ex = hi[0] - lo[0] + 1, ey = hi[1] - lo[1] + 1, ez = hi[2] - lo[2] + 1;
    std::vector<double> out(static_cast<size_t>(ex) * ey * ez);
    const int64_t span = (int64_t(v.end[0] - v.begin[0])) + int64_t(v.end[1] - v.begin[1] - 1) * v.jstride +
                         int64_t(v.end[2] - v.begin[2] - 1) * v.kstride;
    std::vector<double> whole(static_cast<size_t>(span));
    GUST_HIP(hipMemcpyAsync(whole.data(), v.p, size_t(span) * 8, hipMemcpyDeviceToHost, st));
    GUST_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < ez; ++k)
        for (int j = 0; j < ey; ++j)
            for (int i = 0; i < ex; ++i) {
                out[i + size_t(j) * ex + size_t(k) * ex * ey] =
                    whole[size_t(offset_of(v, lo[0] + i, lo[1] + j, lo[2] + k))];
            }
    return out;
}

template <typename R>
void build_coarse_meshes(PoissonImpl& P, std::vector<std::unique_ptr<Level<R>>>& mg, hipStream_t st)
{
    mg.clear();
    // Level 0 of the hierarchy: the fine mesh in precision R.
    int nx = P.nx, ny = P.ny, nz = P.nz;
    std::vector<double> zp = P.zp_host;   // (nx+3)(ny+3)(nz+3)
    double dx = P.geom.dx[0], dy = P.geom.dx[1];
    const double dz = P.geom.dx[2];
    int rx = 1, ry = 1;
    for (int lev = 0;; ++lev) {
        auto L = std::make_unique<Level<R>>();
        L->nx = nx; L->ny = ny; L->nz = nz;
        L->rx = rx; L->ry = ry;
        L->dx[0] = dx; L->dx[1] = dy; L->dx[2] = dz;
        std::vector<R> zr(zp.begin(), zp.end());
        upload(L->zp, zr, st);
        fill_metric<R>(*L, st);
        // J from the node heights (extern_erf/ERF_TerrainMetrics.cpp:679-715)
        const double dzInv = 1.0 / dz;
        const int sj = nx + 3;
        const int64_t sk = int64_t(nx + 3) * (ny + 3);
        auto Z = [&](int i, int j, int k) { return zp[size_t((i + 1) + (j + 1) * sj + (k + 1) * sk)]; };
        std::vector<R> J(size_t(nx) * ny * nz);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    J[i + size_t(j) * nx + size_t(k) * nx * ny] =
                        R(0.25 * dzInv * (Z(i, j, k + 1) + Z(i + 1, j, k + 1) + Z(i, j + 1, k + 1) + Z(i + 1, j + 1, k + 1)
                                          - Z(i, j, k) - Z(i + 1, j, k) - Z(i, j + 1, k) - Z(i + 1, j + 1, k)));
                }
        upload(L->J, J, st);
        const Split sp = L->split();
        const int64_t ns = L->slots();
        std::vector<R> Js(size_t(ns), R(0));   // padding slots hold J = 0
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) { Js[size_t(sp.at(i, j, k))] = J[i + size_t(j) * nx + size_t(k) * nx * ny]; }
        upload(L->Js, Js, st);
        L->ax.ensure(size_t(nx + 1) * ny * nz * sizeof(R));
        L->ay.ensure(size_t(nx) * (ny + 1) * nz * sizeof(R));
        L->az.ensure(size_t(nx) * ny * (nz + 1) * sizeof(R));
        L->coef.ensure(size_t(kBundle + 1) * size_t(ns) * sizeof(R));   // bundles, then the c' plane
        for (DevBuf* b : {&L->x, &L->b, &L->r}) {
            b->ensure(size_t(ns) * sizeof(R));
            GUST_HIP(hipMemsetAsync(b->p, 0, size_t(ns) * sizeof(R), st));   // padding slots stay 0
        }
        GUST_HIP(hipMemsetAsync(L->coef.p, 0, size_t(kBundle + 1) * size_t(ns) * sizeof(R), st));
        mg.push_back(std::move(L));
        // Coarsen x and y by two where the size is at least 4: an even size
        // halves; an odd one (a non-periodic direction only) becomes
        // (n + 1) / 2, its last coarse cell over one fine cell and the ghost
        // node beyond the domain (the coarse grid reaches half a fine cell
        // past the side: a preconditioner's approximation).
        const int cx = (nx >= 4 && (nx % 2 == 0 || (!P.per[0] && nx >= 5))) ? 2 : 1;
        const int cy = (ny >= 4 && (ny % 2 == 0 || (!P.per[1] && ny >= 5))) ? 2 : 1;
        if ((cx == 1 && cy == 1) || lev >= 12 || int64_t(nx) * ny <= 16) { break; }
        if (P.opt.max_levels > 0 && lev + 1 >= P.opt.max_levels) { break; }
        const int ncx = (nx + cx - 1) / cx, ncy = (ny + cy - 1) / cy;
        std::vector<double> zc(size_t(ncx + 3) * (ncy + 3) * (nz + 3));
        const int csj = ncx + 3;
        const int64_t csk = int64_t(ncx + 3) * (ncy + 3);
        auto C = [&](int i, int j, int k) -> double& { return zc[size_t((i + 1) + (j + 1) * csj + (k + 1) * csk)]; };
        for (int k = 0; k <= nz; ++k)
            for (int j = -1; j <= ncy + 1; ++j)
                for (int i = -1; i <= ncx + 1; ++i) {
                    int ii = i, jj = j;
                    if (ii < 0) { ii = P.per[0] ? ii + ncx : 0; }
                    if (ii > ncx) { ii = P.per[0] ? ii - ncx : ncx; }
                    if (jj < 0) { jj = P.per[1] ? jj + ncy : 0; }
                    if (jj > ncy) { jj = P.per[1] ? jj - ncy : ncy; }
                    C(i, j, k) = Z(ii * cx, jj * cy, k);
                }
        for (int j = -1; j <= ncy + 1; ++j)
            for (int i = -1; i <= ncx + 1; ++i) {
                C(i, j, -1) = 2.0 * C(i, j, 0) - C(i, j, 1);
                C(i, j, nz + 1) = C(i, j, nz) + 1 * (C(i, j, nz) - C(i, j, nz - 1));
            }
        zp.swap(zc);
        nx = ncx; ny = ncy;
        dx *= cx; dy *= cy;
        rx = cx; ry = cy;
    }
}

template <typename R>
void launch_apply(const Level<R>& L, const BcInfo& bc, const int* per, const R* x, const R* b, R* y, hipStream_t st)
{
    const Ctx<R> c = L.ctx(bc);
    const MirrorAcc<R> m = mirror<R>(x, L.nx, L.ny, L.nz, per, bc);
    const int64_t n = L.n();
    if (b) { apply_kernel<R, 1><<<blocks_for(n), kThreads, 0, st>>>(c, m, b, y); }
    else { apply_kernel<R, 0><<<blocks_for(n), kThreads, 0, st>>>(c, m, nullptr, y); }
    GUST_HIP(hipGetLastError());
}

// A periodic direction of odd length (more than one cell) puts two columns
// of one colour side by side across the seam. Such a level either reads
// the seam neighbours from a copy (the wave-per-column kernels) or forms
// each colour's right-hand side in a pass of its own before the line solves
// (the lane-per-column kernel).
template <typename R>
bool seam_joins_colour(const Level<R>& L, const int* per)
{
    return (per[0] && L.nx > 1 && (L.nx & 1)) || (per[1] && L.ny > 1 && (L.ny & 1));
}

// e' of one wave's 64 columns in LDS when it fits this budget.
constexpr size_t kZebraLds = 64 * 1024;

// Levels with fewer columns per colour than this run the half sweeps one
// wave per column (zebra_scan_kernel); larger ones one lane per column.
int64_t scan_cols()
{
    static const int64_t v = [] {
        const char* e = std::getenv("GUST_ZEBRA_SCAN_COLS");
        return e ? std::max<int64_t>(0, std::atoll(e)) : int64_t(16384);
    }();
    return v;
}

template <typename R>
const R* bundle_of(const Level<R>& L) { return ptr<R>(L.coef); }
template <typename R>
const R* cp_of(const Level<R>& L) { return ptr<R>(L.coef) + size_t(kBundle) * size_t(L.slots()); }

// The levels whose half sweeps run one wave per column.
template <typename R>
bool scan_level(const Level<R>& L)
{
    return int64_t(L.split().H) * L.ny < scan_cols() && L.nz <= 256;
}

// A seam level smoothed by the wave-per-column kernels keeps its seam
// columns' copy in r (free on such a level: no separate right-hand side).
template <typename R>
R* seam_copy(Level<R>& L, const int* per, bool on)
{
    return on && seam_joins_colour(L, per) && scan_level(L) ? ptr<R>(L.r) : nullptr;
}

template <typename R, bool kZero, bool kFromR, int M, bool kShadow = false>
void launch_scan(const Level<R>& L, const int* per, int color, const R* rhs, hipStream_t st)
{
    const Split s = L.split();
    const int64_t rowc = int64_t(s.H) * s.ny;
    zebra_scan_kernel<R, kZero, kFromR, M, kShadow><<<int((rowc + 3) / 4), 256, 0, st>>>(
        s, color, per[0], per[1], bundle_of(L), cp_of(L), rhs, ptr<R>(L.x), kShadow ? ptr<R>(L.r) : nullptr);
}

// One zebra sweep of a seam level by the wave-per-column kernels.
template <typename R, int M>
void shadow_sweep(const Level<R>& L, const int* per, bool first, hipStream_t st)
{
    if (first) { launch_scan<R, true, false, M, true>(L, per, 0, ptr<R>(L.b), st); }
    else { launch_scan<R, false, false, M, true>(L, per, 0, ptr<R>(L.b), st); }
    launch_scan<R, false, false, M, true>(L, per, 1, ptr<R>(L.b), st);
    GUST_HIP(hipGetLastError());
}

template <typename R, bool kZero, bool kFromR>
void launch_zebra(const Level<R>& L, const int* per, int color, const R* rhs, hipStream_t st)
{
    const Split s = L.split();
    const int64_t rowc = int64_t(s.H) * s.ny;
    if (scan_level(L)) {
        const int M = (L.nz + 63) / 64;
        if (M == 1) { launch_scan<R, kZero, kFromR, 1>(L, per, color, rhs, st); }
        else if (M == 2) { launch_scan<R, kZero, kFromR, 2>(L, per, color, rhs, st); }
        else if (M == 3) { launch_scan<R, kZero, kFromR, 3>(L, per, color, rhs, st); }
        else { launch_scan<R, kZero, kFromR, 4>(L, per, color, rhs, st); }
        GUST_HIP(hipGetLastError());
        return;
    }
    const int blocks = int((rowc + 63) / 64);
    const size_t lds = size_t(64) * size_t(L.nz) * sizeof(R);
    constexpr int kP = sizeof(R) == 8 ? 2 : 6;
    if (lds <= kZebraLds) {
        zebra_kernel<R, kZero, kFromR, true, kP><<<blocks, 64, lds, st>>>(s, color, per[0], per[1], bundle_of(L),
                                                                          cp_of(L), rhs, ptr<R>(L.x), nullptr);
    } else {
        zebra_kernel<R, kZero, kFromR, false, kP><<<blocks, 64, 0, st>>>(s, color, per[0], per[1], bundle_of(L),
                                                                         cp_of(L), rhs, ptr<R>(L.x), ptr<R>(L.r));
    }
    GUST_HIP(hipGetLastError());
}

// One zebra sweep (colour 0, then colour 1); first: x is zero on entry;
// copy: a seam level smoothed one wave per column reads its seam neighbours
// from their copy (else: a pass of its own forms the neighbour terms).
template <typename R>
void zebra(Level<R>& L, const int* per, bool first, bool copy, hipStream_t st)
{
    if (!seam_joins_colour(L, per)) {
        if (first) { launch_zebra<R, true, false>(L, per, 0, ptr<R>(L.b), st); }
        else { launch_zebra<R, false, false>(L, per, 0, ptr<R>(L.b), st); }
        launch_zebra<R, false, false>(L, per, 1, ptr<R>(L.b), st);
        return;
    }
    if (copy && scan_level(L)) {
        const int M = (L.nz + 63) / 64;
        if (M == 1) { shadow_sweep<R, 1>(L, per, first, st); }
        else if (M == 2) { shadow_sweep<R, 2>(L, per, first, st); }
        else if (M == 3) { shadow_sweep<R, 3>(L, per, first, st); }
        else { shadow_sweep<R, 4>(L, per, first, st); }
        return;
    }
    const Split s = L.split();
    if (first) { fill_kernel<R><<<blocks_for(L.slots()), kThreads, 0, st>>>(ptr<R>(L.x), R(0), L.slots()); }
    for (int color = 0; color < 2; ++color) {
        offcol_kernel<R><<<blocks_for(s.plane), kThreads, 0, st>>>(s, color, per[0], per[1], bundle_of(L),
                                                                   ptr<R>(L.b), ptr<R>(L.x), ptr<R>(L.r));
        GUST_HIP(hipGetLastError());
        launch_zebra<R, false, true>(L, per, color, ptr<R>(L.r), st);
    }
}

// J-weighted mean removal of a level's x (the coarsest solve of a singular
// problem): x -= sum J x / sum J, over the split slots (padding has J = 0).
template <typename R>
void remove_mean(PoissonImpl& P, Level<R>& L, hipStream_t st);

template <>
void remove_mean<double>(PoissonImpl& P, Level<double>& L, hipStream_t st)
{
    const int64_t n = L.slots();
    double* d = ptr<double>(P.dots);
    dot_partial_kernel<<<dim3(kRedBlocks, 1), kThreads, 0, st>>>(ptr<double>(L.Js), 0, ptr<double>(L.x), n,
                                                                 ptr<double>(P.partial));
    sum_partial_kernel<<<1, kThreads, 0, st>>>(ptr<double>(P.partial), kRedBlocks, d);
    // sum J: dot(J, ones), with L.r (free at the coarsest level) as the ones
    fill_kernel<double><<<blocks_for(n), kThreads, 0, st>>>(ptr<double>(L.r), 1.0, n);
    dot_partial_kernel<<<dim3(kRedBlocks, 1), kThreads, 0, st>>>(ptr<double>(L.Js), 0, ptr<double>(L.r), n,
                                                                 ptr<double>(P.partial));
    sum_partial_kernel<<<1, kThreads, 0, st>>>(ptr<double>(P.partial), kRedBlocks, d + 1);
    add_scalar_kernel_r<double><<<blocks_for(n), kThreads, 0, st>>>(ptr<double>(L.x), d, d + 1, n);
    GUST_HIP(hipGetLastError());
}

template <>
void remove_mean<float>(PoissonImpl& P, Level<float>& L, hipStream_t st)
{
    // promote to double for the reduction
    const int64_t n = L.slots();
    double* tmp = ptr<double>(P.w);   // 2 n doubles of scratch (w holds 2 level-0 vectors)
    double* tj = tmp + n;
    convert_kernel<float, double><<<blocks_for(n), kThreads, 0, st>>>(ptr<float>(L.x), tmp, n);
    convert_kernel<float, double><<<blocks_for(n), kThreads, 0, st>>>(ptr<float>(L.Js), tj, n);
    double* d = ptr<double>(P.dots);
    dot_partial_kernel<<<dim3(kRedBlocks, 1), kThreads, 0, st>>>(tj, 0, tmp, n, ptr<double>(P.partial));
    sum_partial_kernel<<<1, kThreads, 0, st>>>(ptr<double>(P.partial), kRedBlocks, d);
    fill_kernel<double><<<blocks_for(n), kThreads, 0, st>>>(tmp, 1.0, n);
    dot_partial_kernel<<<dim3(kRedBlocks, 1), kThreads, 0, st>>>(tj, 0, tmp, n, ptr<double>(P.partial));
    sum_partial_kernel<<<1, kThreads, 0, st>>>(ptr<double>(P.partial), kRedBlocks, d + 1);
    add_scalar_kernel_r<float><<<blocks_for(n), kThreads, 0, st>>>(ptr<float>(L.x), d, d + 1, n);
    GUST_HIP(hipGetLastError());
}

template <typename R>
void vcycle(PoissonImpl& P, std::vector<std::unique_ptr<Level<R>>>& mg, size_t l, const BcInfo& bc,
            hipStream_t st)
{
    Level<R>& L = *mg[l];
    if (l + 1 == mg.size()) {
        for (int s = 0; s < P.opt.coarse_sweeps; ++s) { zebra<R>(L, P.per, s == 0, P.opt.seam_copy != 0, st); }
        if (bc.singular) { remove_mean<R>(P, L, st); }
        return;
    }
    if (P.opt.pre_sweeps == 0) {
        fill_kernel<R><<<blocks_for(L.slots()), kThreads, 0, st>>>(ptr<R>(L.x), R(0), L.slots());
    }
    for (int s = 0; s < P.opt.pre_sweeps; ++s) { zebra<R>(L, P.per, s == 0, P.opt.seam_copy != 0, st); }
    Level<R>& C = *mg[l + 1];
    const Split sf = L.split(), sc = C.split();
    resid_restrict_kernel<R><<<blocks_for(C.slots()), kThreads, 0, st>>>(sf, P.per[0], P.per[1], bundle_of(L),
                                                                         ptr<R>(L.b), ptr<R>(L.x), ptr<R>(L.Js), sc,
                                                                         C.rx, C.ry, ptr<R>(C.b));
    GUST_HIP(hipGetLastError());
    vcycle<R>(P, mg, l + 1, bc, st);
    SplitMirror<R> mc;
    mc.m = mirror<R>(nullptr, C.nx, C.ny, C.nz, P.per, bc);
    mc.s = sc;
    mc.p = ptr<R>(C.x);
    prolong_kernel<R><<<blocks_for(L.slots()), kThreads, 0, st>>>(sf, C.rx, C.ry, mc, ptr<R>(L.x),
                                                                   seam_copy(L, P.per, P.opt.seam_copy != 0), P.per[0],
                                                                   P.per[1]);
    GUST_HIP(hipGetLastError());
    for (int s = 0; s < P.opt.post_sweeps; ++s) { zebra<R>(L, P.per, false, P.opt.seam_copy != 0, st); }
}

// Premultiplied areas and the explicit stencil (with its column Thomas
// factors) on every level of a hierarchy, from level 0's FP64 areas.
template <typename R>
void setup_hierarchy(PoissonImpl& P, std::vector<std::unique_ptr<Level<R>>>& mg, const BcInfo& bc, hipStream_t st)
{
    for (size_t l = 0; l < mg.size(); ++l) {
        Level<R>& L = *mg[l];
        if (l == 0) {
            const int64_t nxf = int64_t(L.nx + 1) * L.ny * L.nz, nyf = int64_t(L.nx) * (L.ny + 1) * L.nz,
                          nzf = int64_t(L.nx) * L.ny * (L.nz + 1);
            convert_kernel<double, R><<<blocks_for(nxf), kThreads, 0, st>>>(ptr<double>(P.L0.ax), ptr<R>(L.ax), nxf);
            convert_kernel<double, R><<<blocks_for(nyf), kThreads, 0, st>>>(ptr<double>(P.L0.ay), ptr<R>(L.ay), nyf);
            convert_kernel<double, R><<<blocks_for(nzf), kThreads, 0, st>>>(ptr<double>(P.L0.az), ptr<R>(L.az), nzf);
            L.has_cross = P.L0.has_cross;
            if (L.has_cross) {
                L.dcross.ensure(size_t(nzf) * sizeof(R));
                L.azm.ensure(size_t(nzf) * sizeof(R));
                convert_kernel<double, R><<<blocks_for(nzf), kThreads, 0, st>>>(ptr<double>(P.L0.dcross),
                                                                                ptr<R>(L.dcross), nzf);
                convert_kernel<double, R><<<blocks_for(nzf), kThreads, 0, st>>>(ptr<double>(P.L0.azm),
                                                                                ptr<R>(L.azm), nzf);
            }
        } else {
            Level<R>& F = *mg[l - 1];
            for (int dir = 0; dir < 3; ++dir) {
                const int64_t ne = int64_t(L.nx + (dir == 0)) * (L.ny + (dir == 1)) * (L.nz + (dir == 2));
                const R* f = dir == 0 ? ptr<R>(F.ax) : dir == 1 ? ptr<R>(F.ay) : ptr<R>(F.az);
                R* c = dir == 0 ? ptr<R>(L.ax) : dir == 1 ? ptr<R>(L.ay) : ptr<R>(L.az);
                coarsen_faces_kernel<R><<<blocks_for(ne), kThreads, 0, st>>>(dir, L.nx, L.ny, L.nz, L.rx, L.ry, F.nx,
                                                                              F.ny, f, c);
            }
            L.has_cross = F.has_cross;
            if (L.has_cross) {
                const int64_t nzf = int64_t(L.nx) * L.ny * (L.nz + 1);
                L.dcross.ensure(size_t(nzf) * sizeof(R));
                L.azm.ensure(size_t(nzf) * sizeof(R));
                coarsen_faces_kernel<R><<<blocks_for(nzf), kThreads, 0, st>>>(2, L.nx, L.ny, L.nz, L.rx, L.ry, F.nx,
                                                                              F.ny, ptr<R>(F.dcross),
                                                                              ptr<R>(L.dcross));
                coarsen_faces_kernel<R><<<blocks_for(nzf), kThreads, 0, st>>>(2, L.nx, L.ny, L.nz, L.rx, L.ry, F.nx,
                                                                              F.ny, ptr<R>(F.azm), ptr<R>(L.azm));
            }
        }
        GUST_HIP(hipGetLastError());
        const Ctx<R> c = L.ctx(bc);
        const MirrorAcc<R> m = mirror<R>(nullptr, L.nx, L.ny, L.nz, P.per, bc);
        const Split sp = L.split();
        coef_seed_kernel<R><<<blocks_for(L.slots()), kThreads, 0, st>>>(c, m, sp, P.per[0], P.per[1], ptr<R>(L.coef));
        GUST_HIP(hipGetLastError());
        coef_factor_kernel<R><<<blocks_for(2 * int64_t(sp.H) * sp.ny), kThreads, 0, st>>>(
            sp, ptr<R>(L.coef), ptr<R>(L.coef) + size_t(kBundle) * size_t(L.slots()));
        GUST_HIP(hipGetLastError());
    }
}

// z = M^{-1} v (one V-cycle) in the configured precision.
void precondition(PoissonImpl& P, const double* v, double* z, const BcInfo& bc, hipStream_t st)
{
    const int64_t n = P.n();
    if (P.opt.precision == 1) {
        Level<float>& L = *P.mg32[0];
        to_split_kernel<double, float><<<blocks_for(n), kThreads, 0, st>>>(v, ptr<float>(L.b), L.split());
        vcycle<float>(P, P.mg32, 0, bc, st);
        from_split_kernel<float, double><<<blocks_for(n), kThreads, 0, st>>>(ptr<float>(L.x), z, L.split());
    } else {
        Level<double>& L = *P.mg64[0];
        to_split_kernel<double, double><<<blocks_for(n), kThreads, 0, st>>>(v, ptr<double>(L.b), L.split());
        vcycle<double>(P, P.mg64, 0, bc, st);
        from_split_kernel<double, double><<<blocks_for(n), kThreads, 0, st>>>(ptr<double>(L.x), z, L.split());
    }
    GUST_HIP(hipGetLastError());
}

// dots[i] = <V_i, w> for i < nv; copied to host.
void multi_dot(PoissonImpl& P, const double* V, int64_t ldv, int nv, const double* w, double* host, hipStream_t st)
{
    const int64_t n = P.n();
    dot_partial_kernel<<<dim3(kRedBlocks, nv), kThreads, 0, st>>>(V, ldv, w, n, ptr<double>(P.partial));
    sum_partial_kernel<<<nv, kThreads, 0, st>>>(ptr<double>(P.partial), kRedBlocks, ptr<double>(P.dots));
    GUST_HIP(hipGetLastError());
    auto* h = static_cast<double*>(Workspace::get_host_pinned(1, size_t(nv) * 8));
    GUST_HIP(hipMemcpyAsync(h, P.dots.p, size_t(nv) * 8, hipMemcpyDeviceToHost, st));
    GUST_HIP(hipStreamSynchronize(st));
    std::memcpy(host, h, size_t(nv) * 8);
}

// out[i] = <V_i, w> for i < nv, left on the device.
void multi_dot_dev(PoissonImpl& P, const double* V, int64_t ldv, int nv, const double* w, double* out, hipStream_t st)
{
    const int64_t n = P.n();
    dot_partial_kernel<<<dim3(kRedBlocks, nv), kThreads, 0, st>>>(V, ldv, w, n, ptr<double>(P.partial));
    sum_partial_kernel<<<nv, kThreads, 0, st>>>(ptr<double>(P.partial), kRedBlocks, out);
    GUST_HIP(hipGetLastError());
}

// v -= sum J v / sum J on the device (a level-0 compact field); nd[0], nd[1]
// keep the sums.
void remove_weighted_mean(PoissonImpl& P, double* v, double* nd, hipStream_t st)
{
    const int64_t n = P.n();
    if (!P.have_sum_j) {
        P.sum_j.ensure(8);
        fill_kernel<double><<<blocks_for(n), kThreads, 0, st>>>(ptr<double>(P.w), 1.0, n);
        multi_dot_dev(P, ptr<double>(P.L0.J), 0, 1, ptr<double>(P.w), ptr<double>(P.sum_j), st);
        P.have_sum_j = true;
    }
    multi_dot_dev(P, ptr<double>(P.L0.J), 0, 1, v, nd, st);
    add_scalar_kernel_r<double><<<blocks_for(n), kThreads, 0, st>>>(v, nd, ptr<double>(P.sum_j), n);
    GUST_HIP(hipGetLastError());
}

double norm2(PoissonImpl& P, const double* v, hipStream_t st)
{
    double d;
    multi_dot(P, v, 0, 1, v, &d, st);
    return std::sqrt(d);
}

void set_coef(PoissonImpl& P, const double* h, int nv, hipStream_t st)
{
    // Every upload is followed by a wait (the next residual norm) before the
    // pinned buffer is written again.
    auto* hp = static_cast<double*>(Workspace::get_host_pinned(2, size_t(nv) * 8));
    std::memcpy(hp, h, size_t(nv) * 8);
    GUST_HIP(hipMemcpyAsync(P.hcoef.p, hp, size_t(nv) * 8, hipMemcpyHostToDevice, st));
}

void check_poisson_options(const gust_poisson_options& o)
{
    std::string err;
    const double eps = std::numeric_limits<double>::epsilon();
    if (!std::isfinite(o.reltol) || !(o.reltol >= 10 * eps) || !(o.reltol < 1)) {
        char b[160];
        std::snprintf(b, sizeof b, "\n  solver.poisson_reltol must be >= %g (10 machine epsilon) and < 1", 10 * eps);
        err += b;
    }
    if (!std::isfinite(o.abstol) || !(o.abstol >= 0)) { err += "\n  solver.poisson_abstol must be finite and >= 0"; }
    if (o.max_iter < 1) { err += "\n  solver.poisson_max_iter must be >= 1"; }
    if (o.restart < 1) { err += "\n  solver.poisson_restart must be >= 1"; }
    if (o.precision != 0 && o.precision != 1) { err += "\n  precision must be 0 (fp64) or 1 (fp32 V-cycle)"; }
    if (o.max_levels < 0) { err += "\n  the multigrid's max_levels must be >= 0 (0: no limit)"; }
    if (o.fixed_steps < 0 || o.fixed_steps > 200) {
        GUST_FAIL("gust: PoissonOptions: fixed_steps must be 0 (adaptive GMRES) or 1..200 (got %d)", o.fixed_steps);
    }
    if (o.pre_sweeps < 0 || o.post_sweeps < 0 || o.coarse_sweeps < 1) {
        err += "\n  the V-cycle needs pre/post sweeps >= 0 and coarse sweeps >= 1";
    }
    if (o.seam_copy != 0 && o.seam_copy != 1) { err += "\n  the multigrid's seam_copy must be 0 or 1"; }
    if (!err.empty()) { GUST_FAIL("gust: invalid inputs:%s", err.c_str()); }
}

void check_bc(const PoissonImpl& P, const int* bc)
{
    if (!bc) { GUST_FAIL("gust: PoissonSolver: the boundary conditions are null"); }
    std::string err;
    const char* nm[3] = {"x", "y", "z"};
    for (int d = 0; d < 3; ++d) {
        for (int s = 0; s < 2; ++s) {
            const int f = bc[d + 3 * s];
            if (f < 0 || f > 2) { GUST_FAIL("gust: PoissonSolver: unknown boundary code %d", f); }
            const bool per = d < 2 ? P.per[d] != 0 : false;
            if (per != (f == PF_PERIODIC)) {
                err += std::string("\n  Poisson BC ") + nm[d] + (s ? "hi" : "lo") +
                       ": periodic exactly when the domain is periodic in " + nm[d];
            }
        }
    }
    if (!err.empty()) { GUST_FAIL("gust: invalid inputs:%s", err.c_str()); }
}

// Check and bring in the face coefficients; form a beta on level 0 and the
// z_cross correction (Source/Poisson.cpp:421-515).
void prepare(PoissonImpl& P, const gust_face_coeffs* fc, const BcInfo& bc, Deferred& dq, hipStream_t st)
{
    const gust_array4* b[4] = {nullptr, nullptr, nullptr, nullptr};
    if (fc) { b[0] = fc->x; b[1] = fc->y; b[2] = fc->z; b[3] = fc->z_cross; }
    const char* bname[4] = {"FaceCoeffs::x", "FaceCoeffs::y", "FaceCoeffs::z", "FaceCoeffs::z_cross"};
    MinmaxSet mm{};
    int nmm = 0;
    int64_t most = 1;
    for (int d = 0; d < 4; ++d) {
        if (!b[d]) { continue; }
        const int dir = d < 3 ? d : 2;
        IBox fb;
        for (int q = 0; q < 3; ++q) {
            fb.lo[q] = P.geom.lo[q];
            fb.hi[q] = P.geom.hi[q] + (q == dir ? 1 : 0);
        }
        if (!covers(b[d], fb, 1)) { GUST_FAIL("gust: %s must be a face MultiFab on the mesh's box array", bname[d]); }
        // the flag is read with the batch's other values (at the latest when
        // the solve reads its initial residual)
        const char* name = bname[d];
        mm.flag[nmm] = dq.zero_slot([name](unsigned long long h) {
            if (h) { GUST_FAIL("gust: %s must be finite and > 0 everywhere", name); }
        });
        mm.p[nmm] = static_cast<const double*>(b[d]->p) + offset_of(*b[d], fb.lo[0], fb.lo[1], fb.lo[2]);
        mm.sj[nmm] = b[d]->jstride;
        mm.sk[nmm] = b[d]->kstride;
        mm.nx[nmm] = fb.len(0);
        mm.ny[nmm] = fb.len(1);
        mm.nz[nmm] = fb.len(2);
        most = std::max<int64_t>(most, fb.npts());
        ++nmm;
    }
    if (nmm > 0) {
        minmax_multi_kernel<<<dim3(blocks_for(most), nmm), kThreads, 0, st>>>(mm);
        GUST_HIP(hipGetLastError());
    }
    if (b[3] && bc.dir_ground) { GUST_FAIL("gust: FaceCoeffs::z_cross with a Dirichlet ground is not supported"); }
    if (fc && (!std::isfinite(fc->precond_ratio) || !(fc->precond_ratio > 0))) {
        GUST_FAIL("gust: FaceCoeffs::precond_ratio must be finite and > 0");
    }
    // a beta, as compact valid-face arrays.
    DevBuf* mesh[3] = {&P.ax_mesh, &P.ay_mesh, &P.az_mesh};
    DevBuf* out[3] = {&P.L0.ax, &P.L0.ay, &P.L0.az};
    PremulSet pm{};
    int64_t most_f = 1;
    for (int d = 0; d < 3; ++d) {
        pm.ex[d] = P.nx + (d == 0);
        pm.ey[d] = P.ny + (d == 1);
        pm.ez[d] = P.nz + (d == 2);
        pm.mesh[d] = ptr<double>(*mesh[d]);
        pm.out[d] = ptr<double>(*out[d]);
        pm.view[d] = b[d] ? static_cast<const double*>(b[d]->p) +
                                offset_of(*b[d], P.geom.lo[0], P.geom.lo[1], P.geom.lo[2])
                          : nullptr;
        pm.vj[d] = b[d] ? b[d]->jstride : 0;
        pm.vk[d] = b[d] ? b[d]->kstride : 0;
        most_f = std::max<int64_t>(most_f, int64_t(pm.ex[d]) * pm.ey[d] * pm.ez[d]);
    }
    premul_kernel<<<dim3(blocks_for(most_f), 3), kThreads, 0, st>>>(pm);
    GUST_HIP(hipGetLastError());
    P.L0.has_cross = b[3] != nullptr;
    if (P.L0.has_cross) {
        const int64_t ne = int64_t(P.nx) * P.ny * (P.nz + 1);
        P.L0.dcross.ensure(size_t(ne) * 8);
        P.L0.azm.ensure(size_t(ne) * 8);
        double* d = ptr<double>(P.L0.dcross);
        double* tmp = ptr<double>(P.w);
        const gust_array4* bzc = b[3];
        if (b[2]) {
            view_copy_kernel<double><<<blocks_for(ne), kThreads, 0, st>>>(
                static_cast<double*>(b[2]->p) + offset_of(*b[2], P.geom.lo[0], P.geom.lo[1], P.geom.lo[2]),
                b[2]->jstride, b[2]->kstride, d, P.nx, P.ny, P.nz + 1, 0);
        } else {
            fill_kernel<double><<<blocks_for(ne), kThreads, 0, st>>>(d, 1.0, ne);
        }
        view_copy_kernel<double><<<blocks_for(ne), kThreads, 0, st>>>(
            static_cast<double*>(bzc->p) + offset_of(*bzc, P.geom.lo[0], P.geom.lo[1], P.geom.lo[2]), bzc->jstride,
            bzc->kstride, tmp, P.nx, P.ny, P.nz + 1, 0);
        // d = beta_z - z_cross
        multi_axpy_kernel<<<blocks_for(ne), kThreads, 0, st>>>(d, tmp, 0, ptr<double>(P.scal), 1, -1.0, ne);
        GUST_HIP(hipMemcpyAsync(P.L0.azm.p, P.az_mesh.p, size_t(ne) * 8, hipMemcpyDeviceToDevice, st));
        GUST_HIP(hipGetLastError());
    }
}

void phi_in_out(PoissonImpl& P, gust_array4* phi, double* compact, int dir, hipStream_t st)
{
    const int64_t n = P.n();
    view_copy_kernel<double><<<blocks_for(n), kThreads, 0, st>>>(
        static_cast<double*>(phi->p) + offset_of(*phi, P.geom.lo[0], P.geom.lo[1], P.geom.lo[2]), phi->jstride,
        phi->kstride, compact, P.nx, P.ny, P.nz, dir);
    GUST_HIP(hipGetLastError());
}

void fill_phi_ghosts(PoissonImpl& P, gust_array4* phi, const BcInfo& bc, hipStream_t st)
{
    const MirrorAcc<double> m = mirror<double>(nullptr, P.nx, P.ny, P.nz, P.per, bc);
    const int64_t n = int64_t(P.nx + 2) * (P.ny + 2) * (P.nz + 2);
    fill_view_ghosts_kernel<double><<<blocks_for(n), kThreads, 0, st>>>(
        static_cast<double*>(phi->p) + offset_of(*phi, P.geom.lo[0], P.geom.lo[1], P.geom.lo[2]), phi->jstride,
        phi->kstride, m);
    GUST_HIP(hipGetLastError());
}

void check_phi(const PoissonImpl& P, const gust_array4* phi, const char* what)
{
    IBox V;
    for (int q = 0; q < 3; ++q) { V.lo[q] = P.geom.lo[q]; V.hi[q] = P.geom.hi[q]; }
    if (!covers(phi, grow(V, 1, 1, 1), 1)) { GUST_FAIL("gust: PoissonSolver: %s needs a ghost cell", what); }
}

void check_cells(const PoissonImpl& P, const gust_array4* a, const char* what)
{
    IBox V;
    for (int q = 0; q < 3; ++q) { V.lo[q] = P.geom.lo[q]; V.hi[q] = P.geom.hi[q]; }
    if (!covers(a, V, 1)) {
        GUST_FAIL("gust: PoissonSolver: %s must be on the mesh's BoxArray and DistributionMapping", what);
    }
}

void ensure_work(PoissonImpl& P, int m)
{
    const int64_t n = P.n();
    // two level-0 vectors, or two split-layout vectors of any level
    // (remove_mean<float>): (nx + 1) ny nz covers a split level's slots
    const int64_t slots = int64_t(P.nx + 1) * P.ny * P.nz;
    P.w.ensure(size_t(std::max(n, slots) + 1) * 8 * 2 + 64);
    P.krylov_v.ensure(size_t(n) * 8 * size_t(m + 1));
    P.krylov_z.ensure(size_t(n) * 8 * size_t(m));
    P.partial.ensure(size_t(kRedBlocks) * 8 * size_t(m + 2));
    P.dots.ensure(size_t(2 * (m + 2) + 8) * 8);
    P.hcoef.ensure(size_t(m + 2) * 8 + 64);
}

void apply_fp64(PoissonImpl& P, const BcInfo& bc, const double* x, double* y, hipStream_t st)
{
    launch_apply<double>(P.L0, bc, P.per, x, nullptr, y, st);
}

// The fluxes F = -beta grad phi of the solution x on the valid faces.
void solve_fluxes(PoissonImpl& P, const BcInfo& b, const gust_face_coeffs* coeffs, const double* x, gust_array4* fx,
                  gust_array4* fy, gust_array4* fz, hipStream_t st)
{
    const Ctx<double> c = P.L0.ctx(b);
    const MirrorAcc<double> mm = mirror<double>(x, P.nx, P.ny, P.nz, P.per, b);
    gust_array4* F[3] = {fx, fy, fz};
    const gust_array4* beta3[3] = {coeffs ? coeffs->x : nullptr, coeffs ? coeffs->y : nullptr,
                                   coeffs ? coeffs->z : nullptr};
    for (int d = 0; d < 3; ++d) {
        IBox fb;
        for (int q = 0; q < 3; ++q) { fb.lo[q] = P.geom.lo[q]; fb.hi[q] = P.geom.hi[q] + (q == d); }
        const double* bp = beta3[d] ? static_cast<const double*>(beta3[d]->p) +
                                          offset_of(*beta3[d], fb.lo[0], fb.lo[1], fb.lo[2])
                                    : nullptr;
        const int64_t ne = int64_t(fb.len(0)) * fb.len(1) * fb.len(2);
        flux_kernel<double><<<blocks_for(ne), kThreads, 0, st>>>(
            d, c, mm, bp, beta3[d] ? beta3[d]->jstride : 0, beta3[d] ? beta3[d]->kstride : 0,
            static_cast<double*>(F[d]->p) + offset_of(*F[d], fb.lo[0], fb.lo[1], fb.lo[2]), F[d]->jstride,
            F[d]->kstride);
        GUST_HIP(hipGetLastError());
    }
}

}   // namespace

}   // namespace gust_amd

using namespace gust_amd;

struct gust_poisson {
    PoissonImpl impl;
};

extern "C" {

int gust_poisson_default_options(gust_poisson_options* o)
{
    return guarded([&] {
        if (!o) { GUST_FAIL("gust: poisson options are null"); }
        o->reltol = 1e-10;
        o->abstol = 0;
        o->max_iter = 200;
        o->restart = 50;
        o->verbose = 0;
        o->precision = 0;
        o->pre_sweeps = 2;
        o->post_sweeps = 2;
        o->coarse_sweeps = 16;
        o->max_levels = 0;
        o->fixed_steps = 0;
        o->seam_copy = 1;
    });
}

int gust_poisson_set_fixed_steps(gust_poisson* p, int32_t steps)
{
    return guarded([&] {
        if (!p) { GUST_FAIL("gust: PoissonSolver: null solver"); }
        gust_poisson_options o = p->impl.opt;
        o.fixed_steps = steps;
        check_poisson_options(o);
        p->impl.opt.fixed_steps = steps;
    });
}

int gust_poisson_fixed_steps_misses(gust_poisson* p, int32_t soft, int64_t* misses)
{
    return guarded([&] {
        if (!p) { GUST_FAIL("gust: PoissonSolver: null solver"); }
        p->impl.soft_miss = soft != 0;
        if (misses) { *misses = p->impl.misses; }
    });
}

int gust_poisson_create(const gust_geometry* geom, const gust_array4* z_nd, const gust_array4* detJ,
                        const gust_array4* ax, const gust_array4* ay, const gust_array4* az,
                        const gust_poisson_options* opt, gust_poisson** out, void* stream)
{
    return guarded([&] {
        if (!geom || !z_nd || !detJ || !ax || !ay || !az || !opt || !out) {
            GUST_FAIL("gust: PoissonSolver: null argument");
        }
        *out = nullptr;
        hipStream_t st = (hipStream_t)stream;
        check_poisson_options(*opt);
        if (geom->periodic[2]) { GUST_FAIL("gust: PoissonSolver: the domain cannot be periodic in z"); }
        auto h = std::make_unique<gust_poisson>();
        PoissonImpl& P = h->impl;
        P.geom = *geom;
        P.opt = *opt;
        P.nx = geom->hi[0] - geom->lo[0] + 1;
        P.ny = geom->hi[1] - geom->lo[1] + 1;
        P.nz = geom->hi[2] - geom->lo[2] + 1;
        if (P.nx < 1 || P.ny < 1 || P.nz < 2) { GUST_FAIL("gust: PoissonSolver: the domain needs nz >= 2"); }
        P.per[0] = geom->periodic[0] != 0;
        P.per[1] = geom->periodic[1] != 0;
        // node heights with one ghost node
        IBox nodes;
        for (int q = 0; q < 3; ++q) { nodes.lo[q] = geom->lo[q] - 1; nodes.hi[q] = geom->hi[q] + 2; }
        if (!covers(z_nd, nodes, 1)) { GUST_FAIL("gust: PoissonSolver: z_phys_nd needs one ghost node"); }
        IBox cells;
        for (int q = 0; q < 3; ++q) { cells.lo[q] = geom->lo[q]; cells.hi[q] = geom->hi[q]; }
        if (!covers(detJ, cells, 1)) { GUST_FAIL("gust: PoissonSolver: detJ must cover the domain"); }
        P.zp_host = view_to_host(*z_nd, nodes.lo, nodes.hi, st);
        auto faces = [&](const gust_array4* a, int dir, DevBuf& dst) {
            IBox fb = cells;
            fb.hi[dir] += 1;
            if (!covers(a, fb, 1)) { GUST_FAIL("gust: PoissonSolver: the mesh areas must cover the faces"); }
            upload(dst, view_to_host(*a, fb.lo, fb.hi, st), st);
        };
        faces(ax, 0, P.ax_mesh);
        faces(ay, 1, P.ay_mesh);
        faces(az, 2, P.az_mesh);
        const int64_t n = P.n();
        // level 0 FP64: the mesh as given (J from the caller, zp from the caller)
        Level<double>& L = P.L0;
        L.nx = P.nx; L.ny = P.ny; L.nz = P.nz;
        for (int q = 0; q < 3; ++q) { L.dx[q] = geom->dx[q]; }
        upload(L.zp, P.zp_host, st);
        fill_metric<double>(L, st);
        upload(L.J, view_to_host(*detJ, cells.lo, cells.hi, st), st);
        L.ax.ensure(size_t(P.nx + 1) * P.ny * P.nz * 8);
        L.ay.ensure(size_t(P.nx) * (P.ny + 1) * P.nz * 8);
        L.az.ensure(size_t(P.nx) * P.ny * (P.nz + 1) * 8);
        for (DevBuf* b : {&L.x, &L.b, &L.r}) { b->ensure(size_t(n) * 8); }
        if (opt->precision == 1) { build_coarse_meshes<float>(P, P.mg32, st); }
        else { build_coarse_meshes<double>(P, P.mg64, st); }
        P.scal.ensure(64);
        fill_kernel<double><<<1, 1, 0, st>>>(ptr<double>(P.scal), 1.0, 1);
        ensure_work(P, std::max(1, opt->restart));
        GUST_HIP(hipStreamSynchronize(st));
        *out = h.release();
    });
}

void gust_poisson_destroy(gust_poisson* p) { delete p; }

int gust_poisson_levels(const gust_poisson* p, int32_t* levels)
{
    return guarded([&] {
        if (!p || !levels) { GUST_FAIL("gust: poisson handle is null"); }
        *levels = int32_t(p->impl.opt.precision == 1 ? p->impl.mg32.size() : p->impl.mg64.size());
    });
}

int gust_poisson_fill_ghosts(gust_poisson* p, gust_array4* phi, const int32_t* bc, void* stream)
{
    return guarded([&] {
        if (!p || !phi) { GUST_FAIL("gust: PoissonSolver::fill_ghosts: null argument"); }
        PoissonImpl& P = p->impl;
        check_phi(P, phi, "phi");
        check_bc(P, bc);
        fill_phi_ghosts(P, phi, bc_info(bc), (hipStream_t)stream);
    });
}

int gust_poisson_apply(gust_poisson* p, gust_array4* Aphi, gust_array4* phi, const gust_face_coeffs* coeffs,
                       const int32_t* bc, void* stream)
{
    return guarded([&] {
        if (!p || !phi || !Aphi) { GUST_FAIL("gust: PoissonSolver::apply: null argument"); }
        hipStream_t st = (hipStream_t)stream;
        PoissonImpl& P = p->impl;
        check_phi(P, phi, "phi");
        check_cells(P, Aphi, "Aphi");
        check_bc(P, bc);
        const BcInfo b = bc_info(bc);
        DeferredUse dq(st);
        prepare(P, coeffs, b, *dq, st);
        phi_in_out(P, phi, ptr<double>(P.L0.x), 0, st);
        apply_fp64(P, b, ptr<double>(P.L0.x), ptr<double>(P.L0.r), st);
        view_copy_kernel<double><<<blocks_for(P.n()), kThreads, 0, st>>>(
            static_cast<double*>(Aphi->p) + offset_of(*Aphi, P.geom.lo[0], P.geom.lo[1], P.geom.lo[2]),
            Aphi->jstride, Aphi->kstride, ptr<double>(P.L0.r), P.nx, P.ny, P.nz, 1);
        fill_phi_ghosts(P, phi, b, st);
        GUST_HIP(hipGetLastError());
        dq.finish();
    });
}

int gust_poisson_fluxes(gust_poisson* p, gust_array4* phi, const gust_face_coeffs* coeffs, const int32_t* bc,
                        gust_array4* fx, gust_array4* fy, gust_array4* fz, void* stream)
{
    return guarded([&] {
        if (!p || !phi || !fx || !fy || !fz) { GUST_FAIL("gust: PoissonSolver fluxes: null argument"); }
        hipStream_t st = (hipStream_t)stream;
        PoissonImpl& P = p->impl;
        check_phi(P, phi, "phi");
        check_bc(P, bc);
        const BcInfo b = bc_info(bc);
        DeferredUse dq(st);
        prepare(P, coeffs, b, *dq, st);
        phi_in_out(P, phi, ptr<double>(P.L0.x), 0, st);
        const Ctx<double> c = P.L0.ctx(b);
        const MirrorAcc<double> m = mirror<double>(ptr<double>(P.L0.x), P.nx, P.ny, P.nz, P.per, b);
        gust_array4* F[3] = {fx, fy, fz};
        const gust_array4* beta[3] = {coeffs ? coeffs->x : nullptr, coeffs ? coeffs->y : nullptr,
                                      coeffs ? coeffs->z : nullptr};
        for (int d = 0; d < 3; ++d) {
            IBox fb;
            for (int q = 0; q < 3; ++q) { fb.lo[q] = P.geom.lo[q]; fb.hi[q] = P.geom.hi[q] + (q == d); }
            if (!covers(F[d], fb, 1)) {
                GUST_FAIL("gust: PoissonSolver::solve: fluxes must be face MultiFabs on the mesh's box array, "
                          "without ghost cells");
            }
            const double* bp = beta[d] ? static_cast<const double*>(beta[d]->p) +
                                             offset_of(*beta[d], fb.lo[0], fb.lo[1], fb.lo[2])
                                       : nullptr;
            const int64_t ne = int64_t(fb.len(0)) * fb.len(1) * fb.len(2);
            flux_kernel<double><<<blocks_for(ne), kThreads, 0, st>>>(
                d, c, m, bp, beta[d] ? beta[d]->jstride : 0, beta[d] ? beta[d]->kstride : 0,
                static_cast<double*>(F[d]->p) + offset_of(*F[d], fb.lo[0], fb.lo[1], fb.lo[2]), F[d]->jstride,
                F[d]->kstride);
            GUST_HIP(hipGetLastError());
        }
        fill_phi_ghosts(P, phi, b, st);
        dq.finish();
    });
}

int gust_poisson_solve(gust_poisson* p, gust_array4* rhs, gust_array4* phi, const gust_face_coeffs* coeffs,
                       const int32_t* bc, gust_array4* fx, gust_array4* fy, gust_array4* fz,
                       gust_poisson_result* result, void* stream)
{
    return guarded([&] {
        if (!p || !rhs || !phi) { GUST_FAIL("gust: PoissonSolver::solve: null argument"); }
        hipStream_t st = (hipStream_t)stream;
        PoissonImpl& P = p->impl;
        check_phi(P, phi, "phi");
        check_cells(P, rhs, "rhs");
        check_bc(P, bc);
        if ((fx || fy || fz) && !(fx && fy && fz)) {
            GUST_FAIL("gust: PoissonSolver::solve: give all three flux arrays or none");
        }
        for (int d = 0; d < 3 && fx; ++d) {
            gust_array4* F[3] = {fx, fy, fz};
            IBox fb;
            for (int q = 0; q < 3; ++q) { fb.lo[q] = P.geom.lo[q]; fb.hi[q] = P.geom.hi[q] + (q == d); }
            if (!covers(F[d], fb, 1)) {
                GUST_FAIL("gust: PoissonSolver::solve: fluxes must be face MultiFabs on the mesh's box array, "
                          "without ghost cells");
            }
        }
        const BcInfo b = bc_info(bc);
        // Host reads: the coefficient checks and the reported values go with
        // the caller's batch; the initial residual resolves it (one wait),
        // each Arnoldi step reads its dot products with one more.
        DeferredUse dq(st);
        prepare(P, coeffs, b, *dq, st);
        const int64_t n = P.n();
        const int m = std::max(1, P.opt.restart);
        ensure_work(P, m);
        auto acc = std::make_shared<gust_poisson_result>();
        double* bvec = ptr<double>(P.L0.b);
        double* x = ptr<double>(P.L0.x);
        double* r = ptr<double>(P.L0.r);
        double* hv_d = ptr<double>(P.dots);
        double* hv2_d = hv_d + (m + 2);
        double* hn_d = hv2_d + (m + 2);
        double* mean_d = hn_d + 2;
        phi_in_out(P, rhs, bvec, 0, st);
        acc->singular = b.singular;
        if (b.singular) {
            // J is the left null vector: only the mean-free part of rhs is in
            // the range of A (Source/Poisson.cpp:559-567); rhs is modified.
            remove_weighted_mean(P, bvec, mean_d, st);
            unsigned long long* slot = dq->zero_slot([acc](unsigned long long v) { acc->removed_mean = bits_to_double(v); });
            ratio_bits_kernel<<<1, 1, 0, st>>>(mean_d, ptr<double>(P.sum_j), slot);
            phi_in_out(P, rhs, bvec, 1, st);
        }
        if (P.opt.fixed_steps > 0) {
            // GMRES(k) on the device, k = fixed_steps Arnoldi steps whatever
            // the residual: no host reads, so a captured graph can hold the
            // solve; the tolerance is checked at the end (a deferred check).
            const int k = P.opt.fixed_steps;
            ensure_work(P, std::max(m, k));
            P.gm.ensure(size_t(4 + (k + 1) * k + 3 * k + 1) * 8);
            double* hv_k = ptr<double>(P.dots);
            double* hv2_k = hv_k + (k + 2);
            double* hn_k = hv2_k + (k + 2);
            const GmDev gm{ptr<double>(P.gm), k};
            double* Vk = ptr<double>(P.krylov_v);
            double* Zk = ptr<double>(P.krylov_z);
            double* wk = ptr<double>(P.w);
            fill_kernel<double><<<blocks_for(n), kThreads, 0, st>>>(x, 0.0, n);
            multi_dot_dev(P, bvec, 0, 1, bvec, hn_k, st);
            gm_start_kernel<<<1, 1, 0, st>>>(gm, hn_k, P.opt.reltol, P.opt.abstol);
            scale_copy_dev_kernel<<<blocks_for(n), kThreads, 0, st>>>(bvec, gm.g + 2, Vk, n);
            GUST_HIP(hipGetLastError());
            if (P.opt.precision == 1) { setup_hierarchy<float>(P, P.mg32, b, st); }
            else { setup_hierarchy<double>(P, P.mg64, b, st); }
            for (int j = 0; j < k; ++j) {
                double* vj = Vk + int64_t(j) * n;
                double* zj = Zk + int64_t(j) * n;
                if (j > 0) { scale_copy_dev_kernel<<<blocks_for(n), kThreads, 0, st>>>(wk, gm.g + 2, vj, n); }
                precondition(P, vj, zj, b, st);
                apply_fp64(P, b, zj, wk, st);
                multi_dot_dev(P, Vk, n, j + 1, wk, hv_k, st);
                multi_axpy_kernel<<<blocks_for(n), kThreads, 0, st>>>(wk, Vk, n, hv_k, j + 1, -1.0, n);
                multi_dot_dev(P, Vk, n, j + 1, wk, hv2_k, st);
                multi_axpy_kernel<<<blocks_for(n), kThreads, 0, st>>>(wk, Vk, n, hv2_k, j + 1, -1.0, n);
                multi_dot_dev(P, wk, 0, 1, wk, hn_k, st);
                gm_arnoldi_kernel<<<1, 1, 0, st>>>(gm, j, hv_k, hv2_k, hn_k);
                GUST_HIP(hipGetLastError());
            }
            gm_backsolve_kernel<<<1, 1, 0, st>>>(gm, ptr<double>(P.hcoef));
            multi_axpy_kernel<<<blocks_for(n), kThreads, 0, st>>>(x, Zk, n, ptr<double>(P.hcoef), k, 1.0, n);
            GUST_HIP(hipGetLastError());
            acc->iterations = k;
            if (b.singular) { remove_weighted_mean(P, x, mean_d, st); }
            launch_apply<double>(P.L0, b, P.per, x, bvec, r, st);
            multi_dot_dev(P, r, 0, 1, r, hn_k, st);
            unsigned long long* fin = dq->zero_slot([acc](unsigned long long v) { acc->final_residual = bits_to_double(v); });
            unsigned long long* r0b = dq->zero_slot([acc](unsigned long long v) { acc->initial_residual = bits_to_double(v); });
            const double reltol = P.opt.reltol;
            PoissonImpl* impl = &P;
            unsigned long long* miss = dq->zero_slot([acc, k, reltol, impl](unsigned long long v) {
                if (v && impl->soft_miss) {
                    ++impl->misses;
                } else if (v) {
                    GUST_FAIL("gust: the Poisson solve did not converge in its %d fixed GMRES steps (residual %g of %g; "
                              "solver.poisson_reltol = %g)", k, acc->final_residual, acc->initial_residual, reltol);
                }
            });
            gm_check_kernel<<<1, 1, 0, st>>>(gm, hn_k, fin, miss, r0b);
            GUST_HIP(hipGetLastError());
            phi_in_out(P, phi, x, 1, st);
            fill_phi_ghosts(P, phi, b, st);
            if (result) {
                dq->then([acc, result] { *result = *acc; });
            }
            if (fx) { solve_fluxes(P, b, coeffs, x, fx, fy, fz, st); }
            dq.finish();
            return;
        }
        // The multigrid hierarchy is set up at the first preconditioner call:
        // a right-hand side already within tolerance needs none.
        bool hierarchy = false;
        // GMRES(m), right preconditioned, zero initial guess.
        fill_kernel<double><<<blocks_for(n), kThreads, 0, st>>>(x, 0.0, n);
        double* V = ptr<double>(P.krylov_v);
        double* Z = ptr<double>(P.krylov_z);
        double* w = ptr<double>(P.w);
        double r0sq = 0;
        {
            multi_dot_dev(P, bvec, 0, 1, bvec, hn_d, st);
            unsigned long long* slot = dq->zero_slot([&r0sq](unsigned long long v) { r0sq = bits_to_double(v); });
            copy_bits_kernel<<<1, 1, 0, st>>>(hn_d, slot);
            GUST_HIP(hipGetLastError());
            dq->resolve();
        }
        const double r0 = std::sqrt(r0sq);
        acc->initial_residual = r0;
        const double target = std::max(P.opt.reltol * r0, P.opt.abstol);
        int iters = 0;
        bool converged = r0 <= target;
        std::vector<double> H(size_t(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), hv(m + 2), hv2(m + 2);
        GUST_HIP(hipMemcpyAsync(r, bvec, size_t(n) * 8, hipMemcpyDeviceToDevice, st));
        double beta = r0;
        auto* hdots = static_cast<double*>(Workspace::get_host_pinned(1, size_t(2 * (m + 2) + 1) * 8));
        while (!converged && iters < P.opt.max_iter) {
            scale_copy_kernel<<<blocks_for(n), kThreads, 0, st>>>(r, 1.0 / beta, V, n);
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = beta;
            int j = 0;
            double hprev = 0;   // |w| of the previous step: V_j = w / |w| is formed at the start of step j
            for (; j < m && iters < P.opt.max_iter; ++j) {
                double* vj = V + int64_t(j) * n;
                double* zj = Z + int64_t(j) * n;
                if (j > 0 && hprev > 0) { scale_copy_kernel<<<blocks_for(n), kThreads, 0, st>>>(w, 1.0 / hprev, vj, n); }
                if (!hierarchy) {
                    if (P.opt.precision == 1) { setup_hierarchy<float>(P, P.mg32, b, st); }
                    else { setup_hierarchy<double>(P, P.mg64, b, st); }
                    hierarchy = true;
                }
                precondition(P, vj, zj, b, st);
                apply_fp64(P, b, zj, w, st);
                // classical Gram-Schmidt, twice, the coefficients staying on
                // the device; one read of both passes and |w|^2
                multi_dot_dev(P, V, n, j + 1, w, hv_d, st);
                multi_axpy_kernel<<<blocks_for(n), kThreads, 0, st>>>(w, V, n, hv_d, j + 1, -1.0, n);
                multi_dot_dev(P, V, n, j + 1, w, hv2_d, st);
                multi_axpy_kernel<<<blocks_for(n), kThreads, 0, st>>>(w, V, n, hv2_d, j + 1, -1.0, n);
                multi_dot_dev(P, w, 0, 1, w, hn_d, st);
                GUST_HIP(hipMemcpyAsync(hdots, hv_d, size_t(2 * (m + 2) + 1) * 8, hipMemcpyDeviceToHost, st));
                GUST_HIP(hipStreamSynchronize(st));
                for (int i = 0; i <= j; ++i) {
                    hv[i] = hdots[i];
                    hv2[i] = hdots[(m + 2) + i];
                }
                for (int i = 0; i <= j; ++i) { H[size_t(i) + size_t(j) * (m + 1)] = hv[i] + hv2[i]; }
                const double hn = std::sqrt(hdots[2 * (m + 2)]);
                H[size_t(j + 1) + size_t(j) * (m + 1)] = hn;
                hprev = hn;
                // Givens rotations
                for (int i = 0; i < j; ++i) {
                    double& a = H[size_t(i) + size_t(j) * (m + 1)];
                    double& c2 = H[size_t(i + 1) + size_t(j) * (m + 1)];
                    const double t = cs[i] * a + sn[i] * c2;
                    c2 = -sn[i] * a + cs[i] * c2;
                    a = t;
                }
                double& a = H[size_t(j) + size_t(j) * (m + 1)];
                double& c2 = H[size_t(j + 1) + size_t(j) * (m + 1)];
                const double den = std::hypot(a, c2);
                cs[j] = den > 0 ? a / den : 1.0;
                sn[j] = den > 0 ? c2 / den : 0.0;
                a = den;
                c2 = 0.0;
                g[j + 1] = -sn[j] * g[j];
                g[j] = cs[j] * g[j];
                ++iters;
                if (P.opt.verbose > 1) { std::fprintf(stderr, "gust: GMRES %d: |r| ~ %.6e\n", iters, std::fabs(g[j + 1])); }
                if (std::fabs(g[j + 1]) <= target || hn == 0) {
                    ++j;
                    break;
                }
            }
            // y = H^{-1} g, x += Z y
            for (int i = j - 1; i >= 0; --i) {
                double s = g[i];
                for (int q = i + 1; q < j; ++q) { s -= H[size_t(i) + size_t(q) * (m + 1)] * y[q]; }
                y[i] = s / H[size_t(i) + size_t(i) * (m + 1)];
            }
            set_coef(P, y.data(), j, st);
            multi_axpy_kernel<<<blocks_for(n), kThreads, 0, st>>>(x, Z, n, ptr<double>(P.hcoef), j, 1.0, n);
            // true residual
            launch_apply<double>(P.L0, b, P.per, x, bvec, r, st);
            beta = norm2(P, r, st);
            converged = beta <= target;
        }
        acc->iterations = iters;
        if (b.singular) { remove_weighted_mean(P, x, mean_d, st); }
        launch_apply<double>(P.L0, b, P.per, x, bvec, r, st);
        {
            multi_dot_dev(P, r, 0, 1, r, hn_d, st);
            unsigned long long* slot = dq->zero_slot([acc](unsigned long long v) {
                acc->final_residual = std::sqrt(bits_to_double(v));
            });
            copy_bits_kernel<<<1, 1, 0, st>>>(hn_d, slot);
            GUST_HIP(hipGetLastError());
        }
        phi_in_out(P, phi, x, 1, st);
        fill_phi_ghosts(P, phi, b, st);
        if (result) {
            dq->then([acc, result] { *result = *acc; });
        }
        if (!converged) {
            dq->resolve();
            const int max_iter = P.opt.max_iter;
            const double reltol = P.opt.reltol, abstol = P.opt.abstol;
            GUST_FAIL("gust: the Poisson solve did not converge in solver.poisson_max_iter = %d iterations (residual "
                      "%g of %g; solver.poisson_reltol = %g, solver.poisson_abstol = %g)",
                      max_iter, acc->final_residual, acc->initial_residual, reltol, abstol);
        }
        if (fx) { solve_fluxes(P, b, coeffs, x, fx, fy, fz, st); }
        dq.finish();
    });
}

int gust_poisson_vcycle(gust_poisson* p, const double* b, double* x, const gust_face_coeffs* coeffs,
                        const int32_t* bc, int setup, void* stream)
{
    return guarded([&] {
        if (!p || !b || !x) { GUST_FAIL("gust: poisson_vcycle: null argument"); }
        hipStream_t st = (hipStream_t)stream;
        PoissonImpl& P = p->impl;
        check_bc(P, bc);
        const BcInfo bi = bc_info(bc);
        DeferredUse dq(st);
        if (setup) {
            prepare(P, coeffs, bi, *dq, st);
            if (P.opt.precision == 1) { setup_hierarchy<float>(P, P.mg32, bi, st); }
            else { setup_hierarchy<double>(P, P.mg64, bi, st); }
        }
        precondition(P, b, x, bi, st);
        dq.finish();
    });
}

int gust_face_divergence(const gust_geometry* geom, const gust_array4* ax, const gust_array4* ay,
                         const gust_array4* az, const gust_array4* detJ, const gust_array4* fx,
                         const gust_array4* fy, const gust_array4* fz, gust_array4* div, void* stream)
{
    return guarded([&] {
        if (!geom || !ax || !ay || !az || !detJ || !fx || !fy || !fz || !div) {
            GUST_FAIL("gust: face_divergence: null argument");
        }
        IBox cells;
        for (int q = 0; q < 3; ++q) { cells.lo[q] = geom->lo[q]; cells.hi[q] = geom->hi[q]; }
        const gust_array4* F[3] = {fx, fy, fz};
        const gust_array4* A[3] = {ax, ay, az};
        for (int d = 0; d < 3; ++d) {
            IBox fb = cells;
            fb.hi[d] += 1;
            if (!covers(F[d], fb, 1) || !covers(A[d], fb, 1)) {
                GUST_FAIL("gust: face_divergence: F must be face MultiFabs on the mesh's box array");
            }
        }
        if (!covers(div, cells, 1) || !covers(detJ, cells, 1)) {
            GUST_FAIL("gust: face_divergence: div must be on the mesh's box array");
        }
        const int nx = cells.len(0), ny = cells.len(1), nz = cells.len(2);
        auto P0 = [&](const gust_array4* a) {
            return static_cast<const double*>(a->p) + offset_of(*a, cells.lo[0], cells.lo[1], cells.lo[2]);
        };
        face_div_kernel<<<blocks_for(int64_t(nx) * ny * nz), kThreads, 0, (hipStream_t)stream>>>(
            nx, ny, nz, 1.0 / geom->dx[0], 1.0 / geom->dx[1], 1.0 / geom->dx[2], P0(ax), ax->jstride, ax->kstride,
            P0(ay), ay->jstride, ay->kstride, P0(az), az->jstride, az->kstride, P0(fx), fx->jstride, fx->kstride,
            P0(fy), fy->jstride, fy->kstride, P0(fz), fz->jstride, fz->kstride, P0(detJ), detJ->jstride,
            detJ->kstride,
            static_cast<double*>(div->p) + offset_of(*div, cells.lo[0], cells.lo[1], cells.lo[2]), div->jstride,
            div->kstride);
        GUST_HIP(hipGetLastError());
    });
}

}
