// terrain_stencil.hpp -- the terrain-following Poisson stencil on the device.
// Derived in part from ERF (BSD 3-Clause, Copyright (c) 2022 The Regents of the University of California
// et al.); the full notice is in THIRD_PARTY_NOTICES.md at the repository root.
//
// The face gradients of a cell-centred field on the terrain-following mesh
// (the physical d/dx, d/dy with the terrain cross terms, and the z-face
// normal gradient), written once and instantiated twice:
//   * V = R (double or float): the operator apply, expression by expression
//     in the order of extern_erf/ERF_TerrainPoisson_3D_K.H:22-196 and
//     318-530, so FP64 results equal the CPU build's bit for bit;
//   * V = Lin3<R>: the same expressions on values that carry their linear
//     dependence on the cell's own column (k-1, k, k+1), which yields the
//     exact tridiagonal part of the operator for the z-line smoother without
//     a second hand-derived stencil.
// The accessors X (field, ghosts filled) and Z (node heights, one ghost node)
// are small functors so one body serves every layout and level.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gust_amd::stencil {

// A value linear in the own column's x(k-1), x(k), x(k+1): c[0..2].
template <typename R>
struct Lin3 {
    R c[3];
    __device__ __forceinline__ Lin3() : c{R(0), R(0), R(0)} {}
    __device__ __forceinline__ explicit Lin3(R v) : c{v, v, v} {}
};
template <typename R>
__device__ __forceinline__ Lin3<R> operator+(Lin3<R> a, const Lin3<R>& b)
{
    for (int q = 0; q < 3; ++q) { a.c[q] += b.c[q]; }
    return a;
}
template <typename R>
__device__ __forceinline__ Lin3<R> operator-(Lin3<R> a, const Lin3<R>& b)
{
    for (int q = 0; q < 3; ++q) { a.c[q] -= b.c[q]; }
    return a;
}
template <typename R>
__device__ __forceinline__ Lin3<R> operator-(Lin3<R> a)
{
    for (int q = 0; q < 3; ++q) { a.c[q] = -a.c[q]; }
    return a;
}
template <typename R>
__device__ __forceinline__ Lin3<R> operator*(Lin3<R> a, R s)
{
    for (int q = 0; q < 3; ++q) { a.c[q] *= s; }
    return a;
}
template <typename R>
__device__ __forceinline__ Lin3<R> operator*(R s, Lin3<R> a)
{
    for (int q = 0; q < 3; ++q) { a.c[q] = s * a.c[q]; }
    return a;
}
template <typename R>
__device__ __forceinline__ Lin3<R> operator/(Lin3<R> a, R s)
{
    for (int q = 0; q < 3; ++q) { a.c[q] /= s; }
    return a;
}
template <typename R>
__device__ __forceinline__ Lin3<R>& operator*=(Lin3<R>& a, R s)
{
    for (int q = 0; q < 3; ++q) { a.c[q] *= s; }
    return a;
}
template <typename R>
__device__ __forceinline__ Lin3<R>& operator-=(Lin3<R>& a, const Lin3<R>& b)
{
    for (int q = 0; q < 3; ++q) { a.c[q] -= b.c[q]; }
    return a;
}

template <typename V>
__device__ __forceinline__ V zero_of() { return V(0); }

// Face-mask bits (extern_erf/ERF_TerrainPoisson_3D_K.H:281-285).
enum : int { kXlo = 1, kXhi = 2, kYlo = 4, kYhi = 8, kZlo = 16, kZhi = 32 };

// The mesh-only factors of the stencil, from the node heights zp (one ghost
// node) on the fly (ZMetric) or from arrays filled once per mesh
// (ArrayMetric). Every division of the operator is one of these, so the
// array form removes them from the per-apply work while the stored values
// are the same quotients, bit for bit.
//   ry(i, j, k): h_xi / h_zeta on the y-edge at node x i, cell y j, node z k
//   rx(i, j, k): h_eta / h_zeta on the x-edge at cell x i, node y j, node z k
//   hzi(i, j, k): 8 / (column depth over k-1..k+1) on the z-face (i, j, k)
//   hxz, hez(i, j, k): the face slopes h_xi, h_eta on that z-face
template <typename R, class Z>
struct ZMetric {
    Z zp;
    R dxinv, dyinv;
    __device__ __forceinline__ R ry(int i, int j, int k) const
    {
        const R fourth = R(0.25);
        const R h_xi = fourth * (zp(i + 1, j, k) - zp(i - 1, j, k) + zp(i + 1, j + 1, k) - zp(i - 1, j + 1, k)) * dxinv;
        const R h_zeta = fourth * (zp(i, j, k + 1) - zp(i, j, k - 1) + zp(i, j + 1, k + 1) - zp(i, j + 1, k - 1));
        return h_xi / h_zeta;
    }
    __device__ __forceinline__ R rx(int i, int j, int k) const
    {
        const R fourth = R(0.25);
        const R h_eta = fourth * (zp(i, j + 1, k) - zp(i, j - 1, k) + zp(i + 1, j + 1, k) - zp(i + 1, j - 1, k)) * dyinv;
        const R h_zeta = fourth * (zp(i, j, k + 1) - zp(i, j, k - 1) + zp(i + 1, j, k + 1) - zp(i + 1, j, k - 1));
        return h_eta / h_zeta;
    }
    __device__ __forceinline__ R hzi(int i, int j, int k) const
    {
        return R(8.0) / ((zp(i, j, k + 1) + zp(i + 1, j, k + 1) + zp(i, j + 1, k + 1) + zp(i + 1, j + 1, k + 1))
                         - (zp(i, j, k - 1) + zp(i + 1, j, k - 1) + zp(i, j + 1, k - 1) + zp(i + 1, j + 1, k - 1)));
    }
    __device__ __forceinline__ R hxz(int i, int j, int k) const
    {
        return R(0.5) * (zp(i + 1, j + 1, k) + zp(i + 1, j, k) - zp(i, j + 1, k) - zp(i, j, k)) * dxinv;
    }
    __device__ __forceinline__ R hez(int i, int j, int k) const
    {
        return R(0.5) * (zp(i + 1, j + 1, k) + zp(i, j + 1, k) - zp(i + 1, j, k) - zp(i, j, k)) * dyinv;
    }
};

template <typename R>
struct ArrayMetric {
    const R* ryp;   // (nx+1) x ny x (nz+1)
    const R* rxp;   // nx x (ny+1) x (nz+1)
    const R* zfp;   // 3 x nx x ny x (nz+1): hzi, hxz, hez
    int nx, ny, nz;
    __device__ __forceinline__ R ry(int i, int j, int k) const
    {
        return ryp[i + int64_t(j) * (nx + 1) + int64_t(k) * (nx + 1) * ny];
    }
    __device__ __forceinline__ R rx(int i, int j, int k) const
    {
        return rxp[i + int64_t(j) * nx + int64_t(k) * nx * (ny + 1)];
    }
    __device__ __forceinline__ R zf(int c, int i, int j, int k) const
    {
        return zfp[i + int64_t(j) * nx + int64_t(k) * nx * ny + int64_t(c) * nx * ny * (nz + 1)];
    }
    __device__ __forceinline__ R hzi(int i, int j, int k) const { return zf(0, i, j, k); }
    __device__ __forceinline__ R hxz(int i, int j, int k) const { return zf(1, i, j, k); }
    __device__ __forceinline__ R hez(int i, int j, int k) const { return zf(2, i, j, k); }
};

// The low x-face gradient of cell (i, j, k) (ERF's px_lo; flux_x = -this).
template <typename V, typename R, class X, class Mt>
__device__ __forceinline__ V grad_x(const X& x, const Mt& mt, int i, int j, int k, R dxinv)
{
    const R half = R(0.5);
    V px_lo = (x(i, j, k) - x(i - 1, j, k)) * dxinv;
    V pz_lo_md_hi = half * (x(i, j, k + 1) + x(i - 1, j, k + 1) - x(i, j, k) - x(i - 1, j, k));
    pz_lo_md_hi *= mt.ry(i, j, k + 1);
    V pz_lo_md_lo = half * (x(i, j, k) + x(i - 1, j, k) - x(i, j, k - 1) - x(i - 1, j, k - 1));
    pz_lo_md_lo *= mt.ry(i, j, k);
    px_lo -= half * (pz_lo_md_hi + pz_lo_md_lo);
    return px_lo;
}

template <typename V, typename R, class X, class Mt>
__device__ __forceinline__ V grad_y(const X& x, const Mt& mt, int i, int j, int k, R dyinv)
{
    const R half = R(0.5);
    V py_lo = (x(i, j, k) - x(i, j - 1, k)) * dyinv;
    V pz_md_lo_hi = half * (x(i, j, k + 1) + x(i, j - 1, k + 1) - x(i, j, k) - x(i, j - 1, k));
    pz_md_lo_hi *= mt.rx(i, j, k + 1);
    V pz_md_lo_lo = half * (x(i, j, k) + x(i, j - 1, k) - x(i, j, k - 1) - x(i, j - 1, k - 1));
    pz_md_lo_lo *= mt.rx(i, j, k);
    py_lo -= half * (pz_md_lo_hi + pz_md_lo_lo);
    return py_lo;
}

// The low z-face gradient of cell (i, j, k) (ERF's pz_lo before the ground
// zeroing).
template <typename V, typename R, class X, class Mt>
__device__ __forceinline__ V grad_z(const X& x, const Mt& mt, int i, int j, int k, R dxinv, R dyinv)
{
    const R half = R(0.5);
    V pz_lo = (x(i, j, k) - x(i, j, k - 1));
    pz_lo *= mt.hzi(i, j, k);
    const V px_hi_md_lo = half * (x(i + 1, j, k) - x(i, j, k) + x(i + 1, j, k - 1) - x(i, j, k - 1)) * dxinv;
    const V px_lo_md_lo = half * (x(i, j, k) - x(i - 1, j, k) + x(i, j, k - 1) - x(i - 1, j, k - 1)) * dxinv;
    const V py_md_hi_lo = half * (x(i, j + 1, k) - x(i, j, k) + x(i, j + 1, k - 1) - x(i, j, k - 1)) * dyinv;
    const V py_md_lo_lo = half * (x(i, j, k) - x(i, j - 1, k) + x(i, j, k - 1) - x(i, j - 1, k - 1)) * dyinv;
    V pz_hi_md_lo = half * (x(i + 1, j, k) + x(i, j, k) - x(i + 1, j, k - 1) - x(i, j, k - 1));
    pz_hi_md_lo *= mt.ry(i + 1, j, k);
    V pz_lo_md_lo = half * (x(i, j, k) + x(i - 1, j, k) - x(i, j, k - 1) - x(i - 1, j, k - 1));
    pz_lo_md_lo *= mt.ry(i, j, k);
    V pz_md_hi_lo = half * (x(i, j + 1, k) + x(i, j, k) - x(i, j + 1, k - 1) - x(i, j, k - 1));
    pz_md_hi_lo *= mt.rx(i, j + 1, k);
    V pz_md_lo_lo = half * (x(i, j, k) + x(i, j - 1, k) - x(i, j, k - 1) - x(i, j - 1, k - 1));
    pz_md_lo_lo *= mt.rx(i, j, k);
    const R h_xi_on_zlo = mt.hxz(i, j, k);
    const R h_eta_on_zlo = mt.hez(i, j, k);
    pz_lo -= half * h_xi_on_zlo * ((px_hi_md_lo + px_lo_md_lo) - (pz_hi_md_lo + pz_lo_md_lo));
    pz_lo -= half * h_eta_on_zlo * ((py_md_hi_lo + py_md_lo_lo) - (pz_md_hi_lo + pz_md_lo_lo));
    return pz_lo;
}

// The one-sided ground gradient with phi = 0 on the ground
// (extern_erf/ERF_TerrainPoisson_3D_K.H:212-275; flux = -this).
template <typename V, typename R, class X, class Z>
__device__ __forceinline__ V grad_z_dir(const X& x, const Z& zp, int i, int j, int k, R dxinv, R dyinv)
{
    const R half = R(0.5), fourth = R(0.25);
    R h_xi, h_eta, h_zeta;
    V pz_lo = x(i, j, k);
    const R hzeta_inv_on_zlo = R(8.0) / ((zp(i, j, k + 1) + zp(i + 1, j, k + 1) + zp(i, j + 1, k + 1)
                                          + zp(i + 1, j + 1, k + 1))
                                         - (zp(i, j, k) + zp(i + 1, j, k) + zp(i, j + 1, k) + zp(i + 1, j + 1, k)));
    pz_lo *= hzeta_inv_on_zlo;
    const V zero = zero_of<V>();
    const V px_hi_md_lo = zero, px_lo_md_lo = zero, py_md_hi_lo = zero, py_md_lo_lo = zero;
    V pz_hi_md_lo = half * (x(i + 1, j, k) + x(i, j, k));
    h_xi = fourth * (zp(i + 2, j, k) - zp(i, j, k) + zp(i + 2, j + 1, k) - zp(i, j + 1, k)) * dxinv;
    h_zeta = fourth * (zp(i + 1, j, k + 1) - zp(i + 1, j, k) + zp(i + 1, j + 1, k + 1) - zp(i + 1, j + 1, k));
    pz_hi_md_lo *= h_xi / h_zeta;
    V pz_lo_md_lo = half * (x(i, j, k) + x(i - 1, j, k));
    h_xi = fourth * (zp(i + 1, j, k) - zp(i - 1, j, k) + zp(i + 1, j + 1, k) - zp(i - 1, j + 1, k)) * dxinv;
    h_zeta = fourth * (zp(i, j, k + 1) - zp(i, j, k) + zp(i, j + 1, k + 1) - zp(i, j + 1, k));
    pz_lo_md_lo *= h_xi / h_zeta;
    V pz_md_hi_lo = half * (x(i, j + 1, k) + x(i, j, k));
    h_eta = fourth * (zp(i, j + 2, k) - zp(i, j, k) + zp(i + 1, j + 2, k) - zp(i + 1, j, k)) * dyinv;
    h_zeta = fourth * (zp(i, j + 1, k + 1) - zp(i, j + 1, k) + zp(i + 1, j + 1, k + 1) - zp(i + 1, j + 1, k));
    pz_md_hi_lo *= h_eta / h_zeta;
    V pz_md_lo_lo = half * (x(i, j, k) + x(i, j - 1, k));
    h_eta = fourth * (zp(i, j + 1, k) - zp(i, j - 1, k) + zp(i + 1, j + 1, k) - zp(i + 1, j - 1, k)) * dyinv;
    h_zeta = fourth * (zp(i, j, k + 1) - zp(i, j, k) + zp(i + 1, j, k + 1) - zp(i + 1, j, k));
    pz_md_lo_lo *= h_eta / h_zeta;
    const R h_xi_on_zlo = half * (zp(i + 1, j + 1, k) + zp(i + 1, j, k) - zp(i, j + 1, k) - zp(i, j, k)) * dxinv;
    const R h_eta_on_zlo = half * (zp(i + 1, j + 1, k) + zp(i, j + 1, k) - zp(i + 1, j, k) - zp(i, j, k)) * dyinv;
    pz_lo -= half * h_xi_on_zlo * ((px_hi_md_lo + px_lo_md_lo) - (pz_hi_md_lo + pz_lo_md_lo));
    pz_lo -= half * h_eta_on_zlo * ((py_md_hi_lo + py_md_lo_lo) - (pz_md_hi_lo + pz_md_lo_lo));
    return pz_lo;
}

// The terrain cross part of the z-face gradient (Source/PoissonCross_K.H:16-58).
template <typename V, typename R, class X, class Mt>
__device__ __forceinline__ V z_face_cross(const X& x, const Mt& mt, int i, int j, int k, R dxinv, R dyinv)
{
    const R half = R(0.5);
    const V px_hi_md_lo = half * (x(i + 1, j, k) - x(i, j, k) + x(i + 1, j, k - 1) - x(i, j, k - 1)) * dxinv;
    const V px_lo_md_lo = half * (x(i, j, k) - x(i - 1, j, k) + x(i, j, k - 1) - x(i - 1, j, k - 1)) * dxinv;
    const V py_md_hi_lo = half * (x(i, j + 1, k) - x(i, j, k) + x(i, j + 1, k - 1) - x(i, j, k - 1)) * dyinv;
    const V py_md_lo_lo = half * (x(i, j, k) - x(i, j - 1, k) + x(i, j, k - 1) - x(i, j - 1, k - 1)) * dyinv;
    V pz_hi_md_lo = half * (x(i + 1, j, k) + x(i, j, k) - x(i + 1, j, k - 1) - x(i, j, k - 1));
    pz_hi_md_lo *= mt.ry(i + 1, j, k);
    V pz_lo_md_lo = half * (x(i, j, k) + x(i - 1, j, k) - x(i, j, k - 1) - x(i - 1, j, k - 1));
    pz_lo_md_lo *= mt.ry(i, j, k);
    V pz_md_hi_lo = half * (x(i, j + 1, k) + x(i, j, k) - x(i, j + 1, k - 1) - x(i, j, k - 1));
    pz_md_hi_lo *= mt.rx(i, j + 1, k);
    V pz_md_lo_lo = half * (x(i, j, k) + x(i, j - 1, k) - x(i, j, k - 1) - x(i, j - 1, k - 1));
    pz_md_lo_lo *= mt.rx(i, j, k);
    const R h_xi_on_zlo = mt.hxz(i, j, k);
    const R h_eta_on_zlo = mt.hez(i, j, k);
    return half * h_xi_on_zlo * ((px_hi_md_lo + px_lo_md_lo) - (pz_hi_md_lo + pz_lo_md_lo))
         + half * h_eta_on_zlo * ((py_md_hi_lo + py_md_lo_lo) - (pz_md_hi_lo + pz_md_lo_lo));
}

// Everything the operator at one cell needs besides the field.
template <typename R, class Z, class F, class Mt>
struct OpCtx {
    Z zp;                 // node heights (one ghost node; the one-sided ground face reads them)
    Mt mt;                // the mesh-only factors
    F ax, ay, az, J;      // premultiplied face areas (valid faces), cell J
    F dcross, az_mesh;    // z_cross correction (beta_z - z_cross) and mesh az; dcross.p null: none
    R dxinv, dyinv, dzinv;
    int nx, ny, nz;       // cells
    int zf;               // zero-flux face mask
    int dir_ground;       // Dirichlet ground: the one-sided ground flux
    int top_open;         // the top carries flux (not Neumann)
};

// (A x)(i, j, k): extern_erf/ERF_TerrainPoisson_3D_K.H:318-530 plus the
// Dirichlet ground face (Source/Poisson.cpp:162-190) and the z_cross
// correction (Source/Poisson.cpp:231-254).
template <typename V, typename R, class X, class Z, class F, class Mt>
__device__ __forceinline__ V apply_cell(const X& x, const OpCtx<R, Z, F, Mt>& c, int i, int j, int k)
{
    V px_hi = grad_x<V, R>(x, c.mt, i + 1, j, k, c.dxinv);
    V px_lo = grad_x<V, R>(x, c.mt, i, j, k, c.dxinv);
    V py_hi = grad_y<V, R>(x, c.mt, i, j + 1, k, c.dyinv);
    V py_lo = grad_y<V, R>(x, c.mt, i, j, k, c.dyinv);
    V pz_hi = grad_z<V, R>(x, c.mt, i, j, k + 1, c.dxinv, c.dyinv);
    V pz_lo = grad_z<V, R>(x, c.mt, i, j, k, c.dxinv, c.dyinv);
    const V zero = zero_of<V>();
    if (k == 0) { pz_lo = zero; }
    if ((c.zf & kXlo) && i == 0) { px_lo = zero; }
    if ((c.zf & kXhi) && i == c.nx - 1) { px_hi = zero; }
    if ((c.zf & kYlo) && j == 0) { py_lo = zero; }
    if ((c.zf & kYhi) && j == c.ny - 1) { py_hi = zero; }
    if ((c.zf & kZhi) && k == c.nz - 1) { pz_hi = zero; }
    V y = (c.ax(i + 1, j, k) * px_hi - c.ax(i, j, k) * px_lo) * c.dxinv
          + (c.ay(i, j + 1, k) * py_hi - c.ay(i, j, k) * py_lo) * c.dyinv
          + (c.az(i, j, k + 1) * pz_hi - c.az(i, j, k) * pz_lo) * c.dzinv;
    const R Jc = c.J(i, j, k);
    y = y / Jc;
    if (c.dir_ground && k == 0) {
        const V f = -grad_z_dir<V, R>(x, c.zp, i, j, k, c.dxinv, c.dyinv);
        y = y + c.az(i, j, k) * f * c.dzinv / Jc;
    }
    if (c.dcross.p) {
        V fhi = zero, flo = zero;
        if (!(k + 1 == c.nz && !c.top_open)) {
            fhi = c.az_mesh(i, j, k + 1) * c.dcross(i, j, k + 1)
                  * z_face_cross<V, R>(x, c.mt, i, j, k + 1, c.dxinv, c.dyinv);
        }
        if (!(k == 0)) {
            flo = c.az_mesh(i, j, k) * c.dcross(i, j, k) * z_face_cross<V, R>(x, c.mt, i, j, k, c.dxinv, c.dyinv);
        }
        y = y + (fhi - flo) * c.dzinv / Jc;
    }
    return y;
}

}   // namespace gust_amd::stencil
