/*
 * stencil.h -- the oracle's restatement of Source/Stencil_K.H:23-140: second-
 * order interpolation and index differences between the staggered locations
 * of the C-grid, with the location types as run-time bit masks (bit d: nodal
 * along d). The reference evaluates the same loops with compile-time bounds;
 * the offsets, loop order (z outer, then y, then x), the accumulation from 0
 * and the final scale by 1 / (number of points) are the same, so the values
 * are the same bits.
 *
 * TEST INFRASTRUCTURE ONLY (see gust_oracle.h).
 */
#ifndef GUST_ORACLE_STENCIL_H_
#define GUST_ORACLE_STENCIL_H_

#include "gust_oracle.h"

#define OAT(v, i, j, k)                                                                                            \
    ((v)->p[((int64_t)(i) - (v)->begin[0]) + ((int64_t)(j) - (v)->begin[1]) * (v)->jstride                        \
            + ((int64_t)(k) - (v)->begin[2]) * (v)->kstride])

enum { OS_CC = 0, OS_XF = 1, OS_YF = 2, OS_ZF = 4, OS_NODE = 7, OS_NONE = 3 };

static inline int os_nodal(int t, int d) { return (t >> d) & 1; }
static inline int os_lo_off(int tq, int ts, int d)
{
    return (os_nodal(tq, d) == os_nodal(ts, d) || os_nodal(tq, d)) ? 0 : -1;
}
static inline int os_npts(int tq, int ts, int d) { return os_nodal(tq, d) == os_nodal(ts, d) ? 1 : 2; }

/* interp<TQ, TS> (Stencil_K.H:65-70). */
static inline double os_interp(const oa4* q, int tq, int ts, int i, int j, int k)
{
    const int lx = os_lo_off(tq, ts, 0), ly = os_lo_off(tq, ts, 1), lz = os_lo_off(tq, ts, 2);
    const int nx = os_npts(tq, ts, 0), ny = os_npts(tq, ts, 1), nz = os_npts(tq, ts, 2);
    double s = 0;
    for (int c = 0; c < nz; ++c)
        for (int b = 0; b < ny; ++b)
            for (int a = 0; a < nx; ++a) { s += OAT(q, i + lx + a, j + ly + b, k + lz + c); }
    return s * (1.0 / (double)(nx * ny * nz));
}

/* deriv<TQ, TS, M> per unit index (Stencil_K.H:72-91). */
static inline double os_deriv(const oa4* q, int tq, int ts, int M, int i, int j, int k)
{
    const int same = os_nodal(tq, M) == os_nodal(ts, M);
    const int m0 = same ? -1 : os_lo_off(tq, ts, M);
    const int m1 = same ? 1 : m0 + 1;
    const double w = same ? 0.5 : 1.0;
    const int ex = M == 0, ey = M == 1, ez = M == 2;
    const int lx = M == 0 ? 0 : os_lo_off(tq, ts, 0), ly = M == 1 ? 0 : os_lo_off(tq, ts, 1),
              lz = M == 2 ? 0 : os_lo_off(tq, ts, 2);
    const int nx = M == 0 ? 1 : os_npts(tq, ts, 0), ny = M == 1 ? 1 : os_npts(tq, ts, 1),
              nz = M == 2 ? 1 : os_npts(tq, ts, 2);
    double s = 0;
    for (int c = 0; c < nz; ++c)
        for (int b = 0; b < ny; ++b)
            for (int a = 0; a < nx; ++a) {
                const int ii = i + lx + a, jj = j + ly + b, kk = k + lz + c;
                s += OAT(q, ii + m1 * ex, jj + m1 * ey, kk + m1 * ez) - OAT(q, ii + m0 * ex, jj + m0 * ey, kk + m0 * ez);
            }
    return w * (s * (1.0 / (double)(nx * ny * nz)));
}

/* metric_at<TS> (Stencil_K.H:103-109). */
typedef struct { double zx, zy, zz; } os_metric;
static inline os_metric os_metric_at(const oa4* z, int ts, int i, int j, int k, const double dxinv[3])
{
    os_metric m;
    m.zx = os_deriv(z, OS_NODE, ts, 0, i, j, k) * dxinv[0];
    m.zy = os_deriv(z, OS_NODE, ts, 1, i, j, k) * dxinv[1];
    m.zz = os_deriv(z, OS_NODE, ts, 2, i, j, k) * dxinv[2];
    return m;
}

/* flux_lo (Stencil_K.H:133-136). */
static inline int os_flux_lo(int ts, int d) { return os_nodal(ts, d) ? -1 : 0; }

/* std::max / std::min as the reference's amrex::max / amrex::min. */
static inline double os_max(double a, double b) { return (a < b) ? b : a; }
static inline double os_min(double a, double b) { return (b < a) ? b : a; }

#endif
