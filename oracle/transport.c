/*
 * transport.c -- oracle of the convection and diffusion operators over the
 * control volumes of the C-grid.
 *
 * TEST INFRASTRUCTURE ONLY (see gust_oracle.h).
 *
 *   Source/CVFluxes.cpp:16-39, 54-67, 69-80    divergence, cv_flux_box, flux_divergence
 *   Source/Limiter_K.H:23-69                   limited_slope, muscl_face
 *   Source/Convection.cpp:16-40                flux_kernel (convective fluxes)
 *   Source/Convection.cpp:60-84                transport_divergence (bounded form)
 *   Source/Diffusion.cpp:10-59                 scalar_kernel, momentum_kernel
 *   Source/Diffusion_K.H:66-100                area_dot, stress_row
 *   Source/Stencil_K.H:111-127                 grad_at
 *
 * Flux arrays G[d] are views at the field's control-volume face location
 * along d (loc ^ (1 << d)) covering cv_flux_box(loc, valid, d).
 */
#include "stencil.h"

#include <float.h>
#include <math.h>

enum { OLIM_VAN_LEER = 0, OLIM_VAN_ALBADA = 1, OLIM_UPWIND = 2 };

static double limited_slope(int L, double a, double b, double eps2)
{
    if (L == OLIM_VAN_LEER) {
        const double ab = a * b;
        return (ab > 0) ? 2.0 * ab / (a + b) : 0.0;
    }
    if (L == OLIM_VAN_ALBADA) { return ((b * b + eps2) * a + (a * a + eps2) * b) / (a * a + b * b + 2.0 * eps2); }
    return 0.0;
}

static double muscl_face(int L, double tr, double qm, double q0, double q1, double qp, double eps2)
{
    if (L == OLIM_UPWIND) { return (tr >= 0) ? q0 : q1; }
    const double half = 0.5;
    return (tr >= 0) ? q0 + half * limited_slope(L, q0 - qm, q1 - q0, eps2)
                     : q1 - half * limited_slope(L, q1 - q0, qp - q1, eps2);
}

/* The index range of cv_flux_box(loc, valid, d) along axis a (inclusive). */
static void cv_range(int T, int d, const ogeom* g, int a, int* lo, int* hi)
{
    const int nod = os_nodal(T, a);
    *lo = g->lo[a];
    *hi = g->hi[a] + nod;
    if (a == d) {
        if (nod) {   /* points on d-nodes [lo, hi+1]: fluxes at centres [lo-1, hi+1] */
            *lo = g->lo[a] - 1;
            *hi = g->hi[a] + 1;
        } else {     /* points at d-centres [lo, hi]: fluxes on d-nodes [lo, hi+1] */
            *lo = g->lo[a];
            *hi = g->hi[a] + 1;
        }
    }
}

int oracle_convective_fluxes(int loc, const oa4* q, const oa4* fx, const oa4* fy, const oa4* fz, int limiter,
                             double eps, const ogeom* g, oa4* G0, oa4* G1, oa4* G2)
{
    if (!(loc == 0 || loc == 1 || loc == 2 || loc == 4) || limiter < 0 || limiter > 2) { return -1; }
    if (isnan(eps) || !(eps >= sqrt(DBL_MIN) && eps <= sqrt(DBL_MAX))) { return -2; }
    const double eps2 = eps * eps;
    const int T = loc;
    const oa4* F[3] = {fx, fy, fz};
    oa4* G[3] = {G0, G1, G2};
    for (int D = 0; D < 3; ++D) {
        const int S = T ^ (1 << D), TF = 1 << D;
        const int l = os_lo_off(T, S, D);
        const int ex = D == 0, ey = D == 1, ez = D == 2;
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) { cv_range(T, D, g, a, &lo[a], &hi[a]); }
        for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int i = lo[0]; i <= hi[0]; ++i) {
                    const double tr = os_interp(F[D], TF, S, i, j, k);
                    const int i0 = i + l * ex, j0 = j + l * ey, k0 = k + l * ez;
                    OAT(G[D], i, j, k) =
                        tr * muscl_face(limiter, tr, OAT(q, i0 - ex, j0 - ey, k0 - ez), OAT(q, i0, j0, k0),
                                        OAT(q, i0 + ex, j0 + ey, k0 + ez), OAT(q, i0 + 2 * ex, j0 + 2 * ey, k0 + 2 * ez),
                                        eps2);
                }
    }
    return 0;
}

int oracle_flux_divergence(int loc, const oa4* G0, const oa4* G1, const oa4* G2, const oa4* J, const ogeom* g,
                           oa4* out)
{
    if (!(loc == 0 || loc == 1 || loc == 2 || loc == 4)) { return -1; }
    const int T = loc;
    const int lx = os_flux_lo(T, 0), ly = os_flux_lo(T, 1), lz = os_flux_lo(T, 2);
    const double dxinv[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
    for (int k = g->lo[2]; k <= g->hi[2] + os_nodal(T, 2); ++k)
        for (int j = g->lo[1]; j <= g->hi[1] + os_nodal(T, 1); ++j)
            for (int i = g->lo[0]; i <= g->hi[0] + os_nodal(T, 0); ++i) {
                OAT(out, i, j, k) = ((OAT(G0, i + lx + 1, j, k) - OAT(G0, i + lx, j, k)) * dxinv[0]
                                     + (OAT(G1, i, j + ly + 1, k) - OAT(G1, i, j + ly, k)) * dxinv[1]
                                     + (OAT(G2, i, j, k + lz + 1) - OAT(G2, i, j, k + lz)) * dxinv[2])
                                    / os_interp(J, OS_CC, T, i, j, k);
            }
    return 0;
}

int oracle_subtract_transport_divergence(int loc, const oa4* q, const oa4* fx, const oa4* fy, const oa4* fz,
                                         const oa4* J, const ogeom* g, oa4* out)
{
    if (!(loc == 0 || loc == 1 || loc == 2 || loc == 4)) { return -1; }
    const int T = loc;
    const int SX = T ^ OS_XF, SY = T ^ OS_YF, SZ = T ^ OS_ZF;
    const int lx = os_flux_lo(T, 0), ly = os_flux_lo(T, 1), lz = os_flux_lo(T, 2);
    const double dxinv[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
    for (int k = g->lo[2]; k <= g->hi[2] + os_nodal(T, 2); ++k)
        for (int j = g->lo[1]; j <= g->hi[1] + os_nodal(T, 1); ++j)
            for (int i = g->lo[0]; i <= g->hi[0] + os_nodal(T, 0); ++i) {
                const double div = (os_interp(fx, OS_XF, SX, i + lx + 1, j, k) - os_interp(fx, OS_XF, SX, i + lx, j, k)) * dxinv[0]
                                   + (os_interp(fy, OS_YF, SY, i, j + ly + 1, k) - os_interp(fy, OS_YF, SY, i, j + ly, k)) * dxinv[1]
                                   + (os_interp(fz, OS_ZF, SZ, i, j, k + lz + 1) - os_interp(fz, OS_ZF, SZ, i, j, k + lz)) * dxinv[2];
                OAT(out, i, j, k) -= OAT(q, i, j, k) * div / os_interp(J, OS_CC, T, i, j, k);
            }
    return 0;
}

/* grad_at<TQ, TS> (Stencil_K.H:114-127). */
static void grad_at(const oa4* q, int tq, int ts, int i, int j, int k, const os_metric* m, const double dxinv[3],
                    double out[3])
{
    const double qx = os_deriv(q, tq, ts, 0, i, j, k) * dxinv[0];
    const double qy = os_deriv(q, tq, ts, 1, i, j, k) * dxinv[1];
    const double qz = os_deriv(q, tq, ts, 2, i, j, k) * dxinv[2] / m->zz;
    out[0] = qx - m->zx * qz;
    out[1] = qy - m->zy * qz;
    out[2] = qz;
}

static double area_dot(int D, const os_metric* m, const double v[3])
{
    if (D == 0) { return m->zz * v[0]; }
    if (D == 1) { return m->zz * v[1]; }
    return v[2] - m->zx * v[0] - m->zy * v[1];
}

/* log_mean / face_gamma are the implicit operator's (oracle/implicit.c). */
double oracle_face_gamma(const oa4* g, int ts, int i, int j, int k, int logz);

int oracle_scalar_diffusive_fluxes(const oa4* q, const oa4* gamma, const oa4* z, const ogeom* g, int logz, oa4* G0,
                                   oa4* G1, oa4* G2)
{
    oa4* G[3] = {G0, G1, G2};
    const double dxinv[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
    for (int D = 0; D < 3; ++D) {
        const int S = 1 << D;
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) { cv_range(OS_CC, D, g, a, &lo[a], &hi[a]); }
        for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int i = lo[0]; i <= hi[0]; ++i) {
                    const os_metric m = os_metric_at(z, S, i, j, k, dxinv);
                    double gr[3];
                    grad_at(q, OS_CC, S, i, j, k, &m, dxinv, gr);
                    OAT(G[D], i, j, k) = oracle_face_gamma(gamma, S, i, j, k, logz) * area_dot(D, &m, gr);
                }
    }
    return 0;
}

/* The fluxes of c-momentum through the control-volume faces of component c
 * (Source/Diffusion.cpp:31-59). */
int oracle_momentum_diffusive_fluxes(int c, const oa4* u, const oa4* v, const oa4* w, const oa4* nu, const oa4* z,
                                     const ogeom* g, int logz, oa4* G0, oa4* G1, oa4* G2)
{
    if (c < 0 || c > 2) { return -1; }
    oa4* G[3] = {G0, G1, G2};
    const double dxinv[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
    const int T = 1 << c;
    for (int D = 0; D < 3; ++D) {
        const int S = T ^ (1 << D);
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) { cv_range(T, D, g, a, &lo[a], &hi[a]); }
        for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int i = lo[0]; i <= hi[0]; ++i) {
                    const os_metric m = os_metric_at(z, S, i, j, k, dxinv);
                    double gu[3], gv[3], gw[3];
                    grad_at(u, OS_XF, S, i, j, k, &m, dxinv, gu);
                    grad_at(v, OS_YF, S, i, j, k, &m, dxinv, gv);
                    grad_at(w, OS_ZF, S, i, j, k, &m, dxinv, gw);
                    const double* gc = c == 0 ? gu : c == 1 ? gv : gw;
                    const double tau[3] = {gc[0] + gu[c], gc[1] + gv[c], gc[2] + gw[c]};
                    OAT(G[D], i, j, k) = oracle_face_gamma(nu, S, i, j, k, logz) * area_dot(D, &m, tau);
                }
    }
    return 0;
}
