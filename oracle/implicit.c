/*
 * implicit.c -- oracle of the first-order implicit operator A1.
 *
 * TEST INFRASTRUCTURE ONLY (see gust_oracle.h).
 *
 *   Source/ImplicitOperator.cpp:13-39    face_weight
 *   Source/ImplicitOperator.cpp:41-76    coefficients
 *   Source/ImplicitOperator.cpp:78-93    apply
 *   Source/Diffusion_K.H:35-64           log_mean, face_gamma
 */
#include "stencil.h"

#include <float.h>
#include <math.h>

/* Source/Diffusion_K.H:33-47 (double: window 0.01). */
static double log_mean(double a, double b)
{
    const double tiny = DBL_MIN;
    const double sum = a + b;
    const double x = (a - b) / (sum + tiny);
    const double s = (fabs(x) < 0.01) ? 1.0 : 0.0;
    const double x2 = x * x;
    const double series = 0.5 * sum * (1.0 - x2 * (1.0 / 3 + x2 * (4.0 / 45 + x2 * (44.0 / 945))));
    const double direct = (a - b) / (log(a + tiny) - log(b + tiny) + s);
    return (1.0 - s) * direct + s * series;
}

/* face_gamma<S> (Source/Diffusion_K.H:49-64); shared with transport.c. */
double oracle_face_gamma(const oa4* g, int ts, int i, int j, int k, int logz);
#define face_gamma oracle_face_gamma
double oracle_face_gamma(const oa4* g, int ts, int i, int j, int k, int logz)
{
    if (!logz || !os_nodal(ts, 2)) { return os_interp(g, OS_CC, ts, i, j, k); }
    const int lx = os_lo_off(OS_CC, ts, 0), ly = os_lo_off(OS_CC, ts, 1);
    const int nx = os_npts(OS_CC, ts, 0), ny = os_npts(OS_CC, ts, 1);
    double sum = 0;
    for (int b = 0; b < ny; ++b)
        for (int a = 0; a < nx; ++a) {
            sum += log_mean(OAT(g, i + lx + a, j + ly + b, k - 1), OAT(g, i + lx + a, j + ly + b, k));
        }
    return sum * (1.0 / (double)(nx * ny));
}

/* face_weight<T, C, D> (ImplicitOperator.cpp:16-39); C = -1 for a scalar. */
static double face_weight(int T, int C, int D, int side, int i, int j, int k, const oa4* f, const oa4* g, const oa4* z,
                          const double dxinv[3], int logz)
{
    const int S = T ^ (1 << D);
    const os_metric m = os_metric_at(z, S, i, j, k, dxinv);
    const double gs = face_gamma(g, S, i, j, k, logz);
    const double tr = os_interp(f, 1 << D, S, i, j, k);
    double dif;
    if (D < 2) {
        const double own = (C == D) ? 2.0 : 1.0;
        dif = gs * m.zz * own * dxinv[D];
    } else {
        double nc2 = 0;
        if (C == 0) { nc2 = m.zx * m.zx; }
        if (C == 1) { nc2 = m.zy * m.zy; }
        if (C == 2) { nc2 = 1; }
        dif = gs * (1 + m.zx * m.zx + m.zy * m.zy + nc2) / m.zz * dxinv[2];
    }
    const double inflow = (side == 0) ? os_max(tr, 0.0) : os_max(-tr, 0.0);
    return (dif + inflow) * dxinv[D];
}

int oracle_implicit_coefficients(int loc, const oa4* fx, const oa4* fy, const oa4* fz, const oa4* g, const oa4* z,
                                 const oa4* J, const ogeom* geo, int logz, oa4* A)
{
    if (!(loc == 0 || loc == 1 || loc == 2 || loc == 4)) { return -1; }
    const int C = loc == 0 ? -1 : loc == 1 ? 0 : loc == 2 ? 1 : 2;
    const int T = loc;
    const int lx = os_flux_lo(T, 0), ly = os_flux_lo(T, 1), lz = os_flux_lo(T, 2);
    const double dxinv[3] = {1.0 / geo->dx[0], 1.0 / geo->dx[1], 1.0 / geo->dx[2]};
    for (int k = geo->lo[2]; k <= geo->hi[2] + ((loc >> 2) & 1); ++k)
        for (int j = geo->lo[1]; j <= geo->hi[1] + ((loc >> 1) & 1); ++j)
            for (int i = geo->lo[0]; i <= geo->hi[0] + (loc & 1); ++i) {
                const double Jinv = 1.0 / os_interp(J, OS_CC, T, i, j, k);
                const double xm = face_weight(T, C, 0, 0, i + lx, j, k, fx, g, z, dxinv, logz) * Jinv;
                const double xp = face_weight(T, C, 0, 1, i + lx + 1, j, k, fx, g, z, dxinv, logz) * Jinv;
                const double ym = face_weight(T, C, 1, 0, i, j + ly, k, fy, g, z, dxinv, logz) * Jinv;
                const double yp = face_weight(T, C, 1, 1, i, j + ly + 1, k, fy, g, z, dxinv, logz) * Jinv;
                const double zm = face_weight(T, C, 2, 0, i, j, k + lz, fz, g, z, dxinv, logz) * Jinv;
                const double zp = face_weight(T, C, 2, 1, i, j, k + lz + 1, fz, g, z, dxinv, logz) * Jinv;
                const int64_t n = A->nstride;
                double* a = &OAT(A, i, j, k);
                a[OC_XM * n] = xm;
                a[OC_XP * n] = xp;
                a[OC_YM * n] = ym;
                a[OC_YP * n] = yp;
                a[OC_ZM * n] = zm;
                a[OC_ZP * n] = zp;
                a[OC_P * n] = ((xm + xp) + (ym + yp)) + (zm + zp);
            }
    return 0;
}

int oracle_apply_implicit(int loc, const oa4* A, const oa4* q, const ogeom* geo, oa4* out)
{
    if (!(loc == 0 || loc == 1 || loc == 2 || loc == 4)) { return -1; }
    const int64_t n = A->nstride;
    for (int k = geo->lo[2]; k <= geo->hi[2] + ((loc >> 2) & 1); ++k)
        for (int j = geo->lo[1]; j <= geo->hi[1] + ((loc >> 1) & 1); ++j)
            for (int i = geo->lo[0]; i <= geo->hi[0] + (loc & 1); ++i) {
                const double* a = &OAT(A, i, j, k);
                OAT(out, i, j, k) = a[OC_P * n] * OAT(q, i, j, k)
                                    - (a[OC_XM * n] * OAT(q, i - 1, j, k) + a[OC_XP * n] * OAT(q, i + 1, j, k))
                                    - (a[OC_YM * n] * OAT(q, i, j - 1, k) + a[OC_YP * n] * OAT(q, i, j + 1, k))
                                    - (a[OC_ZM * n] * OAT(q, i, j, k - 1) + a[OC_ZP * n] * OAT(q, i, j, k + 1));
            }
    return 0;
}
