/*
 * pseudotime.c -- oracle of the local pseudo-time step and its face means.
 *
 * TEST INFRASTRUCTURE ONLY (see gust_oracle.h).
 *
 *   Source/PseudoTime.cpp:12-51     local_dtau
 *   Source/PseudoTime.cpp:53-61     ser_cfl
 *   Source/Predictor.cpp:30-77      face_dtau
 */
#include "gust_oracle.h"

#include <math.h>

#define AT(v, i, j, k)                                                                                             \
    ((v)->p[((int64_t)(i) - (v)->begin[0]) + ((int64_t)(j) - (v)->begin[1]) * (v)->jstride                        \
            + ((int64_t)(k) - (v)->begin[2]) * (v)->kstride])

static const double kMinRate = 1.0e-12;   /* Source/PseudoTime.H:27 */

/* interp<Face(d), CC>: the mean of a cell's two d-faces. */
static double cell_mean(const oa4* f, int i, int j, int k, int d)
{
    double s = 0.0;
    s += AT(f, i, j, k);
    s += AT(f, i + (d == 0), j + (d == 1), k + (d == 2));
    return s * 0.5;
}

int oracle_local_dtau(oa4* dtau, const oa4* u, const oa4* v, const oa4* w, const oa4* z, const oa4* J, const oa4* nu,
                      double cfl, const ogeom* g, oa4* omega_scratch)
{
    if (isnan(cfl) || !isfinite(cfl) || !(cfl > 0)) { return -1; }
    oracle_omega_from_w(omega_scratch, u, v, w, z, g, 0);
    const double dxinv[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
    const double diff_h = 2 * (dxinv[0] * dxinv[0] + dxinv[1] * dxinv[1]);
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                const double uc = cell_mean(u, i, j, k, 0), vc = cell_mean(v, i, j, k, 1),
                             oc = cell_mean(omega_scratch, i, j, k, 2);
                double rate = fabs(uc) * dxinv[0] + fabs(vc) * dxinv[1] + fabs(oc) * dxinv[2] / AT(J, i, j, k);
                if (nu) { rate += AT(nu, i, j, k) * diff_h; }
                AT(dtau, i, j, k) = cfl / (rate < kMinRate ? kMinRate : rate);   /* std::max(rate, kMinRate) */
            }
    return 0;
}

double oracle_ser_cfl(double cfl, double r, double cfl0, double cfl_max, double p, double growth)
{
    const double gc = growth * cfl;
    const double capped = gc < cfl_max ? gc : cfl_max;   /* amrex::min = std::min(a, b): b < a ? b : a */
    if (p == 0) { return cfl0 < capped ? cfl0 : capped; }
    const double r_min = pow(cfl0 / cfl_max, 1.0 / p);
    if (!(r > r_min)) { return capped; }
    const double law = cfl0 * pow(r, -p);
    return law < capped ? law : capped;
}

int oracle_face_dtau(oa4* fx, oa4* fy, oa4* fz, const oa4* d, const ogeom* g)
{
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                const double v = AT(d, i, j, k);
                if (isnan(v) || !(v > 0) || !isfinite(v)) { return -1; }
            }
    const int perx = g->periodic[0], pery = g->periodic[1];
    const int nx = g->hi[0] - g->lo[0] + 1, ny = g->hi[1] - g->lo[1] + 1;
    /* periodic images of the cell field (FillBoundary) */
#define DW(i, j, k) AT(d, ((i) < g->lo[0] ? (i) + nx : (i) > g->hi[0] ? (i) - nx : (i)),                             \
                      ((j) < g->lo[1] ? (j) + ny : (j) > g->hi[1] ? (j) - ny : (j)), (k))
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0] + 1; ++i) {
                const int lo = !perx && i == g->lo[0], hi = !perx && i == g->hi[0] + 1;
                AT(fx, i, j, k) = lo ? DW(i, j, k) : hi ? DW(i - 1, j, k) : 0.5 * (DW(i - 1, j, k) + DW(i, j, k));
            }
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1]; j <= g->hi[1] + 1; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                const int lo = !pery && j == g->lo[1], hi = !pery && j == g->hi[1] + 1;
                AT(fy, i, j, k) = lo ? DW(i, j, k) : hi ? DW(i, j - 1, k) : 0.5 * (DW(i, j - 1, k) + DW(i, j, k));
            }
    for (int k = g->lo[2]; k <= g->hi[2] + 1; ++k)
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                const int lo = k == g->lo[2], hi = k == g->hi[2] + 1;
                AT(fz, i, j, k) = lo ? AT(d, i, j, k) : hi ? AT(d, i, j, k - 1) : 0.5 * (AT(d, i, j, k - 1) + AT(d, i, j, k));
            }
#undef DW
    return 0;
}
