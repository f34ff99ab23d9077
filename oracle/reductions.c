/*
 * reductions.c -- oracle of the control loop's reductions.
 *
 * TEST INFRASTRUCTURE ONLY (see gust_oracle.h).
 *
 *   Source/Reductions.cpp:11-53     deterministic_sum (one box: columns in k
 *                                   order, column sums in storage order)
 *   Source/Reductions.cpp:55-66     global_max
 *   Source/Residual.cpp:101-150     scaled_sums (one box, owner mask)
 *   Source/Simplec.cpp:374-409      residual_scales, stage_norms
 *   Source/Simplec.cpp:411-445      SimplecStage::confidence, with
 *   Source/Closure.cpp:261-293      height_above_ground
 *
 * stage_norms' Dot and sums have no fixed order in the reference (a device
 * reduction); here they run in plain loop order (k, j, i).
 */
#include "gust_oracle.h"

#include <math.h>

#define AT(v, i, j, k)                                                                                             \
    ((v)->p[((int64_t)(i) - (v)->begin[0]) + ((int64_t)(j) - (v)->begin[1]) * (v)->jstride                        \
            + ((int64_t)(k) - (v)->begin[2]) * (v)->kstride])

double oracle_box_sum(const oa4* a, const int lo[3], const int hi[3])
{
    double total = 0;
    for (int j = lo[1]; j <= hi[1]; ++j)
        for (int i = lo[0]; i <= hi[0]; ++i) {
            double s = 0;
            for (int k = lo[2]; k <= hi[2]; ++k) { s += AT(a, i, j, k); }
            total += s;
        }
    return total;
}

double oracle_global_max(const oa4* a, const int lo[3], const int hi[3])
{
    double m = -INFINITY;
    for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i) {
                const double x = AT(a, i, j, k);
                if (isnan(x)) { return NAN; }
                if (x > m) { m = x; }
            }
    return (m == 0.0) ? 0.0 : m;
}

static void loc_hi(const ogeom* g, int loc, int hi[3])
{
    for (int d = 0; d < 3; ++d) { hi[d] = g->hi[d] + ((loc >> d) & 1); }
}

void oracle_scaled_sums(const oa4* q, const oa4* v, const oa4* off, double s, int loc, const ogeom* g,
                        double out[3])
{
    int hi[3];
    loc_hi(g, loc, hi);
    /* owner mask of one periodic box: the hi face along a periodic nodal axis is the lo face's image */
    const int skip_i = (g->periodic[0] && (loc & 1)) ? hi[0] : g->lo[0] - 1;
    const int skip_j = (g->periodic[1] && (loc & 2)) ? hi[1] : g->lo[1] - 1;
    const double sinv = 1.0 / s;
    double t[3] = {0, 0, 0};
    for (int j = g->lo[1]; j <= hi[1]; ++j)
        for (int i = g->lo[0]; i <= hi[0]; ++i) {
            double s0 = 0, s1 = 0, s2 = 0;
            for (int k = g->lo[2]; k <= hi[2]; ++k) {
                if (i == skip_i || j == skip_j) { continue; }
                const double qq = off ? AT(q, i, j, k) - AT(off, i, j, k) : AT(q, i, j, k);
                const double vv = AT(v, i, j, k) * sinv;
                s0 += fabs(qq) * sinv;
                s1 += vv * vv;
                s2 += 1.0;
            }
            t[0] += s0;
            t[1] += s1;
            t[2] += s2;
        }
    out[0] = t[0];
    out[1] = t[1];
    out[2] = t[2];
}

void oracle_residual_scales(double U, double theta_s, double k_s, double H, double scale[6])
{
    scale[0] = scale[1] = scale[2] = U * U / H;
    scale[3] = U / H;
    scale[4] = U * theta_s / H;
    scale[5] = U * k_s / H;
}

static const int kLoc[6] = {OLOC_XFACE, OLOC_YFACE, OLOC_ZFACE, OLOC_CELL, OLOC_CELL, OLOC_CELL};

int oracle_stage_norms(const oa4* const R[6], const oa4* u, const oa4* v, const oa4* w, const oa4* tke,
                       const oa4* ustar, double U, double theta_s, double k_s, double H, const ogeom* g,
                       double res[6], double mon[3])
{
    double scale[6];
    oracle_residual_scales(U, theta_s, k_s, H, scale);
    for (int n = 0; n < 6; ++n) {
        int hi[3];
        loc_hi(g, kLoc[n], hi);
        double dot = 0;
        for (int k = g->lo[2]; k <= hi[2]; ++k)
            for (int j = g->lo[1]; j <= hi[1]; ++j)
                for (int i = g->lo[0]; i <= hi[0]; ++i) {
                    const double x = AT(R[n], i, j, k);
                    dot += x * x;
                }
        const double npts = (double)(hi[0] - g->lo[0] + 1) * (hi[1] - g->lo[1] + 1) * (hi[2] - g->lo[2] + 1);
        res[n] = sqrt(dot) / sqrt(npts) / scale[n];
    }
    const oa4* vel[3] = {u, v, w};
    double speed = 0;
    for (int d = 0; d < 3; ++d) {
        int hi[3];
        loc_hi(g, 1 << d, hi);
        for (int k = g->lo[2]; k <= hi[2]; ++k)
            for (int j = g->lo[1]; j <= hi[1]; ++j)
                for (int i = g->lo[0]; i <= hi[0]; ++i) {
                    const double a = fabs(AT(vel[d], i, j, k));
                    if (isnan(a)) { speed = a; }
                    else if (!isnan(speed) && a > speed) { speed = a; }
                }
    }
    mon[0] = speed;
    const int nx = g->hi[0] - g->lo[0] + 1, ny = g->hi[1] - g->lo[1] + 1, nz = g->hi[2] - g->lo[2] + 1;
    double sk = 0;
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) { sk += AT(tke, i, j, k); }
    mon[1] = sk / ((double)nx * ny * nz);
    if (!ustar) {
        mon[2] = 0;
        return 2;
    }
    double su = 0;
    for (int j = g->lo[1]; j <= g->hi[1]; ++j)
        for (int i = g->lo[0]; i <= g->hi[0]; ++i) { su += AT(ustar, i, j, g->lo[2]); }
    mon[2] = su / ((double)nx * ny);
    return 3;
}

static double smax(double a, double b) { return (a < b) ? b : a; }   /* std::max */

void oracle_confidence(const oa4* const R[6], const oa4* z, double U, double theta_s, double k_s, double H,
                       const double threshold[6], const ogeom* g, double* share, double* top)
{
    double scale[6], t[6];
    oracle_residual_scales(U, theta_s, k_s, H, scale);
    for (int n = 0; n < 6; ++n) { t[n] = threshold[n] * scale[n]; }
    long long bad = 0;
    double m = -INFINITY;
    const int klo = g->lo[2];
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                const int ok = smax(fabs(AT(R[0], i, j, k)), fabs(AT(R[0], i + 1, j, k))) <= t[0]
                               && smax(fabs(AT(R[1], i, j, k)), fabs(AT(R[1], i, j + 1, k))) <= t[1]
                               && smax(fabs(AT(R[2], i, j, k)), fabs(AT(R[2], i, j, k + 1))) <= t[2]
                               && fabs(AT(R[3], i, j, k)) <= t[3] && fabs(AT(R[4], i, j, k)) <= t[4]
                               && fabs(AT(R[5], i, j, k)) <= t[5];
                double d = 0;
                if (!ok) {
                    ++bad;
                    double zc = 0, zg = 0;
                    for (int a = 0; a < 2; ++a)
                        for (int b = 0; b < 2; ++b) {
                            zg += AT(z, i + a, j + b, klo);
                            zc += AT(z, i + a, j + b, k) + AT(z, i + a, j + b, k + 1);
                        }
                    d = 0.125 * zc - 0.25 * zg;
                }
                if (d > m) { m = d; }
            }
    const double npts = (double)(g->hi[0] - g->lo[0] + 1) * (g->hi[1] - g->lo[1] + 1) * (g->hi[2] - g->lo[2] + 1);
    *share = 1.0 - (double)bad / npts;
    *top = m;
}
