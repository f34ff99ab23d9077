/*
 * massflux.c -- oracle of the terrain-following C-grid mass-flux operators.
 *
 * TEST INFRASTRUCTURE ONLY (see gust_oracle.h).
 *
 * Restates, in the reference's operation order:
 *   Source/MassFlux_K.H:19-39          slope_x, slope_y
 *   Source/MassFlux.cpp:29-39          slope_dp (one-sided at the ground)
 *   Source/MassFlux.cpp:51-74          omega_from_w
 *   Source/MassFlux.cpp:76-99          w_from_omega
 *   Source/MassFlux.cpp:101-127        transport_fluxes
 *   Source/MassFlux.cpp:129-136        velocity_divergence
 *   Source/MassFlux.cpp:138-179        pressure_gradient
 *   Source/Stencil_K.H:49-99           the averages and differences used
 * The stencil helpers accumulate from 0 in the reference's loop order (z
 * outer, then y, then x) and scale by 1 / (number of points) afterwards, so
 * the sums here start from 0.0 and add term by term in that order.
 */
#include "gust_oracle.h"

#define AT(v, i, j, k)                                                                                             \
    ((v)->p[((int64_t)(i) - (v)->begin[0]) + ((int64_t)(j) - (v)->begin[1]) * (v)->jstride                        \
            + ((int64_t)(k) - (v)->begin[2]) * (v)->kstride])

/* d z / d xi at the z-face (i, j, k): deriv<NODE, ZF, 0> * dxinv. */
static double slope_x(const oa4* z, int i, int j, int k, double dxinv)
{
    double s = 0.0;
    s += AT(z, i + 1, j, k) - AT(z, i, j, k);
    s += AT(z, i + 1, j + 1, k) - AT(z, i, j + 1, k);
    return (1.0 * (s * 0.5)) * dxinv;
}

/* d z / d eta at the z-face (i, j, k): deriv<NODE, ZF, 1> * dyinv. */
static double slope_y(const oa4* z, int i, int j, int k, double dyinv)
{
    double s = 0.0;
    s += AT(z, i, j + 1, k) - AT(z, i, j, k);
    s += AT(z, i + 1, j + 1, k) - AT(z, i + 1, j, k);
    return (1.0 * (s * 0.5)) * dyinv;
}

/* u (x-faces) at the z-face (i, j, k): interp<XF, ZF>. */
static double u_at_zface(const oa4* u, int i, int j, int k)
{
    double s = 0.0;
    s += AT(u, i, j, k - 1);
    s += AT(u, i + 1, j, k - 1);
    s += AT(u, i, j, k);
    s += AT(u, i + 1, j, k);
    return s * 0.25;
}

/* v (y-faces) at the z-face (i, j, k): interp<YF, ZF>. */
static double v_at_zface(const oa4* v, int i, int j, int k)
{
    double s = 0.0;
    s += AT(v, i, j, k - 1);
    s += AT(v, i, j + 1, k - 1);
    s += AT(v, i, j, k);
    s += AT(v, i, j + 1, k);
    return s * 0.25;
}

/* J (cells) at a face normal to d: interp<CC, face(d)>. */
static double J_at_face(const oa4* J, int i, int j, int k, int d)
{
    double s = 0.0;
    s += AT(J, i - (d == 0), j - (d == 1), k - (d == 2));
    s += AT(J, i, j, k);
    return s * 0.5;
}

/* slope times d_zeta p at the z-face (i, j, k), one-sided on the ground
 * face (Source/MassFlux.cpp:29-39); D = 0: the x slope, 1: the y slope. */
static double slope_dp(const oa4* z, const oa4* p, int i, int j, int k, int klo, int D, const double dxinv[3])
{
    const double s = D == 0 ? slope_x(z, i, j, k, dxinv[0]) : slope_y(z, i, j, k, dxinv[1]);
    const double dp = (k == klo) ? 3.0 * AT(p, i, j, k + 1) - 2.0 * AT(p, i, j, k) - AT(p, i, j, k + 2)
                                 : AT(p, i, j, k) - AT(p, i, j, k - 1);
    return s * dp * dxinv[2];
}

static void inv_dx(const ogeom* g, double dxinv[3])
{
    for (int d = 0; d < 3; ++d) { dxinv[d] = 1.0 / g->dx[d]; }
}

void oracle_omega_from_w(oa4* omega, const oa4* u, const oa4* v, const oa4* w, const oa4* z, const ogeom* g,
                         int ngrow)
{
    double dxinv[3];
    inv_dx(g, dxinv);
    for (int k = g->lo[2] - ngrow; k <= g->hi[2] + 1 + ngrow; ++k)
        for (int j = g->lo[1] - ngrow; j <= g->hi[1] + ngrow; ++j)
            for (int i = g->lo[0] - ngrow; i <= g->hi[0] + ngrow; ++i) {
                AT(omega, i, j, k) = AT(w, i, j, k) - slope_x(z, i, j, k, dxinv[0]) * u_at_zface(u, i, j, k)
                                     - slope_y(z, i, j, k, dxinv[1]) * v_at_zface(v, i, j, k);
            }
}

void oracle_w_from_omega(oa4* w, const oa4* u, const oa4* v, const oa4* omega, const oa4* z, const ogeom* g)
{
    double dxinv[3];
    inv_dx(g, dxinv);
    for (int k = g->lo[2]; k <= g->hi[2] + 1; ++k)
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                AT(w, i, j, k) = AT(omega, i, j, k) + slope_x(z, i, j, k, dxinv[0]) * u_at_zface(u, i, j, k)
                                 + slope_y(z, i, j, k, dxinv[1]) * v_at_zface(v, i, j, k);
            }
}

void oracle_transport_fluxes(oa4* Fx, oa4* Fy, oa4* Fz, const oa4* u, const oa4* v, const oa4* w, const oa4* z,
                             const oa4* ax, const oa4* ay, const ogeom* g)
{
    for (int k = g->lo[2] - 1; k <= g->hi[2] + 1; ++k)
        for (int j = g->lo[1] - 1; j <= g->hi[1] + 1; ++j)
            for (int i = g->lo[0] - 1; i <= g->hi[0] + 2; ++i) { AT(Fx, i, j, k) = AT(ax, i, j, k) * AT(u, i, j, k); }
    for (int k = g->lo[2] - 1; k <= g->hi[2] + 1; ++k)
        for (int j = g->lo[1] - 1; j <= g->hi[1] + 2; ++j)
            for (int i = g->lo[0] - 1; i <= g->hi[0] + 1; ++i) { AT(Fy, i, j, k) = AT(ay, i, j, k) * AT(v, i, j, k); }
    oracle_omega_from_w(Fz, u, v, w, z, g, 1);
}

void oracle_velocity_divergence(oa4* div, const oa4* u, const oa4* v, const oa4* w, const oa4* z, const oa4* ax,
                                const oa4* ay, const oa4* az, const oa4* J, const ogeom* g, oa4* omega_scratch)
{
    oracle_omega_from_w(omega_scratch, u, v, w, z, g, 0);
    oracle_face_divergence(div, u, v, omega_scratch, ax, ay, az, J, g);
}

void oracle_pressure_gradient(oa4* gx, oa4* gy, oa4* gz, const oa4* p, const oa4* z, const oa4* J, const oa4* ax,
                              const oa4* ay, const ogeom* g)
{
    double dxinv[3];
    inv_dx(g, dxinv);
    const int klo = g->lo[2];
    const double quarter = 0.25;
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0] + 1; ++i) {
                const double cross =
                    quarter * ((slope_dp(z, p, i - 1, j, k, klo, 0, dxinv) + slope_dp(z, p, i, j, k, klo, 0, dxinv))
                               + (slope_dp(z, p, i - 1, j, k + 1, klo, 0, dxinv)
                                  + slope_dp(z, p, i, j, k + 1, klo, 0, dxinv)));
                AT(gx, i, j, k) = (AT(ax, i, j, k) * (AT(p, i, j, k) - AT(p, i - 1, j, k)) * dxinv[0] - cross)
                                  / J_at_face(J, i, j, k, 0);
            }
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1]; j <= g->hi[1] + 1; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                const double cross =
                    quarter * ((slope_dp(z, p, i, j - 1, k, klo, 1, dxinv) + slope_dp(z, p, i, j, k, klo, 1, dxinv))
                               + (slope_dp(z, p, i, j - 1, k + 1, klo, 1, dxinv)
                                  + slope_dp(z, p, i, j, k + 1, klo, 1, dxinv)));
                AT(gy, i, j, k) = (AT(ay, i, j, k) * (AT(p, i, j, k) - AT(p, i, j - 1, k)) * dxinv[1] - cross)
                                  / J_at_face(J, i, j, k, 1);
            }
    for (int k = g->lo[2]; k <= g->hi[2] + 1; ++k)
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                AT(gz, i, j, k) = (AT(p, i, j, k) - AT(p, i, j, k - 1)) * dxinv[2] / J_at_face(J, i, j, k, 2);
            }
}
