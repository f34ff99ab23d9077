/*
 * physics.c -- oracle of the residual's physics on a tile: heights above the ground, the k-L closure, the neutral surface
 * Derived in part from ERF (BSD 3-Clause, Copyright (c) 2022 The Regents of the University of California
 * et al.); the full notice is in THIRD_PARTY_NOTICES.md at the repository root.
 * layer, the ghost fills of the ground and the top, the momentum forcing,
 * and the wall and boundary rows of the residual and the implicit side.
 *
 * TEST INFRASTRUCTURE ONLY (see gust_oracle.h).
 *
 *   Source/Closure.cpp:261-293          height_above_ground
 *   Source/Closure.cpp:132-148          copy_vertical_ghosts
 *   Source/Closure.cpp:296-332          stratification_and_strain
 *   Source/Closure.cpp:366-466          Closure::evaluate (nut_relax 1, no
 *                                       gate, no cap, switches from the state)
 *   Source/Closure_K.H:94-176           switches_for, closure_point,
 *                                       patankar_diagonal
 *   Source/SurfaceLayer.cpp:35-81       Column, tangential_wind_ground
 *   Source/SurfaceLayer.cpp:114-217     columns<> (first centre, neutral)
 *   Source/SurfaceLayer.cpp:430-464     set_ground_fluxes
 *   Source/SurfaceLayer.cpp:466-492     apply_wall_k
 *   Source/SurfaceLayer.cpp:494-536     apply_wall_to_implicit
 *   Source/SurfaceLayer.cpp:538-567     apply_wall_k_to_implicit
 *   Source/StabilityFunctions.H:37-47   most_phi (Businger-Dyer)
 *   Source/BoundaryConditions.cpp:38-65 rule_for
 *   Source/BoundaryConditions.cpp:83-104, 418-534, 540-576, 597-666, 756-797
 *                                       ghost_slab, fill_face, fill_gradient_z,
 *                                       ground_gradients, fill_ghosts,
 *                                       fill_no_flux_w (every face type)
 *   Source/BoundaryConditions.cpp:800-940
 *                                       apply_rows, dirichlet_rows,
 *                                       apply_to_implicit
 *   Source/InflowProfile.H:137-200      table_interp, View::eval
 *   Source/Heights_K.H:20-68            NodeZ, height_above_ground<T>, point_z<T>
 *   Source/Forcing.cpp:77-170, 285-299  coriolis_parameter, fill_theta_ref,
 *                                       subtract_buoyancy, subtract_coriolis,
 *                                       subtract_pressure_gradient
 *   Source/MassFlux_K.H:31-39           no_flux_w
 */
#include "gust_oracle.h"
#include "stencil.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>

#define AT OAT

/* ---- heights ------------------------------------------------------------ */

/* NodeZ: lateral indices clamped to the stored nodes, vertical ones
 * extrapolated linearly (Heights_K.H:20-33). */
static double nodez(const oa4* z, int i, int j, int k)
{
    const int lo0 = z->begin[0], lo1 = z->begin[1], lo2 = z->begin[2];
    const int hi0 = z->end[0] - 1, hi1 = z->end[1] - 1, hi2 = z->end[2] - 1;
    i = i < lo0 ? lo0 : (i > hi0 ? hi0 : i);
    j = j < lo1 ? lo1 : (j > hi1 ? hi1 : j);
    if (k < lo2) { return AT(z, i, j, lo2) + (lo2 - k) * (AT(z, i, j, lo2) - AT(z, i, j, lo2 + 1)); }
    if (k > hi2) { return AT(z, i, j, hi2) + (k - hi2) * (AT(z, i, j, hi2) - AT(z, i, j, hi2 - 1)); }
    return AT(z, i, j, k);
}

static double hag(const oa4* z, int t, int i, int j, int k, int klo)
{
    const int nx = os_nodal(t, 0) ? 1 : 2, ny = os_nodal(t, 1) ? 1 : 2, nz = os_nodal(t, 2) ? 1 : 2;
    double zp = 0, zg = 0;
    for (int a = 0; a < nx; ++a)
        for (int b = 0; b < ny; ++b) {
            zg += nodez(z, i + a, j + b, klo);
            for (int c = 0; c < nz; ++c) { zp += nodez(z, i + a, j + b, k + c); }
        }
    return zp / (nx * ny * nz) - zg / (nx * ny);
}

static double point_z(const oa4* z, int t, int i, int j, int k)
{
    const int nx = os_nodal(t, 0) ? 1 : 2, ny = os_nodal(t, 1) ? 1 : 2, nz = os_nodal(t, 2) ? 1 : 2;
    double zp = 0;
    for (int a = 0; a < nx; ++a)
        for (int b = 0; b < ny; ++b)
            for (int c = 0; c < nz; ++c) { zp += nodez(z, i + a, j + b, k + c); }
    return zp / (nx * ny * nz);
}

static void copy_vertical_ghosts(oa4* a, const ogeom* g)
{
    for (int j = a->begin[1]; j < a->end[1]; ++j)
        for (int i = a->begin[0]; i < a->end[0]; ++i) {
            AT(a, i, j, g->lo[2] - 1) = AT(a, i, j, g->lo[2]);
            AT(a, i, j, g->hi[2] + 1) = AT(a, i, j, g->hi[2]);
        }
}

void oracle_height_above_ground(const oa4* z, const ogeom* g, oa4* dist)
{
    const int klo = g->lo[2];
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1] - 1; j <= g->hi[1] + 1; ++j)
            for (int i = g->lo[0] - 1; i <= g->hi[0] + 1; ++i) {
                double zc = 0, zg = 0;
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b) {
                        zg += AT(z, i + a, j + b, klo);
                        zc += AT(z, i + a, j + b, k) + AT(z, i + a, j + b, k + 1);
                    }
                AT(dist, i, j, k) = 0.125 * zc - 0.25 * zg;
            }
    copy_vertical_ghosts(dist, g);
}

/* ---- the k-L closure ---------------------------------------------------- */

typedef struct {
    double kappa, Cmu0, Cmu0_pow3, inv_Cb_sq, Rt_crit, Rt_min, l_g_max, tke_floor, n2_eps;
} kconst;

static kconst closure_constants(const oclosure* o)
{
    kconst c;
    c.kappa = o->kappa;
    c.Cmu0 = o->Cmu0;
    c.Cmu0_pow3 = o->Cmu0 * o->Cmu0 * o->Cmu0;
    c.inv_Cb_sq = 1.0 / (o->Cb * o->Cb);
    c.Rt_crit = o->Rt_crit;
    c.Rt_min = o->Rt_min;
    c.l_g_max = o->max_geom_lscale;
    c.tke_floor = o->tke_floor > 0 ? o->tke_floor : DBL_EPSILON;   /* 0: machine epsilon */
    c.n2_eps = DBL_EPSILON;
    return c;
}

static double pos_part(double x) { return 0.5 * (x + fabs(x)); }
static double neg_part(double x) { return 0.5 * (x - fabs(x)); }

static double smooth_rt(double Rt, int smooth, double Rt_crit, double Rt_min)
{
    const double x = neg_part(Rt - Rt_crit);
    const double a = Rt_min - Rt_crit;
    const double s = Rt_crit + a * x / (x + a);
    return smooth ? s : Rt;
}

typedef struct {
    int branch, smooth_g, smooth, floored;
} kswitch;

static kswitch switches_for(double k, double N2, double l_g, const kconst* c)
{
    kswitch s;
    s.floored = !(k > c->tke_floor);
    const double kf = s.floored ? c->tke_floor : k;
    s.branch = (fabs(N2) <= c->n2_eps) ? 0 : (N2 > c->n2_eps) ? 1 : 2;
    const double c6 = c->Cmu0_pow3 * c->Cmu0_pow3;
    const double N2s = pos_part(N2), N2u = neg_part(N2);
    const double Rt_g = l_g * l_g * N2u / (kf * c->Cmu0_pow3 * c->Cmu0_pow3);
    s.smooth_g = (s.branch == 2) && (Rt_g < c->Rt_crit);
    const double L_unstable = l_g * sqrt(1.0 - c6 * c->inv_Cb_sq * smooth_rt(Rt_g, s.smooth_g, c->Rt_crit, c->Rt_min));
    const double L_stable = sqrt(1.0 / (1.0 / (l_g * l_g) + c->inv_Cb_sq * N2s / kf));
    const double L = (s.branch == 0) ? l_g : (s.branch == 1) ? L_stable : L_unstable;
    s.smooth = (L * L * N2 / (kf * c->Cmu0_pow3 * c->Cmu0_pow3) < c->Rt_crit);
    return s;
}

typedef struct {
    double k, L, Rt, cmu, cmu_prime, nut, nut_prime, eps, Ps, Pb;
} kpoint;

static kpoint closure_point(double k, double N2, double S2, double l_g, const kconst* c, kswitch s)
{
    kpoint p;
    p.k = s.floored ? c->tke_floor : k;
    const double c6 = c->Cmu0_pow3 * c->Cmu0_pow3;
    const double N2s = pos_part(N2), N2u = neg_part(N2);
    const double L_stable = sqrt(1.0 / ((1.0 / (l_g * l_g)) + (c->inv_Cb_sq * N2s) / p.k));
    const double Rt_g = (l_g * l_g * N2u) / (p.k * c->Cmu0_pow3 * c->Cmu0_pow3);
    const double L_unstable = l_g * sqrt(1.0 - (c6 * c->inv_Cb_sq) * smooth_rt(Rt_g, s.smooth_g, c->Rt_crit, c->Rt_min));
    p.L = (s.branch == 0) ? l_g : (s.branch == 1) ? L_stable : L_unstable;
    p.eps = c->Cmu0_pow3 * pow(p.k, 1.5) / p.L;
    p.Rt = smooth_rt(p.L * p.L * N2 / (p.k * c->Cmu0_pow3 * c->Cmu0_pow3), s.smooth, c->Rt_crit, c->Rt_min);
    p.cmu = (c->Cmu0 + 0.108 * p.Rt) / (1.0 + 0.308 * p.Rt + 0.00837 * p.Rt * p.Rt);
    p.cmu_prime = c->Cmu0 / (1.0 + 0.277 * p.Rt);
    p.nut = p.cmu * sqrt(p.k) * p.L;
    p.nut_prime = p.cmu_prime / p.cmu * p.nut;
    p.Ps = p.nut * S2;
    p.Pb = -p.nut_prime * N2;
    return p;
}

static double geom_length(double raw, double l_max) { return l_max * raw / (l_max + raw); }

/* N^2 and S^2 at cell (i, j, k) (Closure.cpp:296-332). */
static void strat_strain(const oa4* u, const oa4* v, const oa4* w, const oa4* th, const oa4* z, double gt0,
                         const double dxinv[3], int i, int j, int k, double* N2, double* S2)
{
    const os_metric m = os_metric_at(z, OS_CC, i, j, k, dxinv);
    const oa4* f[4] = {th, u, v, w};
    const int t[4] = {OS_CC, OS_XF, OS_YF, OS_ZF};
    double G[4][3];
    for (int n = 0; n < 4; ++n) {
        const double qx = os_deriv(f[n], t[n], OS_CC, 0, i, j, k) * dxinv[0];
        const double qy = os_deriv(f[n], t[n], OS_CC, 1, i, j, k) * dxinv[1];
        const double qz = os_deriv(f[n], t[n], OS_CC, 2, i, j, k) * dxinv[2] / m.zz;
        G[n][0] = qx - m.zx * qz;
        G[n][1] = qy - m.zy * qz;
        G[n][2] = qz;
    }
    *N2 = gt0 * G[0][2];
    double s = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const double e = G[1 + a][b] + G[1 + b][a];
            s += e * e;
        }
    *S2 = 0.5 * s;
}

int oracle_closure_evaluate(const oa4* u, const oa4* v, const oa4* w, const oa4* theta, const oa4* tke, const oa4* z,
                            const oa4* dist, double z0, const oclosure* opt, double gravity, double theta0,
                            const ogeom* g, oa4* nu, oa4* gth, oa4* gk, oa4* Ps, oa4* Pb, oa4* eps, oa4* dSdk)
{
    const kconst c = closure_constants(opt);
    const double dxinv[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
    const double gt0 = gravity / theta0;
    const double inv_sigma = 1.0 / opt->sigma_k;
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1] - 1; j <= g->hi[1] + 1; ++j)
            for (int i = g->lo[0] - 1; i <= g->hi[0] + 1; ++i) {
                if (!(AT(dist, i, j, k) + z0 > 0)) { return -1; }
            }
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1] - 1; j <= g->hi[1] + 1; ++j)
            for (int i = g->lo[0] - 1; i <= g->hi[0] + 1; ++i) {
                double N2, S2;
                strat_strain(u, v, w, theta, z, gt0, dxinv, i, j, k, &N2, &S2);
                const double l_g = geom_length(c.kappa * (AT(dist, i, j, k) + z0), c.l_g_max);
                const double kk = AT(tke, i, j, k);
                const kswitch s = switches_for(kk, N2, l_g, &c);
                const kpoint p = closure_point(kk, N2, S2, l_g, &c, s);
                const double ratio = p.cmu_prime / p.cmu;
                const double nt = 1.0 * os_min(p.nut, DBL_MAX) + (1.0 - 1.0) * kk;
                AT(nu, i, j, k) = opt->nu_molecular + nt;
                AT(gth, i, j, k) = opt->alpha_molecular + ratio * nt;
                AT(gk, i, j, k) = opt->nu_molecular + nt * inv_sigma;
                const int valid = i >= g->lo[0] && i <= g->hi[0] && j >= g->lo[1] && j <= g->hi[1];
                if (valid) {
                    const double ps = nt * S2, pb = -ratio * nt * N2;
                    AT(Ps, i, j, k) = ps;
                    AT(Pb, i, j, k) = pb;
                    AT(eps, i, j, k) = p.eps;
                    AT(dSdk, i, j, k) = (p.eps + pos_part(-pb)) / p.k;
                }
            }
    copy_vertical_ghosts(nu, g);
    copy_vertical_ghosts(gth, g);
    copy_vertical_ghosts(gk, g);
    return 0;
}

/* ---- the neutral surface layer ---------------------------------------- */

static const double kWsMin = 0.1, kWallKUstarMin = 0.01;

/* One ground column (i, j): first-centre reference, neutral log law. */
static void surface_column(const oa4* u, const oa4* v, const oa4* z, const osurface* o, const double dxinv[3],
                           int klo, int i, int j, double out[10])
{
    const double ground = 0.25 * ((AT(z, i, j, klo) + AT(z, i + 1, j, klo)) +
                                  (AT(z, i, j + 1, klo) + AT(z, i + 1, j + 1, klo)));
    const double h0 = os_interp(z, OS_NODE, OS_CC, i, j, klo) - ground;
    const double zx = os_deriv(z, OS_NODE, OS_ZF, 0, i, j, klo) * dxinv[0];
    const double zy = os_deriv(z, OS_NODE, OS_ZF, 1, i, j, klo) * dxinv[1];
    const double area = sqrt(1 + zx * zx + zy * zy);
    const double n[3] = {-zx / area, -zy / area, 1 / area};
    /* tangential_wind_ground at the ground cell */
    const double uc = 0.5 * (AT(u, i, j, klo) + AT(u, i + 1, j, klo));
    const double vc = 0.5 * (AT(v, i, j, klo) + AT(v, i, j + 1, klo));
    const double U3[3] = {uc, vc, zx * uc + zy * vc};
    const double un = U3[0] * n[0] + U3[1] * n[1] + U3[2] * n[2];
    double ut[3];
    for (int d = 0; d < 3; ++d) { ut[d] = U3[d] - un * n[d]; }
    /* the reference point is the first centre: the interpolation weight is 0 */
    const double ia = 1.0 / area;
    const double la = log(h0 * ia + o->z0);
    const double f = (log(h0 * ia + o->z0) - la) / 1.0;
    for (int d = 0; d < 3; ++d) { ut[d] = ut[d] + f * 0.0; }
    const double mag = sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
    const double U = os_max(mag, kWsMin);
    const double dd = h0 * ia;
    const double lz = log((dd + o->z0) / o->z0);
    const double den_min = o->floor_one ? 1.0 : os_min(lz, 1.0);
    const double ustar = o->kappa * U / os_max(lz, den_min);
    const double d0 = h0 * ia;
    const double cdrag = ustar * ustar / U * area;
    out[0] = ustar;
    out[1] = U;
    out[2] = d0;
    out[3] = cdrag * ut[0];
    out[4] = cdrag * ut[1];
    out[5] = cdrag * ut[2];
    out[6] = cdrag;
    out[7] = ustar * ustar * (1.0 / (o->Cmu0 * o->Cmu0));
    out[8] = area;
    out[9] = 0.0;   /* 1/L, neutral */
}

int oracle_surface_state(const oa4* u, const oa4* v, const oa4* z, const osurface* o, const ogeom* g, oa4* ss)
{
    const double dxinv[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
    const int klo = g->lo[2];
    int bad = 0;
    for (int j = g->lo[1] - 1; j <= g->hi[1] + 1; ++j)
        for (int i = g->lo[0] - 1; i <= g->hi[0] + 1; ++i) {
            double out[10];
            surface_column(u, v, z, o, dxinv, klo, i, j, out);
            for (int n = 0; n < OSS_N; ++n) { ss->p[(i - ss->begin[0]) + (j - ss->begin[1]) * ss->jstride + n * ss->nstride] = out[n]; }
            const double a = out[0], b = out[7];
            if (a != a || b != b || !(a > 0 && a < DBL_MAX)) { ++bad; }
        }
    return bad ? -1 : 0;
}

#define SS(ss, n, i, j) ((ss)->p[((i) - (ss)->begin[0]) + ((j) - (ss)->begin[1]) * (ss)->jstride + (n) * (ss)->nstride])

/* ---- the inflow profile -------------------------------------------------- */

static double table_interp(const double* x, const double* y, int n, double xq)
{
    if (n == 1 || xq <= x[0]) { return y[0]; }
    if (xq >= x[n - 1]) { return y[n - 1]; }
    int m = 1;
    while (m < n - 1 && !(xq < x[m])) { ++m; }
    const double w = (xq - x[m - 1]) / (x[m] - x[m - 1]);
    return y[m - 1] + w * (y[m] - y[m - 1]);
}

void oracle_profile_eval(const oprofile* p, double z, double out[5])
{
    z = os_max(z, 0.0);
    const double c = 1.0 / (1.0 + p->kappa * (z + p->z0) * p->inv_lmax);
    double ustar_l, u, v, th, k;
    if (p->kind == 2) {
        u = table_interp(p->tz, p->tu, p->n, z);
        v = table_interp(p->tz, p->tv, p->n, z);
        th = table_interp(p->tz, p->tt, p->n, z);
        int m = 1;
        while (m < p->n - 1 && !(z < p->tz[m])) { ++m; }
        const double s0 = sqrt(p->tu[m - 1] * p->tu[m - 1] + p->tv[m - 1] * p->tv[m - 1]);
        const double s1 = sqrt(p->tu[m] * p->tu[m] + p->tv[m] * p->tv[m]);
        const double dsdz = (p->n > 1) ? (s1 - s0) / (p->tz[m] - p->tz[m - 1]) : 0.0;
        ustar_l = p->kappa * (z + p->z0) * os_max(dsdz, 0.0) * c;
        if (p->has_k) {
            k = table_interp(p->tz, p->tk, p->n, z);
            ustar_l = p->Cmu0 * sqrt(os_max(k, 0.0));
        } else {
            k = ustar_l * ustar_l / (p->Cmu0 * p->Cmu0);
        }
    } else {
        double S;
        if (p->kind == 0) {
            S = p->ustar / p->kappa * log((z + p->z0) / p->z0);
            ustar_l = p->ustar * c;
        } else {
            S = p->speed_ref * pow(z / p->z_ref, p->alpha);
            const double dsdz = p->alpha * S / (0.5 * (z + p->z0 + fabs(z - p->z0)));
            ustar_l = p->kappa * (z + p->z0) * dsdz * c;
        }
        u = S * p->cosd;
        v = S * p->sind;
        th = table_interp(p->rz, p->rt, p->rn, z);
        k = ustar_l * ustar_l / (p->Cmu0 * p->Cmu0);
    }
    out[0] = u;
    out[1] = v;
    out[2] = th;
    out[3] = k;
    out[4] = p->kappa * ustar_l * (z + p->z0) * c;
}

/* ---- ghost fills ------------------------------------------------------- */

enum { R_EVEN, R_ODD, R_VALUE, R_PROFILE, R_GRADIENT };
enum { F_U, F_V, F_W, F_P, F_THETA, F_TKE };

static int rule_for(int f, int t, int d, int lo)
{
    const int normal_vel = (f == F_U && d == 0) || (f == F_V && d == 1) || (f == F_W && d == 2);
    const int vel = f == F_U || f == F_V || f == F_W;
    switch (t) {
    case OFACE_INFLOW: return f == F_P ? R_EVEN : R_PROFILE;
    case OFACE_OUTFLOW: return f == F_P ? R_ODD : R_EVEN;
    case OFACE_SLIP_WALL:
        if (normal_vel) { return R_VALUE; }
        return (f == F_THETA && d == 2 && !lo) ? R_GRADIENT : R_EVEN;
    case OFACE_NO_SLIP_WALL: return vel ? R_VALUE : R_EVEN;
    case OFACE_SURFACE_LAYER:
        if (normal_vel) { return R_VALUE; }
        return (vel || f == F_THETA) ? R_GRADIENT : R_EVEN;
    case OFACE_GEOSTROPHIC:
        if (normal_vel) { return R_VALUE; }
        if (f == F_U || f == F_V) { return R_VALUE; }
        return f == F_THETA ? R_GRADIENT : R_EVEN;
    default: return R_EVEN;
    }
}

static int dirichlet_normal(int t)
{
    return t == OFACE_INFLOW || t == OFACE_SLIP_WALL || t == OFACE_NO_SLIP_WALL || t == OFACE_SURFACE_LAYER ||
           t == OFACE_GEOSTROPHIC;
}

static int face_of(int d, int lo) { return lo ? d : d + 3; }

static double profile_value(const oprofile* p, int f, double h)
{
    double o[5];
    oracle_profile_eval(p, h, o);
    return f == F_U ? o[0] : f == F_V ? o[1] : f == F_THETA ? o[2] : f == F_TKE ? o[3] : 0.0;
}

/* The wall law's gradient per column (BoundaryConditions.cpp:540-576). */
static void ground_gradient(const oa4* ss, const osurface* o, int i, int j, double G[3])
{
    const double zeta = os_max(os_min(SS(ss, OSS_D0, i, j) * SS(ss, OSS_INVL, i, j), 100.0), -100.0);
    const double az = fabs(zeta);
    const double zs = 0.5 * (az + zeta), zu = 0.5 * (zeta - az);
    const double r = sqrt(1.0 - 16.0 * zu);
    const double phi_m = (1.0 + 5.0 * zs) / (1.0 + 0.0 * zs) + (1.0 / sqrt(r) - 1.0);
    const double phi_h = (1.0 + 5.0 * zs) / (1.0 + 0.0 * zs) + (1.0 / r - 1.0);
    const double s = o->kappa * (SS(ss, OSS_D0, i, j) + o->z0);
    const double t0 = SS(ss, OSS_TAU0, i, j), t1 = SS(ss, OSS_TAU1, i, j), t2 = SS(ss, OSS_TAU2, i, j);
    const double tmag = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
    const double tsafe = tmag + DBL_MIN;
    const double us = SS(ss, OSS_USTAR, i, j);
    G[0] = us / s * (t0 / tsafe) * phi_m;
    G[1] = us / s * (t1 / tsafe) * phi_m;
    G[2] = -0.0 / (us * s) * phi_h;   /* -q/(u* s) phi_h with q = 0 (neutral) */
}

static int wrap(int i, int lo, int n) { return lo + ((i - lo) % n + n) % n; }
static int os_imax(int a, int b) { return a > b ? a : b; }
static int os_imin(int a, int b) { return a < b ? a : b; }
static int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

/* The surface-state column the ground gradient of column (i, j) comes from
 * (BoundaryConditions.cpp:617-647): the gradient is computed on the ground
 * columns and one ghost column (ground_gradients); FillBoundary then gives
 * every ghost column whose periodic image is a ground column that image; a
 * column still unset takes the one clamped into the domain along the
 * non-periodic directions (itself then a ground column, an image, or a
 * computed ghost column). */
static void gg_column(const ogeom* g, int i, int j, int* si, int* sj)
{
    const int nx = g->hi[0] - g->lo[0] + 1, ny = g->hi[1] - g->lo[1] + 1;
    for (int pass = 0; pass < 2; ++pass) {
        const int valid = i >= g->lo[0] && i <= g->hi[0] && j >= g->lo[1] && j <= g->hi[1];
        if (valid) { *si = i; *sj = j; return; }
        const int ii = (g->periodic[0] && (i < g->lo[0] || i > g->hi[0])) ? wrap(i, g->lo[0], nx) : i;
        const int jj = (g->periodic[1] && (j < g->lo[1] || j > g->hi[1])) ? wrap(j, g->lo[1], ny) : j;
        if (ii >= g->lo[0] && ii <= g->hi[0] && jj >= g->lo[1] && jj <= g->hi[1]) { *si = ii; *sj = jj; return; }
        const int computed = i >= g->lo[0] - 1 && i <= g->hi[0] + 1 && j >= g->lo[1] - 1 && j <= g->hi[1] + 1;
        if (computed || pass == 1) { *si = i; *sj = j; return; }
        i = g->periodic[0] ? i : clampi(i, g->lo[0], g->hi[0]);
        j = g->periodic[1] ? j : clampi(j, g->lo[1], g->hi[1]);
    }
}

/* FillBoundary(periodicity) on one box: every ghost point whose periodic
 * image is a valid point takes its value. */
static void fill_boundary(oa4* a, int t, const ogeom* g)
{
    if (!g->periodic[0] && !g->periodic[1]) { return; }
    const int nx = g->hi[0] - g->lo[0] + 1, ny = g->hi[1] - g->lo[1] + 1;
    const int vh[3] = {g->hi[0] + os_nodal(t, 0), g->hi[1] + os_nodal(t, 1), g->hi[2] + os_nodal(t, 2)};
    for (int k = a->begin[2]; k < a->end[2]; ++k)
        for (int j = a->begin[1]; j < a->end[1]; ++j)
            for (int i = a->begin[0]; i < a->end[0]; ++i) {
                const int in_i = i >= g->lo[0] && i <= vh[0], in_j = j >= g->lo[1] && j <= vh[1];
                const int in_k = k >= g->lo[2] && k <= vh[2];
                if (in_i && in_j && in_k) { continue; }
                int si = i, sj = j;
                if (g->periodic[0] && !in_i) { si = i < g->lo[0] ? i + nx : i - nx; }
                if (g->periodic[1] && !in_j) { sj = j < g->lo[1] ? j + ny : j - ny; }
                if (si < g->lo[0] || si > vh[0] || sj < g->lo[1] || sj > vh[1] || !in_k) { continue; }
                AT(a, i, j, k) = AT(a, si, sj, k);
            }
}

/* The ranges of the ghost slab of face (d, lo) along the other directions
 * (BoundaryConditions.cpp:83-104): the whole fab along directions filled
 * before (e < d) and periodic ones, the domain along later non-periodic ones. */
static void slab_ranges(const oa4* a, int t, const ogeom* g, int d, int rlo[3], int rhi[3])
{
    for (int e = 0; e < 3; ++e) {
        if (e == d) { rlo[e] = rhi[e] = 0; continue; }
        if (e < d || (e < 2 && g->periodic[e])) {
            rlo[e] = a->begin[e];
            rhi[e] = a->end[e] - 1;
        } else {
            rlo[e] = g->lo[e];
            rhi[e] = g->hi[e] + os_nodal(t, e);
        }
    }
}

/* fill_face (BoundaryConditions.cpp:418-470) for the even, odd, value and
 * profile rules; homogeneous: V = 0 and the profile's points 0 (427, 442,
 * 455). */
static void fill_face(oa4* a, int t, int f, int d, int lo, int rule, double V, const oprofile* prof, const oa4* z,
                      const ogeom* g, int homogeneous)
{
    if (homogeneous) { V = 0.0; }
    const int nodal_d = os_nodal(t, d);
    const int edge = lo ? g->lo[d] : (nodal_d ? g->hi[d] + 1 : g->hi[d]);
    const int klo = g->lo[2];
    int rlo[3], rhi[3];
    slab_ranges(a, t, g, d, rlo, rhi);
    const int mirror = nodal_d ? 2 * edge : (lo ? 2 * edge - 1 : 2 * edge + 1);
    const int sgn = rule == R_ODD ? -1 : 1;
    for (int layer = (nodal_d && (rule == R_VALUE || rule == R_PROFILE)) ? 0 : 1; layer <= 2; ++layer) {
        const int pd = lo ? edge - layer : edge + layer;
        int lo3[3] = {rlo[0], rlo[1], rlo[2]}, hi3[3] = {rhi[0], rhi[1], rhi[2]};
        lo3[d] = hi3[d] = pd;
        for (int k = lo3[2]; k <= hi3[2]; ++k)
            for (int j = lo3[1]; j <= hi3[1]; ++j)
                for (int i = lo3[0]; i <= hi3[0]; ++i) {
                    if (rule == R_PROFILE) {
                        AT(a, i, j, k) = homogeneous ? 0.0 : profile_value(prof, f, hag(z, t, i, j, k, klo));
                        continue;
                    }
                    if (layer == 0) { AT(a, i, j, k) = V; continue; }
                    const int p = d == 0 ? i : d == 1 ? j : k;
                    const int m = mirror - p;
                    const int mi = d == 0 ? m : i, mj = d == 1 ? m : j, mk = d == 2 ? m : k;
                    AT(a, i, j, k) = rule == R_VALUE ? 2 * V - AT(a, mi, mj, mk) : sgn * AT(a, mi, mj, mk);
                }
    }
}

/* fill_gradient_z (BoundaryConditions.cpp:472-514): the ground (wall law,
 * per column) or the top (a constant); homogeneous: G = 0 (510). */
static void fill_gradient_z(oa4* a, int t, int f, int lo, int wall, double Gc, const oa4* z, const oa4* ss,
                            const osurface* so, const ogeom* g, int homogeneous)
{
    const int edge = lo ? g->lo[2] : g->hi[2];
    const int c1 = lo ? edge + 1 : edge - 1;
    const int g1 = lo ? edge - 1 : edge + 1, g2 = lo ? edge - 2 : edge + 2;
    const int comp = f == F_U ? 0 : f == F_V ? 1 : 2;
    for (int j = a->begin[1]; j < a->end[1]; ++j)
        for (int i = a->begin[0]; i < a->end[0]; ++i) {
            double G = Gc;
            if (wall) {
                const int ib = os_nodal(t, 0) ? i - 1 : i, jb = os_nodal(t, 1) ? j - 1 : j;
                int ai, aj, bi, bj;
                gg_column(g, i, j, &ai, &aj);
                gg_column(g, ib, jb, &bi, &bj);
                double Ga[3], Gb[3];
                ground_gradient(ss, so, ai, aj, Ga);
                ground_gradient(ss, so, bi, bj, Gb);
                G = 0.5 * (Gb[comp] + Ga[comp]);
            }
            if (homogeneous) { G = 0.0; }
            const double zc1 = point_z(z, t, i, j, c1), zg1 = point_z(z, t, i, j, g1), zg2 = point_z(z, t, i, j, g2);
            AT(a, i, j, g1) = AT(a, i, j, c1) + (zg1 - zc1) * G;
            AT(a, i, j, g2) = AT(a, i, j, g1) + (zg2 - zg1) * G;
        }
}

static void fill_field(oa4* a, int t, int f, int d, const obc* bc, const oa4* z, const oa4* ss, const osurface* so,
                       const ogeom* g, int homogeneous)
{
    for (int lo = 1; lo >= 0; --lo) {
        const int ft = bc->face[face_of(d, lo)];
        const int rule = rule_for(f, ft, d, lo);
        double V = 0;
        if (ft == OFACE_GEOSTROPHIC) { V = f == F_U ? bc->ug : f == F_V ? bc->vg : 0.0; }
        if (ft == OFACE_NO_SLIP_WALL && f <= F_W) { V = bc->wall_velocity[face_of(d, lo)][f]; }
        if (rule == R_GRADIENT) {
            if (ft == OFACE_SURFACE_LAYER) {
                fill_gradient_z(a, t, f, lo, !homogeneous, 0.0, z, ss, so, g, homogeneous);
            } else {
                fill_gradient_z(a, t, f, lo, 0, bc->zhi_theta_grad, z, ss, so, g, homogeneous);
            }
        } else {
            fill_face(a, t, f, d, lo, rule, V, bc->profile, z, g, homogeneous);
        }
    }
}

/* fill_no_flux_w (BoundaryConditions.cpp:756-797): the ground and top faces
 * of w take Omega = 0, w = z_x u_f + z_y v_f, ghosts mirrored about it, on
 * the columns whose u, v points and nodes exist. */
static void fill_no_flux_w(oa4* w, const oa4* u, const oa4* v, const oa4* z, const obc* bc, const ogeom* g)
{
    const double dxinv[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
    int i0 = w->begin[0], i1 = w->end[0] - 1, j0 = w->begin[1], j1 = w->end[1] - 1;
    i0 = os_imax(i0, os_imax(u->begin[0], os_imax(v->begin[0], z->begin[0])));
    i1 = os_imin(i1, os_imin(u->end[0] - 2, os_imin(v->end[0] - 1, z->end[0] - 2)));
    j0 = os_imax(j0, os_imax(u->begin[1], os_imax(v->begin[1], z->begin[1])));
    j1 = os_imin(j1, os_imin(u->end[1] - 1, os_imin(v->end[1] - 2, z->end[1] - 2)));
    for (int lo = 1; lo >= 0; --lo) {
        if (!dirichlet_normal(bc->face[face_of(2, lo)])) { continue; }
        const int edge = lo ? g->lo[2] : g->hi[2] + 1;
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) {
                const double sx = os_deriv(z, OS_NODE, OS_ZF, 0, i, j, edge) * dxinv[0];
                const double sy = os_deriv(z, OS_NODE, OS_ZF, 1, i, j, edge) * dxinv[1];
                const double Vw = sx * os_interp(u, OS_XF, OS_ZF, i, j, edge) + sy * os_interp(v, OS_YF, OS_ZF, i, j, edge);
                AT(w, i, j, edge) = Vw;
                const int sg = lo ? -1 : 1;
                for (int m = 1; m <= 2; ++m) { AT(w, i, j, edge + sg * m) = 2 * Vw - AT(w, i, j, edge - sg * m); }
            }
    }
    fill_boundary(w, OS_ZF, g);
}

static int fill_ghosts_impl(oa4* const q[6], const oa4* z, const obc* bc, const osurface* so, const ogeom* g,
                            oa4* ss, int homogeneous)
{
    const int t[6] = {OS_XF, OS_YF, OS_ZF, OS_CC, OS_CC, OS_CC};
    for (int o = 0; o < 6; ++o) {
        if (bc->face[o] == OFACE_INFLOW && !bc->profile && !homogeneous) { return -3; }
    }
    for (int n = 0; n < 6; ++n) { fill_boundary(q[n], t[n], g); }
    const int wall = bc->face[2] == OFACE_SURFACE_LAYER && !homogeneous;
    for (int d = 0; d < 3; ++d) {
        if (d == 2 && wall) {
            if (oracle_surface_state(q[0], q[1], z, so, g, ss) != 0) { return -1; }
        }
        if (d < 2 && g->periodic[d]) { continue; }
        fill_field(q[0], t[0], F_U, d, bc, z, ss, so, g, homogeneous);
        fill_field(q[1], t[1], F_V, d, bc, z, ss, so, g, homogeneous);
        fill_field(q[2], t[2], F_W, d, bc, z, ss, so, g, homogeneous);
        if (d == 2) { fill_no_flux_w(q[2], q[0], q[1], z, bc, g); }
        for (int n = 3; n < 6; ++n) { fill_field(q[n], t[n], n, d, bc, z, ss, so, g, homogeneous); }
    }
    for (int n = 0; n < 6; ++n) { fill_boundary(q[n], t[n], g); }
    return 0;
}

int oracle_fill_ghosts(oa4* const q[6], const oa4* z, const obc* bc, const osurface* so, const ogeom* g, oa4* ss)
{
    return fill_ghosts_impl(q, z, bc, so, g, ss, 0);
}

/* The fill of a perturbation (BoundaryConditions::fill_ghosts with
 * homogeneous = true): every prescribed value, profile and z gradient 0, no
 * surface state. */
int oracle_fill_ghosts_homogeneous(oa4* const q[6], const oa4* z, const obc* bc, const ogeom* g)
{
    return fill_ghosts_impl(q, z, bc, NULL, g, NULL, 1);
}

/* ---- forcing ---------------------------------------------------------- */

double oracle_coriolis_parameter(int enabled, double latitude, double rotation_period)
{
    if (!enabled) { return 0; }
    const double pi = 3.14159265358979323846;
    return 4 * pi * sin(latitude * pi / 180) / rotation_period;
}

void oracle_fill_theta_ref(int n, const double* zs, const double* ts, const oa4* z, oa4* tr)
{
    for (int k = tr->begin[2]; k < tr->end[2]; ++k)
        for (int j = tr->begin[1]; j < tr->end[1]; ++j)
            for (int i = tr->begin[0]; i < tr->end[0]; ++i) {
                const double zc = os_interp(z, OS_NODE, OS_CC, i, j, k);
                double v = ts[0];
                if (zc >= zs[n - 1]) {
                    v = ts[n - 1];
                } else {
                    for (int m = 1; m < n; ++m) {
                        if (zc < zs[m]) {
                            if (zc > zs[m - 1]) {
                                const double w = (zc - zs[m - 1]) / (zs[m] - zs[m - 1]);
                                v = ts[m - 1] + w * (ts[m] - ts[m - 1]);
                            } else {
                                v = ts[m - 1];
                            }
                            break;
                        }
                    }
                }
                AT(tr, i, j, k) = v;
            }
}

void oracle_forcing(const oa4* u, const oa4* v, const oa4* w, const oa4* theta, const oa4* theta_ref,
                    const oforcing* f, const ogeom* g, oa4* Ru, oa4* Rv, oa4* Rw)
{
    /* subtract_buoyancy: every valid z-face */
    const double gt = f->gravity / f->theta0;
    for (int k = g->lo[2]; k <= g->hi[2] + 1; ++k)
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                AT(Rw, i, j, k) -= gt * 0.5 * ((AT(theta, i, j, k - 1) - AT(theta_ref, i, j, k - 1)) +
                                               (AT(theta, i, j, k) - AT(theta_ref, i, j, k)));
            }
    if (f->coriolis) {
        const double pi = 3.14159265358979323846;
        const double phi = f->latitude * pi / 180;
        const double fc = 4 * pi / f->rotation_period;
        const double fs = fc * sin(phi);
        const double fcos = f->three_d ? fc * cos(phi) : 0.0;
        const double ug = f->geostrophic ? f->ug : 0.0, vg = f->geostrophic ? f->vg : 0.0;
        for (int k = g->lo[2]; k <= g->hi[2]; ++k)
            for (int j = g->lo[1]; j <= g->hi[1]; ++j)
                for (int i = g->lo[0]; i <= g->hi[0] + 1; ++i) {
                    AT(Ru, i, j, k) -= fs * (os_interp(v, OS_YF, OS_XF, i, j, k) - vg) -
                                       fcos * os_interp(w, OS_ZF, OS_XF, i, j, k);
                }
        for (int k = g->lo[2]; k <= g->hi[2]; ++k)
            for (int j = g->lo[1]; j <= g->hi[1] + 1; ++j)
                for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                    AT(Rv, i, j, k) += fs * (os_interp(u, OS_XF, OS_YF, i, j, k) - ug);
                }
        if (f->three_d) {
            for (int k = g->lo[2]; k <= g->hi[2] + 1; ++k)
                for (int j = g->lo[1]; j <= g->hi[1]; ++j)
                    for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                        AT(Rw, i, j, k) -= fcos * os_interp(u, OS_XF, OS_ZF, i, j, k);
                    }
        }
    }
    oa4* R[3] = {Ru, Rv, Rw};
    for (int d = 0; d < 3; ++d) {
        if (f->pgrad[d] == 0) { continue; }
        for (int k = g->lo[2]; k <= g->hi[2] + (d == 2); ++k)
            for (int j = g->lo[1]; j <= g->hi[1] + (d == 1); ++j)
                for (int i = g->lo[0]; i <= g->hi[0] + (d == 0); ++i) { AT(R[d], i, j, k) += -f->pgrad[d]; }
    }
}

/* ---- wall and boundary rows ------------------------------------------- */

void oracle_set_ground_fluxes(const oa4* ss, const ogeom* g, oa4* Gu, oa4* Gv, oa4* Gt, oa4* Gk)
{
    const int k0 = g->lo[2];
    for (int j = g->lo[1]; j <= g->hi[1]; ++j)
        for (int i = g->lo[0]; i <= g->hi[0] + 1; ++i) {
            AT(Gu, i, j, k0) = 0.5 * (SS(ss, OSS_TAU0, i - 1, j) + SS(ss, OSS_TAU0, i, j));
        }
    for (int j = g->lo[1]; j <= g->hi[1] + 1; ++j)
        for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
            AT(Gv, i, j, k0) = 0.5 * (SS(ss, OSS_TAU1, i, j - 1) + SS(ss, OSS_TAU1, i, j));
        }
    for (int j = g->lo[1]; j <= g->hi[1]; ++j)
        for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
            AT(Gt, i, j, k0) = -0.0 * SS(ss, OSS_AREA, i, j);
            AT(Gk, i, j, k0) = 0;
        }
}

void oracle_apply_wall_k(const oa4* ss, const oa4* tke, int dirichlet, const ogeom* g, oa4* Rk)
{
    const int k0 = g->lo[2];
    for (int j = g->lo[1]; j <= g->hi[1]; ++j)
        for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
            const double a = os_max(SS(ss, OSS_USTAR, i, j), kWallKUstarMin) / SS(ss, OSS_D0, i, j) *
                             (AT(tke, i, j, k0) - SS(ss, OSS_KWALL, i, j));
            if (dirichlet) { AT(Rk, i, j, k0) = a; }
            else { AT(Rk, i, j, k0) += a; }
        }
}

void oracle_apply_rows(const oa4* u, const oa4* v, const oa4* w, const oa4* z, const obc* bc, const ogeom* g,
                       oa4* Ru, oa4* Rv, oa4* Rw)
{
    const double dxinv[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
    const int klo = g->lo[2];
    const oa4* qf[3] = {u, v, w};
    oa4* rf[3] = {Ru, Rv, Rw};
    for (int d = 0; d < 3; ++d) {
        if (d < 2 && g->periodic[d]) { continue; }
        const int t = d == 0 ? OS_XF : d == 1 ? OS_YF : OS_ZF;
        for (int lo = 1; lo >= 0; --lo) {
            const int ft = bc->face[face_of(d, lo)];
            if (!dirichlet_normal(ft)) { continue; }
            const int face = lo ? g->lo[d] : g->hi[d] + 1;
            int lo3[3] = {g->lo[0], g->lo[1], g->lo[2]};
            int hi3[3] = {g->hi[0] + os_nodal(t, 0), g->hi[1] + os_nodal(t, 1), g->hi[2] + os_nodal(t, 2)};
            lo3[d] = hi3[d] = face;
            for (int k = lo3[2]; k <= hi3[2]; ++k)
                for (int j = lo3[1]; j <= hi3[1]; ++j)
                    for (int i = lo3[0]; i <= hi3[0]; ++i) {
                        double V = 0;
                        if (d < 2) {
                            if (ft == OFACE_INFLOW) { V = profile_value(bc->profile, d, hag(z, t, i, j, k, klo)); }
                        } else {
                            const double sx = os_deriv(z, OS_NODE, OS_ZF, 0, i, j, face) * dxinv[0];
                            const double sy = os_deriv(z, OS_NODE, OS_ZF, 1, i, j, face) * dxinv[1];
                            V = sx * os_interp(u, OS_XF, OS_ZF, i, j, face) + sy * os_interp(v, OS_YF, OS_ZF, i, j, face);
                        }
                        AT(rf[d], i, j, k) = AT(qf[d], i, j, k) - V;
                    }
        }
    }
}

/* ---- the Rayleigh and sponge layers ------------------------------------- */

static int damping_on(const obc* bc) { return bc->damping.rayleigh_zdamp > 0 || bc->damping.sponge_width > 0; }

static double rayleigh_part(const odamping* d, double z)
{
    const double pi = 3.14159265358979323846;
    if (d->rayleigh_zdamp > 0 && z > d->rayleigh_zdamp) {
        const double s = sin(0.5 * pi * (z - d->rayleigh_zdamp) / (d->ztop - d->rayleigh_zdamp));
        return d->rayleigh_dampcoef * s * s;
    }
    return 0;
}

static double sponge_part(const obc* bc, double x, double y)
{
    const odamping* d = &bc->damping;
    double c = 0;
    if (d->sponge_width > 0) {
        const double inv_w = 1 / d->sponge_width;
        const double dist[4] = {x - d->prob_lo[0], d->prob_hi[0] - x, y - d->prob_lo[1], d->prob_hi[1] - y};
        const int lat[4] = {bc->face[0], bc->face[3], bc->face[1], bc->face[4]};
        for (int n = 0; n < 4; ++n) {
            if ((lat[n] == OFACE_INFLOW || lat[n] == OFACE_OUTFLOW) && dist[n] < d->sponge_width) {
                const double r = 1 - dist[n] * inv_w;
                c += d->sponge_strength * r * r;
            }
        }
    }
    return c;
}

void oracle_add_damping(oa4* const q[6], const oa4* z, const obc* bc, const ogeom* g, oa4* const R[6])
{
    if (!damping_on(bc)) { return; }
    const int klo = g->lo[2];
    const int ff[5] = {F_U, F_V, F_W, F_THETA, F_TKE};
    const int qn[5] = {0, 1, 2, 4, 5};
    for (int n = 0; n < 5; ++n) {
        const int f = ff[n];
        const int t = n < 3 ? (1 << n) : OS_CC;
        const int with_rayleigh = bc->damping.rayleigh_fields[n];
        const int vh[3] = {g->hi[0] + os_nodal(t, 0), g->hi[1] + os_nodal(t, 1), g->hi[2] + os_nodal(t, 2)};
        for (int k = g->lo[2]; k <= vh[2]; ++k)
            for (int j = g->lo[1]; j <= vh[1]; ++j)
                for (int i = g->lo[0]; i <= vh[0]; ++i) {
                    const double h = hag(z, t, i, j, k, klo), zc = point_z(z, t, i, j, k);
                    const double x = bc->damping.prob_lo[0] + (i + (os_nodal(t, 0) ? 0.0 : 0.5)) * g->dx[0];
                    const double y = bc->damping.prob_lo[1] + (j + (os_nodal(t, 1) ? 0.0 : 0.5)) * g->dx[1];
                    const double zr = bc->damping.rayleigh_above_ground ? h : zc;
                    const double cr = with_rayleigh ? rayleigh_part(&bc->damping, zr) : 0.0;
                    const double cs = sponge_part(bc, x, y);
                    const double a = AT(q[qn[n]], i, j, k);
                    if (cr > 0) {
                        const double ref = f == F_W ? 0.0 : profile_value(bc->profile, f, f == F_THETA ? zr : h);
                        AT(R[qn[n]], i, j, k) += cr * (a - ref);
                    }
                    if (cs > 0) {
                        const double ref = f == F_W ? 0.0 : profile_value(bc->profile, f, h);
                        AT(R[qn[n]], i, j, k) += cs * (a - ref);
                    }
                }
    }
}

void oracle_dirichlet_rows(int eq, const obc* bc, const ogeom* g, oa4* mask)
{
    const int t = eq == 0 ? OS_XF : eq == 1 ? OS_YF : eq == 2 ? OS_ZF : OS_CC;
    const int f = eq == 0 ? F_U : eq == 1 ? F_V : eq == 2 ? F_W : eq == 3 ? F_THETA : F_TKE;
    const int vh[3] = {g->hi[0] + os_nodal(t, 0), g->hi[1] + os_nodal(t, 1), g->hi[2] + os_nodal(t, 2)};
    for (int k = g->lo[2]; k <= vh[2]; ++k)
        for (int j = g->lo[1]; j <= vh[1]; ++j)
            for (int i = g->lo[0]; i <= vh[0]; ++i) { AT(mask, i, j, k) = 0; }
    for (int d = 0; d < 3; ++d) {
        if (d < 2 && g->periodic[d]) { continue; }
        for (int lo = 1; lo >= 0; --lo) {
            const int rule = rule_for(f, bc->face[face_of(d, lo)], d, lo);
            if (!(os_nodal(t, d) && (rule == R_VALUE || rule == R_PROFILE))) { continue; }
            const int edge = lo ? g->lo[d] : g->hi[d] + 1;
            int lo3[3] = {g->lo[0], g->lo[1], g->lo[2]}, hi3[3] = {vh[0], vh[1], vh[2]};
            lo3[d] = hi3[d] = edge;
            for (int k = lo3[2]; k <= hi3[2]; ++k)
                for (int j = lo3[1]; j <= hi3[1]; ++j)
                    for (int i = lo3[0]; i <= hi3[0]; ++i) { AT(mask, i, j, k) = 1; }
        }
    }
}

/* The wall model's rows and the z faces' folds of the implicit side A
 * (7 components at the field's location): eq 0..4 = u, v, w, theta, k. */
void oracle_implicit_folds(int eq, const oa4* ss, const oa4* dSdk, const oa4* J, const oa4* zn, const obc* bc,
                           const osurface* so, const ogeom* g, oa4* A)
{
    const int t = eq == 0 ? OS_XF : eq == 1 ? OS_YF : eq == 2 ? OS_ZF : OS_CC;
    const int f = eq == 0 ? F_U : eq == 1 ? F_V : eq == 2 ? F_W : eq == 3 ? F_THETA : F_TKE;
    const int xs = g->hi[0] + os_nodal(t, 0), ys = g->hi[1] + os_nodal(t, 1);
    const int k0 = g->lo[2];
#define AC(i, j, k, n) (A->p[((i) - A->begin[0]) + ((j) - A->begin[1]) * A->jstride + ((k) - A->begin[2]) * A->kstride + (n) * A->nstride])
    if (eq == 4 && dSdk) {
        for (int k = g->lo[2]; k <= g->hi[2]; ++k)
            for (int j = g->lo[1]; j <= g->hi[1]; ++j)
                for (int i = g->lo[0]; i <= g->hi[0]; ++i) { AC(i, j, k, OC_P) += AT(dSdk, i, j, k); }
    }
    if (bc->face[2] == OFACE_SURFACE_LAYER && eq != 2) {
        const double dzinv = 1.0 / g->dx[2];
        for (int j = g->lo[1]; j <= ys; ++j)
            for (int i = g->lo[0]; i <= xs; ++i) {
                AC(i, j, k0, OC_P) -= AC(i, j, k0, OC_ZM);
                AC(i, j, k0, OC_ZM) = 0;
            }
        if (eq == 0 || eq == 1) {
            for (int j = g->lo[1]; j <= ys; ++j)
                for (int i = g->lo[0]; i <= xs; ++i) {
                    const double dg = eq == 0 ? 0.5 * (SS(ss, OSS_DRAG, i - 1, j) + SS(ss, OSS_DRAG, i, j))
                                              : 0.5 * (SS(ss, OSS_DRAG, i, j - 1) + SS(ss, OSS_DRAG, i, j));
                    AC(i, j, k0, OC_P) += dg * dzinv / os_interp(J, OS_CC, t, i, j, k0);
                }
        }
        if (eq == 4) {
            for (int j = g->lo[1]; j <= g->hi[1]; ++j)
                for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                    const double a = os_max(SS(ss, OSS_USTAR, i, j), kWallKUstarMin) / SS(ss, OSS_D0, i, j);
                    if (so->wall_k_dirichlet) {
                        for (int n = 0; n < OC_N; ++n) { AC(i, j, k0, n) = 0; }
                        AC(i, j, k0, OC_P) = a;
                    } else {
                        AC(i, j, k0, OC_P) += a;
                    }
                }
        }
    }
    /* apply_to_implicit (BoundaryConditions.cpp:885-940): the damping layers'
     * diagonal first, then every non-periodic face */
    const int vh[3] = {g->hi[0] + os_nodal(t, 0), g->hi[1] + os_nodal(t, 1), g->hi[2] + os_nodal(t, 2)};
    if (damping_on(bc) && zn) {
        const int with_rayleigh = bc->damping.rayleigh_fields[eq];
        for (int k = g->lo[2]; k <= vh[2]; ++k)
            for (int j = g->lo[1]; j <= vh[1]; ++j)
                for (int i = g->lo[0]; i <= vh[0]; ++i) {
                    const double h = hag(zn, t, i, j, k, k0), zc = point_z(zn, t, i, j, k);
                    const double x = bc->damping.prob_lo[0] + (i + (os_nodal(t, 0) ? 0.0 : 0.5)) * g->dx[0];
                    const double y = bc->damping.prob_lo[1] + (j + (os_nodal(t, 1) ? 0.0 : 0.5)) * g->dx[1];
                    AC(i, j, k, OC_P) += (with_rayleigh ? rayleigh_part(&bc->damping,
                                                                        bc->damping.rayleigh_above_ground ? h : zc)
                                                        : 0.0) + sponge_part(bc, x, y);
                }
    }
    for (int d = 0; d < 3; ++d) {
        if (d < 2 && g->periodic[d]) { continue; }
        for (int lo = 1; lo >= 0; --lo) {
            const int ft = bc->face[face_of(d, lo)];
            if (ft == OFACE_PERIODIC) { continue; }   /* no boundary condition given (residual without bc) */
            const int rule = rule_for(f, ft, d, lo);
            const int nodal_d = os_nodal(t, d);
            const int dirichlet_pt = nodal_d && (rule == R_VALUE || rule == R_PROFILE);
            const int edge = lo ? g->lo[d] : (nodal_d ? g->hi[d] + 1 : g->hi[d]);
            const int cm = d == 0 ? OC_XM : d == 1 ? OC_YM : OC_ZM;
            const int cp = d == 0 ? OC_XP : d == 1 ? OC_YP : OC_ZP;
            const int ghost = lo ? cm : cp, other = lo ? cp : cm;
            double fold = 0;
            int to_other = 0;
            if (nodal_d) { to_other = 1; }
            else if (rule == R_EVEN) { fold = -1; }
            else if (rule == R_ODD || rule == R_VALUE) { fold = 1; }
            else if (rule == R_GRADIENT) { to_other = 1; }
            int lo3[3] = {g->lo[0], g->lo[1], g->lo[2]}, hi3[3] = {vh[0], vh[1], vh[2]};
            lo3[d] = hi3[d] = edge;
            for (int k = lo3[2]; k <= hi3[2]; ++k)
                for (int j = lo3[1]; j <= hi3[1]; ++j)
                    for (int i = lo3[0]; i <= hi3[0]; ++i) {
                        if (dirichlet_pt) {
                            for (int n = 0; n < OC_N; ++n) { AC(i, j, k, n) = 0; }
                            AC(i, j, k, OC_P) = 1;
                        } else {
                            const double ag = AC(i, j, k, ghost);
                            AC(i, j, k, ghost) = 0;
                            if (to_other) { AC(i, j, k, other) += ag; }
                            else { AC(i, j, k, OC_P) += fold * ag; }
                        }
                    }
        }
    }
#undef AC
}
