/*
 * physics.c -- oracle of the residual's physics on a tile with periodic
 * sides: heights above the ground, the k-L closure, the neutral surface
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
 *   Source/BoundaryConditions.cpp:418-534, 540-576, 597-666, 756-797
 *                                       fill_face, fill_gradient_z,
 *                                       ground_gradients, fill_ghosts,
 *                                       fill_no_flux_w (z faces; sides periodic)
 *   Source/BoundaryConditions.cpp:800-850, 878-940
 *                                       apply_rows, apply_to_implicit (z faces)
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

/* ---- ghost fills of the ground and the top (periodic sides) --------- */

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

/* Periodic images along x and y for every cell of a, its z ghosts included
 * (FillBoundary, then the ghost-copy exchange). */
static void periodic_fill(oa4* a, int t, const ogeom* g)
{
    const int nx = g->hi[0] - g->lo[0] + 1, ny = g->hi[1] - g->lo[1] + 1;
    const int vx1 = g->hi[0] + os_nodal(t, 0), vy1 = g->hi[1] + os_nodal(t, 1);
    for (int k = a->begin[2]; k < a->end[2]; ++k)
        for (int j = a->begin[1]; j < a->end[1]; ++j)
            for (int i = a->begin[0]; i < a->end[0]; ++i) {
                if (i >= g->lo[0] && i <= vx1 && j >= g->lo[1] && j <= vy1) { continue; }
                const int si = (i >= g->lo[0] && i <= vx1) ? i : wrap(i, g->lo[0], nx);
                const int sj = (j >= g->lo[1] && j <= vy1) ? j : wrap(j, g->lo[1], ny);
                AT(a, i, j, k) = AT(a, si, sj, k);
            }
}

static void fill_z(oa4* a, int t, int f, const obc* bc, const oa4* z, const oa4* ss, const osurface* so,
                   const double dxinv[3], const oa4* u, const oa4* v, const ogeom* g)
{
    const int nodal_z = os_nodal(t, 2);
    const int xs = g->hi[0] + os_nodal(t, 0), ys = g->hi[1] + os_nodal(t, 1);
    for (int lo = 1; lo >= 0; --lo) {
        const int ft = lo ? bc->zlo : bc->zhi;
        const int rule = rule_for(f, ft, 2, lo);
        double V = 0;
        if (ft == OFACE_GEOSTROPHIC) { V = f == F_U ? bc->ug : f == F_V ? bc->vg : 0.0; }
        if (ft == OFACE_NO_SLIP_WALL && f <= F_W) { V = bc->wall_velocity[lo ? 0 : 1][f]; }
        const int edge = lo ? g->lo[2] : (nodal_z ? g->hi[2] + 1 : g->hi[2]);
        for (int j = g->lo[1]; j <= ys; ++j)
            for (int i = g->lo[0]; i <= xs; ++i) {
                if (rule == R_GRADIENT) {
                    double G;
                    if (ft == OFACE_SURFACE_LAYER) {
                        const int comp = f == F_U ? 0 : f == F_V ? 1 : 2;
                        const int ia = i, ja = j;
                        const int ib = os_nodal(t, 0) ? i - 1 : i, jb = os_nodal(t, 1) ? j - 1 : j;
                        double Ga[3], Gb[3];
                        ground_gradient(ss, so, wrap(ia, g->lo[0], g->hi[0] - g->lo[0] + 1),
                                        wrap(ja, g->lo[1], g->hi[1] - g->lo[1] + 1), Ga);
                        ground_gradient(ss, so, wrap(ib, g->lo[0], g->hi[0] - g->lo[0] + 1),
                                        wrap(jb, g->lo[1], g->hi[1] - g->lo[1] + 1), Gb);
                        G = 0.5 * (Gb[comp] + Ga[comp]);
                    } else {
                        G = bc->zhi_theta_grad;
                    }
                    const int c1 = lo ? edge + 1 : edge - 1;
                    const int g1 = lo ? edge - 1 : edge + 1, g2 = lo ? edge - 2 : edge + 2;
                    const double zc1 = point_z(z, t, i, j, c1), zg1 = point_z(z, t, i, j, g1),
                                 zg2 = point_z(z, t, i, j, g2);
                    AT(a, i, j, g1) = AT(a, i, j, c1) + (zg1 - zc1) * G;
                    AT(a, i, j, g2) = AT(a, i, j, g1) + (zg2 - zg1) * G;
                    continue;
                }
                if (nodal_z && rule == R_VALUE) { AT(a, i, j, edge) = V; }
                const int mirror = nodal_z ? 2 * edge : (lo ? 2 * edge - 1 : 2 * edge + 1);
                const int sgn = rule == R_ODD ? -1 : 1;
                for (int layer = 1; layer <= 2; ++layer) {
                    const int kg = lo ? edge - layer : edge + layer;
                    const int m = mirror - kg;
                    AT(a, i, j, kg) = rule == R_VALUE ? 2 * V - AT(a, i, j, m) : sgn * AT(a, i, j, m);
                }
            }
        if (f == F_W && (ft == OFACE_INFLOW || ft == OFACE_SLIP_WALL || ft == OFACE_NO_SLIP_WALL ||
                         ft == OFACE_SURFACE_LAYER || ft == OFACE_GEOSTROPHIC)) {
            /* fill_no_flux_w: the face takes Omega = 0, ghosts mirrored about it */
            for (int j = g->lo[1]; j <= g->hi[1]; ++j)
                for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                    const double sx = os_deriv(z, OS_NODE, OS_ZF, 0, i, j, edge) * dxinv[0];
                    const double sy = os_deriv(z, OS_NODE, OS_ZF, 1, i, j, edge) * dxinv[1];
                    const double Vw = sx * os_interp(u, OS_XF, OS_ZF, i, j, edge) +
                                      sy * os_interp(v, OS_YF, OS_ZF, i, j, edge);
                    AT(a, i, j, edge) = Vw;
                    const int sg = lo ? -1 : 1;
                    for (int m = 1; m <= 2; ++m) { AT(a, i, j, edge + sg * m) = 2 * Vw - AT(a, i, j, edge - sg * m); }
                }
        }
    }
}

int oracle_fill_ghosts(oa4* const q[6], const oa4* z, const obc* bc, const osurface* so, const ogeom* g, oa4* ss)
{
    if (!g->periodic[0] || !g->periodic[1]) { return -2; }
    const int t[6] = {OS_XF, OS_YF, OS_ZF, OS_CC, OS_CC, OS_CC};
    for (int n = 0; n < 6; ++n) { periodic_fill(q[n], t[n], g); }
    if (bc->zlo == OFACE_SURFACE_LAYER) {
        if (oracle_surface_state(q[0], q[1], z, so, g, ss) != 0) { return -1; }
    }
    const double dxinv[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
    for (int n = 0; n < 6; ++n) {
        fill_z(q[n], t[n], n, bc, z, ss, so, dxinv, q[0], q[1], g);
        periodic_fill(q[n], t[n], g);
    }
    return 0;
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
                       oa4* Rw)
{
    const double dxinv[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
    for (int lo = 1; lo >= 0; --lo) {
        const int ft = lo ? bc->zlo : bc->zhi;
        if (!(ft == OFACE_INFLOW || ft == OFACE_SLIP_WALL || ft == OFACE_NO_SLIP_WALL || ft == OFACE_SURFACE_LAYER ||
              ft == OFACE_GEOSTROPHIC)) {
            continue;
        }
        const int face = lo ? g->lo[2] : g->hi[2] + 1;
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                const double sx = os_deriv(z, OS_NODE, OS_ZF, 0, i, j, face) * dxinv[0];
                const double sy = os_deriv(z, OS_NODE, OS_ZF, 1, i, j, face) * dxinv[1];
                const double V = sx * os_interp(u, OS_XF, OS_ZF, i, j, face) + sy * os_interp(v, OS_YF, OS_ZF, i, j, face);
                AT(Rw, i, j, face) = AT(w, i, j, face) - V;
            }
    }
}

/* The wall model's rows and the z faces' folds of the implicit side A
 * (7 components at the field's location): eq 0..4 = u, v, w, theta, k. */
void oracle_implicit_folds(int eq, const oa4* ss, const oa4* dSdk, const oa4* J, const obc* bc, const osurface* so,
                           const ogeom* g, oa4* A)
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
    if (bc->zlo == OFACE_SURFACE_LAYER && eq != 2) {
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
    /* apply_to_implicit along z (the sides are periodic) */
    for (int lo = 1; lo >= 0; --lo) {
        const int ft = lo ? bc->zlo : bc->zhi;
        const int rule = rule_for(f, ft, 2, lo);
        const int nodal_d = os_nodal(t, 2);
        const int dirichlet_pt = nodal_d && (rule == R_VALUE || rule == R_PROFILE);
        const int edge = lo ? g->lo[2] : (nodal_d ? g->hi[2] + 1 : g->hi[2]);
        const int ghost = lo ? OC_ZM : OC_ZP, other = lo ? OC_ZP : OC_ZM;
        double fold = 0;
        int to_other = 0;
        if (nodal_d) { to_other = 1; }
        else if (rule == R_EVEN) { fold = -1; }
        else if (rule == R_ODD || rule == R_VALUE) { fold = 1; }
        else if (rule == R_GRADIENT) { to_other = 1; }
        for (int j = g->lo[1]; j <= ys; ++j)
            for (int i = g->lo[0]; i <= xs; ++i) {
                if (dirichlet_pt) {
                    for (int n = 0; n < OC_N; ++n) { AC(i, j, edge, n) = 0; }
                    AC(i, j, edge, OC_P) = 1;
                } else {
                    const double ag = AC(i, j, edge, ghost);
                    AC(i, j, edge, ghost) = 0;
                    if (to_other) { AC(i, j, edge, other) += ag; }
                    else { AC(i, j, edge, OC_P) += fold * ag; }
                }
            }
    }
#undef AC
}
