// physics.hip -- the residual R(q) of the steady RANS equations on a tile
// Derived in part from ERF (BSD 3-Clause, Copyright (c) 2022 The Regents of the University of California
// et al.); the full notice is in THIRD_PARTY_NOTICES.md at the repository root.
// with periodic sides, and the implicit rows of its five transported
// fields: heights above the ground, the k-L closure (Axell-Liungman, ERF's
// relations), the neutral wall model, the ghost fills of the ground and the
// top, the buoyancy/Coriolis/pressure-gradient forcing, the wall model's
// ground fluxes and rows, and the boundary rows and folds.
//
//   Source/Residual.cpp:159-256, 272-304   residual, implicit_rows
//   Source/Closure.cpp:132-148, 261-466     copy_vertical_ghosts,
//                                           height_above_ground,
//                                           stratification_and_strain,
//                                           Closure::evaluate
//   Source/Closure_K.H:94-176               switches_for, closure_point
//   Source/SurfaceLayer.cpp:35-217, 430-567 columns<>, set_ground_fluxes,
//                                           apply_wall_k, the wall rows
//   Source/BoundaryConditions.cpp:38-65, 418-666, 756-940
//                                           rules, fills, rows, folds (z faces)
//   Source/Forcing.cpp:77-170, 285-299      theta_ref, buoyancy, Coriolis,
//                                           pressure-gradient force
//   Source/Heights_K.H:20-68                (heights.hpp) NodeZ, point_z<T>
//
// Every point value is formed with the reference's operations in its order
// (contraction off), so the fields equal the FP64 CPU restatement bit for
// bit except where the device's log and pow enter (the wall law's u*, the
// closure's k^1.5), which differ from the host's in the last bit.
#include "cgrid.hpp"
#include "deferred.hpp"
#include "heights.hpp"
#include "profile.hpp"
#include "stencil.hpp"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace gust_amd {

namespace {

namespace sc = stencil_cg;
using cgrid::blocks_for;
using cgrid::Box3;
using cgrid::box3;
using cgrid::check_geom;
using cgrid::CV;
using cgrid::kThreads;
using cgrid::V;

// Workspace slots of this file.
constexpr int kSlotDist = 70, kSlotNu = 71, kSlotGth = 72, kSlotGk = 73, kSlotPs = 74, kSlotPb = 75, kSlotEps = 76,
              kSlotDs = 77, kSlotSS = 78, kSlotTref = 79, kSlotF = 80, kSlotG = 83 /* 15 slots */, kSlotBad = 99,
              kSlotTmp = 100 /* 4 slots */;

// ---- heights -------------------------------------------------------------

using heights::node_z;
using heights::NodeZ;
using heights::point_z;

__global__ __launch_bounds__(kThreads) void height_kernel(Box3 b, CV z, V dist, int klo)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        double zc = 0, zg = 0;
        for (int a = 0; a < 2; ++a)
            for (int c = 0; c < 2; ++c) {
                zg += z(i + a, j + c, klo);
                zc += z(i + a, j + c, k) + z(i + a, j + c, k + 1);
            }
        dist(i, j, k) = 0.125 * zc - 0.25 * zg;
    }
}

// The ghost layers below klo and above khi take the first and last cells.
__global__ __launch_bounds__(kThreads) void copy_vertical_kernel(Box3 plane, V a, int klo, int khi)
{
    GUST_GRID_LOOP(plane)
    {
        int i, j, k;
        plane.at(t, i, j, k);
        a(i, j, klo - 1) = a(i, j, klo);
        a(i, j, khi + 1) = a(i, j, khi);
    }
}

// ---- the k-L closure -------------------------------------------------------

struct KConst {
    double kappa, Cmu0, Cmu0_pow3, inv_Cb_sq, Rt_crit, Rt_min, l_g_max, tke_floor, n2_eps;
    double nu_mol, alpha_mol, inv_sigma;
};

__device__ __forceinline__ double pos_part(double x) { return 0.5 * (x + fabs(x)); }
__device__ __forceinline__ double neg_part(double x) { return 0.5 * (x - fabs(x)); }

__device__ __forceinline__ double smooth_rt(double Rt, bool smooth, double Rt_crit, double Rt_min)
{
    const double x = neg_part(Rt - Rt_crit);
    const double a = Rt_min - Rt_crit;
    const double s = Rt_crit + a * x / (x + a);
    return smooth ? s : Rt;
}

struct KPoint {
    double k, L, eps, cmu, cmu_prime, nut;
};

// Closure_K.H:94-176: the decisions from the primal values, then the point.
__device__ __forceinline__ KPoint closure_point(double kin, double N2, double l_g, const KConst& c)
{
    const bool floored = !(kin > c.tke_floor);
    const double k = floored ? c.tke_floor : kin;
    const int branch = (fabs(N2) <= c.n2_eps) ? 0 : (N2 > c.n2_eps) ? 1 : 2;
    const double c6 = c.Cmu0_pow3 * c.Cmu0_pow3;
    const double N2s = pos_part(N2), N2u = neg_part(N2);
    // switches_for (with kf = k)
    const double Rt_g0 = l_g * l_g * N2u / (k * c.Cmu0_pow3 * c.Cmu0_pow3);
    const bool smooth_g = (branch == 2) && (Rt_g0 < c.Rt_crit);
    const double Lu0 = l_g * sqrt(1.0 - c6 * c.inv_Cb_sq * smooth_rt(Rt_g0, smooth_g, c.Rt_crit, c.Rt_min));
    const double Ls0 = sqrt(1.0 / (1.0 / (l_g * l_g) + c.inv_Cb_sq * N2s / k));
    const double L0 = (branch == 0) ? l_g : (branch == 1) ? Ls0 : Lu0;
    const bool smooth = (L0 * L0 * N2 / (k * c.Cmu0_pow3 * c.Cmu0_pow3) < c.Rt_crit);
    // closure_point
    KPoint p;
    p.k = k;
    const double L_stable = sqrt(1.0 / ((1.0 / (l_g * l_g)) + (c.inv_Cb_sq * N2s) / k));
    const double Rt_g = (l_g * l_g * N2u) / (k * c.Cmu0_pow3 * c.Cmu0_pow3);
    const double L_unstable = l_g * sqrt(1.0 - (c6 * c.inv_Cb_sq) * smooth_rt(Rt_g, smooth_g, c.Rt_crit, c.Rt_min));
    p.L = (branch == 0) ? l_g : (branch == 1) ? L_stable : L_unstable;
    p.eps = c.Cmu0_pow3 * pow(k, 1.5) / p.L;
    const double Rt = smooth_rt(p.L * p.L * N2 / (k * c.Cmu0_pow3 * c.Cmu0_pow3), smooth, c.Rt_crit, c.Rt_min);
    p.cmu = (c.Cmu0 + 0.108 * Rt) / (1.0 + 0.308 * Rt + 0.00837 * Rt * Rt);
    p.cmu_prime = c.Cmu0 / (1.0 + 0.277 * Rt);
    p.nut = p.cmu * sqrt(k) * p.L;
    return p;
}

// grad_at<TQ, CC> (Stencil_K.H:115-124).
template <int TQ>
__device__ __forceinline__ void grad_cc(const CV& q, int i, int j, int k, const sc::Metric& m, const double (&dxinv)[3],
                                        double g[3])
{
    const double qx = sc::deriv<TQ, sc::CC, 0>(q, i, j, k) * dxinv[0];
    const double qy = sc::deriv<TQ, sc::CC, 1>(q, i, j, k) * dxinv[1];
    const double qz = sc::deriv<TQ, sc::CC, 2>(q, i, j, k) * dxinv[2] / m.zz;
    g[0] = qx - m.zx * qz;
    g[1] = qy - m.zy * qz;
    g[2] = qz;
}

struct ClosureOut {
    V nu, gth, gk, Ps, Pb, eps, dSdk;
};

__global__ __launch_bounds__(kThreads) void closure_kernel(Box3 b, CV u, CV v, CV w, CV th, CV tke, CV z, CV dist,
                                                          double z0, KConst c, double gt0, double dx0, double dx1,
                                                          double dx2, ClosureOut o, int vlo0, int vhi0, int vlo1,
                                                          int vhi1)
{
    const double dxinv[3] = {dx0, dx1, dx2};
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const sc::Metric m = sc::metric_at<sc::CC>(z, i, j, k, dxinv);
        double G[4][3];
        grad_cc<sc::CC>(th, i, j, k, m, dxinv, G[0]);
        grad_cc<sc::XF>(u, i, j, k, m, dxinv, G[1]);
        grad_cc<sc::YF>(v, i, j, k, m, dxinv, G[2]);
        grad_cc<sc::ZF>(w, i, j, k, m, dxinv, G[3]);
        const double N2 = gt0 * G[0][2];
        double s = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const double x = G[1 + a][e] + G[1 + e][a];
                s += x * x;
            }
        const double S2 = 0.5 * s;
        const double raw = c.kappa * (dist(i, j, k) + z0);
        const double l_g = c.l_g_max * raw / (c.l_g_max + raw);
        const KPoint p = closure_point(tke(i, j, k), N2, l_g, c);
        const double ratio = p.cmu_prime / p.cmu;
        const double nt = 1.0 * sc::dmin(p.nut, DBL_MAX) + (1.0 - 1.0) * tke(i, j, k);
        o.nu(i, j, k) = c.nu_mol + nt;
        o.gth(i, j, k) = c.alpha_mol + ratio * nt;
        o.gk(i, j, k) = c.nu_mol + nt * c.inv_sigma;
        if (i >= vlo0 && i <= vhi0 && j >= vlo1 && j <= vhi1) {
            const double ps = nt * S2, pb = -ratio * nt * N2;
            o.Ps(i, j, k) = ps;
            o.Pb(i, j, k) = pb;
            o.eps(i, j, k) = p.eps;
            o.dSdk(i, j, k) = (p.eps + pos_part(-pb)) / p.k;
        }
    }
}

// ---- the neutral wall model -------------------------------------------------

struct SSV {
    double* p;
    int64_t js, ns;
    int b0, b1;
    __device__ __forceinline__ double& operator()(int n, int i, int j) const
    {
        return p[int64_t(i - b0) + int64_t(j - b1) * js + int64_t(n) * ns];
    }
};

SSV ss_view(const gust_array4* a)
{
    return SSV{static_cast<double*>(a->p), a->jstride, a->nstride, a->begin[0], a->begin[1]};
}

struct SurfOpt {
    double z0, kappa, inv_cmu2;
    int floor_one;
};

constexpr double kWsMin = 0.1, kWallKUstarMin = 0.01;

__global__ __launch_bounds__(kThreads) void surface_kernel(Box3 cols, CV u, CV v, CV z, SurfOpt o, double dxinv0,
                                                          double dxinv1, SSV ss, int klo, unsigned* bad)
{
    unsigned nbad = 0;
    GUST_GRID_LOOP(cols)
    {
        int i, j, kk;
        cols.at(t, i, j, kk);
        const double ground = 0.25 * ((z(i, j, klo) + z(i + 1, j, klo)) + (z(i, j + 1, klo) + z(i + 1, j + 1, klo)));
        const double h0 = sc::interp<sc::NODE, sc::CC>(z, i, j, klo) - ground;
        const double zx = sc::deriv<sc::NODE, sc::ZF, 0>(z, i, j, klo) * dxinv0;
        const double zy = sc::deriv<sc::NODE, sc::ZF, 1>(z, i, j, klo) * dxinv1;
        const double area = sqrt(1 + zx * zx + zy * zy);
        const double n[3] = {-zx / area, -zy / area, 1 / area};
        const double uc = 0.5 * (u(i, j, klo) + u(i + 1, j, klo));
        const double vc = 0.5 * (v(i, j, klo) + v(i, j + 1, klo));
        const double U3[3] = {uc, vc, zx * uc + zy * vc};
        const double un = U3[0] * n[0] + U3[1] * n[1] + U3[2] * n[2];
        double ut[3];
        for (int d = 0; d < 3; ++d) { ut[d] = U3[d] - un * n[d]; }
        // The reference point is the first centre: the interpolation in
        // ln(d + z0) between the first two centres has the weight 0.
        const double ia = 1.0 / area;
        const double mag = sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
        const double U = sc::dmax(mag, kWsMin);
        const double dd = h0 * ia;
        const double lz = log((dd + o.z0) / o.z0);
        const double den_min = o.floor_one ? 1.0 : sc::dmin(lz, 1.0);
        const double ustar = o.kappa * U / sc::dmax(lz, den_min);
        const double cdrag = ustar * ustar / U * area;
        ss(GUST_SS_USTAR, i, j) = ustar;
        ss(GUST_SS_SPEED, i, j) = U;
        ss(GUST_SS_D0, i, j) = h0 * ia;
        ss(GUST_SS_TAU0, i, j) = cdrag * ut[0];
        ss(GUST_SS_TAU1, i, j) = cdrag * ut[1];
        ss(GUST_SS_TAU2, i, j) = cdrag * ut[2];
        ss(GUST_SS_DRAG, i, j) = cdrag;
        const double kw = ustar * ustar * o.inv_cmu2;
        ss(GUST_SS_KWALL, i, j) = kw;
        ss(GUST_SS_AREA, i, j) = area;
        ss(GUST_SS_INVL, i, j) = 0.0;
        if (ustar != ustar || kw != kw || !(ustar > 0 && ustar < DBL_MAX)) { ++nbad; }
    }
    if (nbad) { atomicAdd(bad, nbad); }
}

// ---- ghost fills --------------------------------------------------------------

// R_ZERO: the homogeneous fill's profile rule (the face point and every
// ghost layer 0, BoundaryConditions.cpp:438-457 with homogeneous set).
enum : int { R_EVEN = 0, R_ODD, R_VALUE, R_PROFILE, R_GRADIENT, R_ZERO };
enum : int { F_U = 0, F_V, F_W, F_P, F_THETA, F_TKE };

int rule_for(int f, int t, int d, int lo)
{
    const bool normal_vel = (f == F_U && d == 0) || (f == F_V && d == 1) || (f == F_W && d == 2);
    const bool vel = f == F_U || f == F_V || f == F_W;
    switch (t) {
    case GUST_FACE_INFLOW: return f == F_P ? R_EVEN : R_PROFILE;
    case GUST_FACE_OUTFLOW: return f == F_P ? R_ODD : R_EVEN;
    case GUST_FACE_SLIP_WALL:
        if (normal_vel) { return R_VALUE; }
        return (f == F_THETA && d == 2 && !lo) ? R_GRADIENT : R_EVEN;
    case GUST_FACE_NO_SLIP_WALL: return vel ? R_VALUE : R_EVEN;
    case GUST_FACE_SURFACE_LAYER:
        if (normal_vel) { return R_VALUE; }
        return (vel || f == F_THETA) ? R_GRADIENT : R_EVEN;
    case GUST_FACE_GEOSTROPHIC:
        if (normal_vel) { return R_VALUE; }
        if (f == F_U || f == F_V) { return R_VALUE; }
        return f == F_THETA ? R_GRADIENT : R_EVEN;
    default: return R_EVEN;
    }
}

bool dirichlet_normal(int t)
{
    return t == GUST_FACE_INFLOW || t == GUST_FACE_SLIP_WALL || t == GUST_FACE_NO_SLIP_WALL ||
           t == GUST_FACE_SURFACE_LAYER || t == GUST_FACE_GEOSTROPHIC;
}

struct FaceFill {
    int rule;
    double V;       // value
    double G;       // constant gradient (gradient without the wall model)
    int wall;       // gradient from the wall law (comp)
    int comp;
};

// The wall law's gradient of column (i, j) (BoundaryConditions.cpp:540-576).
__device__ __forceinline__ double ground_gradient(const SSV& ss, const SurfOpt& o, int comp, int i, int j)
{
    const double zeta = sc::dmax(sc::dmin(ss(GUST_SS_D0, i, j) * ss(GUST_SS_INVL, i, j), 100.0), -100.0);
    const double az = fabs(zeta);
    const double zs = 0.5 * (az + zeta), zu = 0.5 * (zeta - az);
    const double r = sqrt(1.0 - 16.0 * zu);
    const double phi_m = (1.0 + 5.0 * zs) / (1.0 + 0.0 * zs) + (1.0 / sqrt(r) - 1.0);
    const double phi_h = (1.0 + 5.0 * zs) / (1.0 + 0.0 * zs) + (1.0 / r - 1.0);
    const double s = o.kappa * (ss(GUST_SS_D0, i, j) + o.z0);
    const double t0 = ss(GUST_SS_TAU0, i, j), t1 = ss(GUST_SS_TAU1, i, j), t2 = ss(GUST_SS_TAU2, i, j);
    const double tmag = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
    const double tsafe = tmag + DBL_MIN;
    const double us = ss(GUST_SS_USTAR, i, j);
    if (comp == 0) { return us / s * (t0 / tsafe) * phi_m; }
    if (comp == 1) { return us / s * (t1 / tsafe) * phi_m; }
    return -0.0 / (us * s) * phi_h;
}

__device__ __forceinline__ int wrap(int i, int lo, int n) { return lo + ((i - lo) % n + n) % n; }
__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

// The surface-state column the ground gradient of column (i, j) comes from
// (BoundaryConditions.cpp:617-647): computed on the ground columns and one
// ghost column; FillBoundary gives each ghost column whose periodic image is
// a ground column that image; a column still unset takes the one clamped
// into the domain along the non-periodic directions.
struct GgMap {
    int lo0, hi0, lo1, hi1, per0, per1, nx, ny;
};

__device__ __forceinline__ void gg_column(const GgMap& m, int i, int j, int& si, int& sj)
{
    for (int pass = 0; pass < 2; ++pass) {
        if (i >= m.lo0 && i <= m.hi0 && j >= m.lo1 && j <= m.hi1) {
            si = i;
            sj = j;
            return;
        }
        const int ii = (m.per0 && (i < m.lo0 || i > m.hi0)) ? wrap(i, m.lo0, m.nx) : i;
        const int jj = (m.per1 && (j < m.lo1 || j > m.hi1)) ? wrap(j, m.lo1, m.ny) : j;
        if (ii >= m.lo0 && ii <= m.hi0 && jj >= m.lo1 && jj <= m.hi1) {
            si = ii;
            sj = jj;
            return;
        }
        const bool computed = i >= m.lo0 - 1 && i <= m.hi0 + 1 && j >= m.lo1 - 1 && j <= m.hi1 + 1;
        if (computed || pass == 1) {
            si = i;
            sj = j;
            return;
        }
        i = m.per0 ? i : clampi(i, m.lo0, m.hi0);
        j = m.per1 ? j : clampi(j, m.lo1, m.hi1);
    }
}

template <int T>
__device__ __forceinline__ void fill_z_columns(const Box3& cols, const V& a, const NodeZ& z, int edge_lo, int edge_hi,
                                               const FaceFill& lo, const FaceFill& hi, const SSV& ss,
                                               const SurfOpt& so, const GgMap& gm)
{
    GUST_GRID_LOOP(cols)
    {
        int i, j, kk;
        cols.at(t, i, j, kk);
        for (int side = 0; side < 2; ++side) {
            const FaceFill& f = side == 0 ? lo : hi;
            const bool is_lo = side == 0;
            const int edge = is_lo ? edge_lo : edge_hi;
            if (f.rule == R_GRADIENT) {
                double G = f.G;
                if (f.wall) {
                    const int ib = sc::nodal(T, 0) ? i - 1 : i, jb = sc::nodal(T, 1) ? j - 1 : j;
                    int ai, aj, bi, bj;
                    gg_column(gm, i, j, ai, aj);
                    gg_column(gm, ib, jb, bi, bj);
                    const double Ga = ground_gradient(ss, so, f.comp, ai, aj);
                    const double Gb = ground_gradient(ss, so, f.comp, bi, bj);
                    G = 0.5 * (Gb + Ga);
                }
                const int c1 = is_lo ? edge + 1 : edge - 1;
                const int g1 = is_lo ? edge - 1 : edge + 1, g2 = is_lo ? edge - 2 : edge + 2;
                const double zc1 = point_z<T>(z, i, j, c1), zg1 = point_z<T>(z, i, j, g1), zg2 = point_z<T>(z, i, j, g2);
                a(i, j, g1) = a(i, j, c1) + (zg1 - zc1) * G;
                a(i, j, g2) = a(i, j, g1) + (zg2 - zg1) * G;
                continue;
            }
            constexpr bool nodal_z = sc::nodal(T, 2);
            if (nodal_z && f.rule == R_VALUE) { a(i, j, edge) = f.V; }
            const int mirror = nodal_z ? 2 * edge : (is_lo ? 2 * edge - 1 : 2 * edge + 1);
            const double sgn = f.rule == R_ODD ? -1.0 : 1.0;
            for (int layer = 1; layer <= 2; ++layer) {
                const int kg = is_lo ? edge - layer : edge + layer;
                const int m = mirror - kg;
                a(i, j, kg) = f.rule == R_VALUE ? 2 * f.V - a(i, j, m) : sgn * a(i, j, m);
            }
        }
    }
}

// The z-face fills of up to six fields in one launch (blockIdx.y: the field;
// the fields are independent of one another).
struct FzField {
    V a;
    Box3 cols;
    int loc, edge_lo, edge_hi;
    FaceFill lo, hi;
};
struct FzSet {
    FzField f[6];
};

__global__ __launch_bounds__(kThreads) void fill_z_kernel(FzSet s, NodeZ z, SSV ss, SurfOpt so, GgMap gm)
{
    const FzField& F = s.f[blockIdx.y];
    switch (F.loc) {
    case sc::XF: fill_z_columns<sc::XF>(F.cols, F.a, z, F.edge_lo, F.edge_hi, F.lo, F.hi, ss, so, gm); break;
    case sc::YF: fill_z_columns<sc::YF>(F.cols, F.a, z, F.edge_lo, F.edge_hi, F.lo, F.hi, ss, so, gm); break;
    case sc::ZF: fill_z_columns<sc::ZF>(F.cols, F.a, z, F.edge_lo, F.edge_hi, F.lo, F.hi, ss, so, gm); break;
    default: fill_z_columns<sc::CC>(F.cols, F.a, z, F.edge_lo, F.edge_hi, F.lo, F.hi, ss, so, gm); break;
    }
}

// w = z_x u_f + z_y v_f on the ground and top faces (Omega = 0), ghosts
// mirrored about it (BoundaryConditions.cpp:756-797; MassFlux_K.H:31-39).
__global__ __launch_bounds__(kThreads) void no_flux_w_kernel(Box3 cols, V w, CV u, CV v, CV z, int edge_lo,
                                                            int edge_hi, int do_lo, int do_hi, double dx0, double dx1)
{
    GUST_GRID_LOOP(cols)
    {
        int i, j, kk;
        cols.at(t, i, j, kk);
        for (int side = 0; side < 2; ++side) {
            if (side == 0 ? !do_lo : !do_hi) { continue; }
            const int edge = side == 0 ? edge_lo : edge_hi;
            const double sx = sc::deriv<sc::NODE, sc::ZF, 0>(z, i, j, edge) * dx0;
            const double sy = sc::deriv<sc::NODE, sc::ZF, 1>(z, i, j, edge) * dx1;
            const double Vw = sx * sc::interp<sc::XF, sc::ZF>(u, i, j, edge) + sy * sc::interp<sc::YF, sc::ZF>(v, i, j, edge);
            w(i, j, edge) = Vw;
            const int sg = side == 0 ? -1 : 1;
            for (int m = 1; m <= 2; ++m) { w(i, j, edge + sg * m) = 2 * Vw - w(i, j, edge - sg * m); }
        }
    }
}

// fill_face along x or y (BoundaryConditions.cpp:418-470): the even, odd,
// value and profile rules on both faces of direction d, one thread per point
// of the ghost slab's plane (the plane's own index along d is unused).
struct SideFill {
    int rule;
    double V;
};

template <int T>
__device__ __forceinline__ void fill_side_plane(const Box3& plane, const V& a, int d, int edge_lo, int edge_hi,
                                                const SideFill& flo, const SideFill& fhi, const NodeZ& z,
                                                const profile::ProfileView& prof, int f, int klo)
{
    GUST_GRID_LOOP(plane)
    {
        int i, j, k;
        plane.at(t, i, j, k);
        const bool nodal_d = ((T >> d) & 1) != 0;
        for (int side = 0; side < 2; ++side) {
            const SideFill& s = side == 0 ? flo : fhi;
            const bool is_lo = side == 0;
            const int edge = is_lo ? edge_lo : edge_hi;
            const int mirror = nodal_d ? 2 * edge : (is_lo ? 2 * edge - 1 : 2 * edge + 1);
            const double sgn = s.rule == R_ODD ? -1.0 : 1.0;
            const int first = (nodal_d && (s.rule == R_VALUE || s.rule == R_PROFILE || s.rule == R_ZERO)) ? 0 : 1;
            for (int layer = first; layer <= 2; ++layer) {
                const int pd = is_lo ? edge - layer : edge + layer;
                const int pi = d == 0 ? pd : i, pj = d == 1 ? pd : j;
                if (s.rule == R_ZERO) {
                    a(pi, pj, k) = 0.0;
                    continue;
                }
                if (s.rule == R_PROFILE) {
                    double u, v, th, kk, nut;
                    profile::profile_eval(prof, heights::height_above_ground<T>(z, pi, pj, k, klo), u, v, th, kk,
                                          nut);
                    a(pi, pj, k) = f == F_U ? u : f == F_V ? v : f == F_THETA ? th : f == F_TKE ? kk : 0.0;
                    continue;
                }
                if (layer == 0) {
                    a(pi, pj, k) = s.V;
                    continue;
                }
                const int m = mirror - pd;
                const int mi = d == 0 ? m : i, mj = d == 1 ? m : j;
                a(pi, pj, k) = s.rule == R_VALUE ? 2 * s.V - a(mi, mj, k) : sgn * a(mi, mj, k);
            }
        }
    }
}

// The side fills of the six fields along one direction in one launch
// (blockIdx.y: the field; the fields are independent of one another).
struct SideField {
    V a;
    Box3 plane;
    int loc, f, edge_lo, edge_hi;
    SideFill lo, hi;
};
struct SideSet {
    SideField f[6];
};

__global__ __launch_bounds__(kThreads) void fill_side_kernel(SideSet s, int d, NodeZ z, profile::ProfileView prof,
                                                            int klo)
{
    const SideField& F = s.f[blockIdx.y];
    switch (F.loc) {
    case sc::XF: fill_side_plane<sc::XF>(F.plane, F.a, d, F.edge_lo, F.edge_hi, F.lo, F.hi, z, prof, F.f, klo); break;
    case sc::YF: fill_side_plane<sc::YF>(F.plane, F.a, d, F.edge_lo, F.edge_hi, F.lo, F.hi, z, prof, F.f, klo); break;
    case sc::ZF: fill_side_plane<sc::ZF>(F.plane, F.a, d, F.edge_lo, F.edge_hi, F.lo, F.hi, z, prof, F.f, klo); break;
    default: fill_side_plane<sc::CC>(F.plane, F.a, d, F.edge_lo, F.edge_hi, F.lo, F.hi, z, prof, F.f, klo); break;
    }
}

// FillBoundary(periodicity) on one tile: a ghost point whose periodic image
// is a valid point takes its value.
// Up to six fields per launch (blockIdx.y: the field).
struct FbField {
    V a;
    Box3 fab;
    int vlo0, vhi0, vlo1, vhi1, vlo2, vhi2;
};
struct FbSet {
    FbField f[6];
};

__global__ __launch_bounds__(kThreads) void fill_boundary_kernel(FbSet s, int per0, int per1, int nx, int ny)
{
    const FbField& F = s.f[blockIdx.y];
    const Box3& fab = F.fab;
    const V& a = F.a;
    const int vlo0 = F.vlo0, vhi0 = F.vhi0, vlo1 = F.vlo1, vhi1 = F.vhi1, vlo2 = F.vlo2, vhi2 = F.vhi2;
    GUST_GRID_LOOP(fab)
    {
        int i, j, k;
        fab.at(t, i, j, k);
        const bool in_i = i >= vlo0 && i <= vhi0, in_j = j >= vlo1 && j <= vhi1, in_k = k >= vlo2 && k <= vhi2;
        if ((in_i && in_j && in_k) || !in_k) { continue; }
        int si = i, sj = j;
        if (per0 && !in_i) { si = i < vlo0 ? i + nx : i - nx; }
        if (per1 && !in_j) { sj = j < vlo1 ? j + ny : j - ny; }
        if (si < vlo0 || si > vhi0 || sj < vlo1 || sj > vhi1) { continue; }
        a(i, j, k) = a(si, sj, k);
    }
}

// ---- forcing ------------------------------------------------------------------

using heights::Sounding;
using heights::theta_ref_kernel;

struct Rot {
    int on, three_d;
    double fs, fcos, ug, vg;
    double pg[3];
};

__global__ __launch_bounds__(kThreads) void force_u_kernel(Box3 b, V r, CV v, CV w, Rot f)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        if (f.on) { r(i, j, k) -= f.fs * (sc::interp<sc::YF, sc::XF>(v, i, j, k) - f.vg) - f.fcos * sc::interp<sc::ZF, sc::XF>(w, i, j, k); }
        if (f.pg[0] != 0) { r(i, j, k) += -f.pg[0]; }
    }
}

__global__ __launch_bounds__(kThreads) void force_v_kernel(Box3 b, V r, CV u, Rot f)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        if (f.on) { r(i, j, k) += f.fs * (sc::interp<sc::XF, sc::YF>(u, i, j, k) - f.ug); }
        if (f.pg[1] != 0) { r(i, j, k) += -f.pg[1]; }
    }
}

__global__ __launch_bounds__(kThreads) void force_w_kernel(Box3 b, V r, CV th, CV tr, CV u, double gt, Rot f)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        r(i, j, k) -= gt * 0.5 * ((th(i, j, k - 1) - tr(i, j, k - 1)) + (th(i, j, k) - tr(i, j, k)));
        if (f.on && f.three_d) { r(i, j, k) -= f.fcos * sc::interp<sc::XF, sc::ZF>(u, i, j, k); }
        if (f.pg[2] != 0) { r(i, j, k) += -f.pg[2]; }
    }
}

// ---- residual rows -----------------------------------------------------------

// R_k -= P_s; R_k -= P_b; R_k += eps (Residual.cpp:242-245).
__global__ __launch_bounds__(kThreads) void k_sources_kernel(Box3 b, V r, CV ps, CV pb, CV eps)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        double x = r(i, j, k);
        x -= ps(i, j, k);
        x -= pb(i, j, k);
        x += eps(i, j, k);
        r(i, j, k) = x;
    }
}

// set_ground_fluxes (SurfaceLayer.cpp:430-464) on the ground faces.
__global__ __launch_bounds__(kThreads) void ground_flux_kernel(Box3 cols, SSV ss, V gu, V gv, V gt, V gk, int k0,
                                                              int hi0, int hi1)
{
    GUST_GRID_LOOP(cols)
    {
        int i, j, kk;
        cols.at(t, i, j, kk);   // i in lo..hi0+1, j in lo..hi1+1
        if (j <= hi1) { gu(i, j, k0) = 0.5 * (ss(GUST_SS_TAU0, i - 1, j) + ss(GUST_SS_TAU0, i, j)); }
        if (i <= hi0) { gv(i, j, k0) = 0.5 * (ss(GUST_SS_TAU1, i, j - 1) + ss(GUST_SS_TAU1, i, j)); }
        if (i <= hi0 && j <= hi1) {
            gt(i, j, k0) = -0.0 * ss(GUST_SS_AREA, i, j);
            gk(i, j, k0) = 0;
        }
    }
}

__global__ __launch_bounds__(kThreads) void wall_k_kernel(Box3 cols, SSV ss, CV tke, V r, int k0, int dirichlet)
{
    GUST_GRID_LOOP(cols)
    {
        int i, j, kk;
        cols.at(t, i, j, kk);
        const double a = sc::dmax(ss(GUST_SS_USTAR, i, j), kWallKUstarMin) / ss(GUST_SS_D0, i, j) *
                         (tke(i, j, k0) - ss(GUST_SS_KWALL, i, j));
        if (dirichlet) { r(i, j, k0) = a; }
        else { r(i, j, k0) += a; }
    }
}

// R_w = w - the no-flux value on the ground and the top (BoundaryConditions.cpp:800-850).
__global__ __launch_bounds__(kThreads) void w_rows_kernel(Box3 cols, V r, CV w, CV u, CV v, CV z, int edge_lo,
                                                         int edge_hi, int do_lo, int do_hi, double dx0, double dx1)
{
    GUST_GRID_LOOP(cols)
    {
        int i, j, kk;
        cols.at(t, i, j, kk);
        for (int side = 0; side < 2; ++side) {
            if (side == 0 ? !do_lo : !do_hi) { continue; }
            const int face = side == 0 ? edge_lo : edge_hi;
            const double sx = sc::deriv<sc::NODE, sc::ZF, 0>(z, i, j, face) * dx0;
            const double sy = sc::deriv<sc::NODE, sc::ZF, 1>(z, i, j, face) * dx1;
            const double V0 = sx * sc::interp<sc::XF, sc::ZF>(u, i, j, face) + sy * sc::interp<sc::YF, sc::ZF>(v, i, j, face);
            r(i, j, face) = w(i, j, face) - V0;
        }
    }
}

// R = q - V on the x (d = 0, R_u) or y (d = 1, R_v) faces with a Dirichlet
// normal velocity: V the inflow profile's at the point's height, else 0
// (BoundaryConditions.cpp:800-834).
template <int T>
__global__ __launch_bounds__(kThreads) void side_rows_kernel(Box3 plane, V r, CV q, int d, int face_lo, int face_hi,
                                                            int do_lo, int do_hi, int prof_lo, int prof_hi, NodeZ z,
                                                            profile::ProfileView prof, int klo)
{
    GUST_GRID_LOOP(plane)
    {
        int i, j, k;
        plane.at(t, i, j, k);
        for (int side = 0; side < 2; ++side) {
            if (side == 0 ? !do_lo : !do_hi) { continue; }
            const int face = side == 0 ? face_lo : face_hi;
            const int pi = d == 0 ? face : i, pj = d == 1 ? face : j;
            double Vd = 0.0;
            if (side == 0 ? prof_lo : prof_hi) {
                double u, v, th, kk, nut;
                profile::profile_eval(prof, heights::height_above_ground<T>(z, pi, pj, k, klo), u, v, th, kk, nut);
                Vd = d == 0 ? u : v;
            }
            r(pi, pj, k) = q(pi, pj, k) - Vd;
        }
    }
}

// ---- the Rayleigh and sponge layers (BoundaryConditions.cpp:942-1101) -------------

struct DampCoef {
    int rayleigh, sponge;
    double zd, rc, ztop, sw, inv_w, ss;
    double plo[2], phi[2];
    int lateral[4];   // xlo, xhi, ylo, yhi hold an inflow or outflow face
    __device__ __forceinline__ double rayleigh_part(double z) const
    {
        const double pi = 3.14159265358979323846;
        if (rayleigh && z > zd) {
            const double s = sin(0.5 * pi * (z - zd) / (ztop - zd));
            return rc * s * s;
        }
        return 0.0;
    }
    __device__ __forceinline__ double sponge_part(double x, double y) const
    {
        double c = 0.0;
        if (sponge) {
            const double dist[4] = {x - plo[0], phi[0] - x, y - plo[1], phi[1] - y};
            for (int n = 0; n < 4; ++n) {
                if (lateral[n] && dist[n] < sw) {
                    const double r = 1 - dist[n] * inv_w;
                    c += ss * r * r;
                }
            }
        }
        return c;
    }
};

// R += c (q - q_ref) (BoundaryConditions.cpp:1022-1068).
template <int T>
__global__ __launch_bounds__(kThreads) void damping_kernel(Box3 b, V r, CV a, DampCoef dc, int f, int with_rayleigh,
                                                          int above_ground, NodeZ z, profile::ProfileView prof,
                                                          int klo, double dx0, double dx1)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double h = heights::height_above_ground<T>(z, i, j, k, klo);
        const double zc = point_z<T>(z, i, j, k);
        const double x = dc.plo[0] + (i + (sc::nodal(T, 0) ? 0.0 : 0.5)) * dx0;
        const double y = dc.plo[1] + (j + (sc::nodal(T, 1) ? 0.0 : 0.5)) * dx1;
        const double zr = above_ground ? h : zc;
        const double cr = with_rayleigh ? dc.rayleigh_part(zr) : 0.0;
        const double cs = dc.sponge_part(x, y);
        auto ref_at = [&](double zz) {
            double u, v, th, kk, nut;
            profile::profile_eval(prof, zz, u, v, th, kk, nut);
            return f == F_U ? u : f == F_V ? v : f == F_THETA ? th : f == F_TKE ? kk : 0.0;
        };
        if (cr > 0) {
            const double ref = (f == F_W) ? 0.0 : ref_at(f == F_THETA ? zr : h);
            r(i, j, k) += cr * (a(i, j, k) - ref);
        }
        if (cs > 0) {
            const double ref = (f == F_W) ? 0.0 : ref_at(h);
            r(i, j, k) += cs * (a(i, j, k) - ref);
        }
    }
}

// The hi face of a periodic nodal axis takes the lo face's value.
__global__ __launch_bounds__(kThreads) void seam_copy_kernel(Box3 plane, V a, int axis, int n)
{
    GUST_GRID_LOOP(plane)
    {
        int i, j, k;
        plane.at(t, i, j, k);
        a(i + (axis == 0 ? n : 0), j + (axis == 1 ? n : 0), k) = a(i, j, k);
    }
}

// y -= x (sub) or y += x: MultiFab::Subtract / MultiFab::Add.
__global__ __launch_bounds__(kThreads) void accumulate_kernel(Box3 b, V y, CV x, int sub)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        if (sub) { y(i, j, k) -= x(i, j, k); }
        else { y(i, j, k) += x(i, j, k); }
    }
}

// The same for up to three fields in one launch (blockIdx.y: the field).
struct AccSet {
    Box3 b[3];
    V y[3];
    CV x[3];
    int sub;
};

__global__ __launch_bounds__(kThreads) void accumulate3_kernel(AccSet s)
{
    const int q = blockIdx.y;
    const Box3& b = s.b[q];
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        if (s.sub) { s.y[q](i, j, k) -= s.x[q](i, j, k); }
        else { s.y[q](i, j, k) += s.x[q](i, j, k); }
    }
}

// ---- implicit rows: wall model and z-face folds -------------------------------

struct AV {
    double* p;
    int64_t js, ks, ns;
    int b0, b1, b2;
    __device__ __forceinline__ double& operator()(int i, int j, int k, int n) const
    {
        return p[int64_t(i - b0) + int64_t(j - b1) * js + int64_t(k - b2) * ks + int64_t(n) * ns];
    }
};

AV a_view(const gust_array4* a)
{
    return AV{static_cast<double*>(a->p), a->jstride, a->kstride, a->nstride, a->begin[0], a->begin[1], a->begin[2]};
}

// The layers' coefficient on the implicit diagonal (BoundaryConditions.cpp:1070-1099).
template <int T>
__global__ __launch_bounds__(kThreads) void damping_diag_kernel(Box3 b, AV A, DampCoef dc, int with_rayleigh,
                                                               int above_ground, NodeZ z, int klo, double dx0,
                                                               double dx1)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double h = heights::height_above_ground<T>(z, i, j, k, klo);
        const double zc = point_z<T>(z, i, j, k);
        const double x = dc.plo[0] + (i + (sc::nodal(T, 0) ? 0.0 : 0.5)) * dx0;
        const double y = dc.plo[1] + (j + (sc::nodal(T, 1) ? 0.0 : 0.5)) * dx1;
        A(i, j, k, GUST_COEF_P) += (with_rayleigh ? dc.rayleigh_part(above_ground ? h : zc) : 0.0) +
                                   dc.sponge_part(x, y);
    }
}

__global__ __launch_bounds__(kThreads) void add_dsdk_kernel(Box3 b, AV A, CV ds)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        A(i, j, k, GUST_COEF_P) += ds(i, j, k);
    }
}

struct Fold {
    int dirichlet_pt, to_other;
    double fold;
};

// The wall model's rows (SurfaceLayer.cpp:494-567) on the ground plane.
template <int T>
__global__ __launch_bounds__(kThreads) void wall_fold_kernel(Box3 cols, AV A, SSV ss, CV J, int k0, int eq,
                                                            int wall_k_dirichlet, double dzinv)
{
    GUST_GRID_LOOP(cols)
    {
        int i, j, kk;
        cols.at(t, i, j, kk);
        A(i, j, k0, GUST_COEF_P) -= A(i, j, k0, GUST_COEF_ZM);
        A(i, j, k0, GUST_COEF_ZM) = 0;
        if (eq == 0) {
            A(i, j, k0, GUST_COEF_P) += 0.5 * (ss(GUST_SS_DRAG, i - 1, j) + ss(GUST_SS_DRAG, i, j)) * dzinv /
                                        sc::interp<sc::CC, T>(J, i, j, k0);
        } else if (eq == 1) {
            A(i, j, k0, GUST_COEF_P) += 0.5 * (ss(GUST_SS_DRAG, i, j - 1) + ss(GUST_SS_DRAG, i, j)) * dzinv /
                                        sc::interp<sc::CC, T>(J, i, j, k0);
        } else if (eq == 4) {
            const double a = sc::dmax(ss(GUST_SS_USTAR, i, j), kWallKUstarMin) / ss(GUST_SS_D0, i, j);
            if (wall_k_dirichlet) {
                for (int n = 0; n < GUST_N_COEFS; ++n) { A(i, j, k0, n) = 0; }
                A(i, j, k0, GUST_COEF_P) = a;
            } else {
                A(i, j, k0, GUST_COEF_P) += a;
            }
        }
    }
}

// apply_to_implicit along one direction (BoundaryConditions.cpp:885-940):
// the identity row at a Dirichlet face point, else the first row's ghost
// coupling folded into the diagonal or onto the opposite neighbour; one
// thread per point of the face plane (its index along d unused).
__global__ __launch_bounds__(kThreads) void bc_fold_kernel(Box3 plane, AV A, int d, int edge_lo, int edge_hi,
                                                          Fold flo, Fold fhi)
{
    GUST_GRID_LOOP(plane)
    {
        int i, j, k;
        plane.at(t, i, j, k);
        const int cm = d == 0 ? GUST_COEF_XM : d == 1 ? GUST_COEF_YM : GUST_COEF_ZM;
        const int cp = d == 0 ? GUST_COEF_XP : d == 1 ? GUST_COEF_YP : GUST_COEF_ZP;
        for (int side = 0; side < 2; ++side) {
            const Fold& f = side == 0 ? flo : fhi;
            const int edge = side == 0 ? edge_lo : edge_hi;
            const int pi = d == 0 ? edge : i, pj = d == 1 ? edge : j, pk = d == 2 ? edge : k;
            const int ghost = side == 0 ? cm : cp, other = side == 0 ? cp : cm;
            if (f.dirichlet_pt) {
                for (int n = 0; n < GUST_N_COEFS; ++n) { A(pi, pj, pk, n) = 0; }
                A(pi, pj, pk, GUST_COEF_P) = 1;
            } else {
                const double ag = A(pi, pj, pk, ghost);
                A(pi, pj, pk, ghost) = 0;
                if (f.to_other) { A(pi, pj, pk, other) += ag; }
                else { A(pi, pj, pk, GUST_COEF_P) += f.fold * ag; }
            }
        }
    }
}

// ---- host helpers ---------------------------------------------------------------

void check_rc(int rc)
{
    if (rc != 0) { throw Fail{rc}; }
}

// A compact device field over box b (from the workspace), ncomp components.
gust_array4 scratch_view(hipStream_t st, int slot, const IBox& b, int ncomp = 1)
{
    const int64_t n = b.npts();
    gust_array4 v;
    v.p = Workspace::get(st, slot, size_t(n) * ncomp * sizeof(double));
    v.jstride = b.len(0);
    v.kstride = int64_t(b.len(0)) * b.len(1);
    v.nstride = n;
    for (int d = 0; d < 3; ++d) {
        v.begin[d] = b.lo[d];
        v.end[d] = b.hi[d] + 1;
    }
    v.ncomp = ncomp;
    return v;
}

IBox slab_box(const gust_geometry& g)
{
    IBox b = grow(loc_box(g, GUST_LOC_CELL), 1, 1, 0);
    b.hi[2] = b.lo[2];
    return b;
}

// y_c -= x_c (sign < 0) or += for the three velocity locations, one launch.
void cgrid_sub3(hipStream_t st, const gust_geometry& g, gust_array4* const* y, const gust_array4* x, double sign)
{
    AccSet a{};
    int64_t most = 1;
    for (int c = 0; c < 3; ++c) {
        a.b[c] = box3(loc_box(g, 1 << c));
        a.y[c] = dview<double>(y[c]);
        a.x[c] = dview<const double>(&x[c]);
        most = std::max<int64_t>(most, a.b[c].npts());
    }
    a.sub = sign < 0 ? 1 : 0;
    accumulate3_kernel<<<dim3(blocks_for(most), 3), kThreads, 0, st>>>(a);
    GUST_HIP(hipGetLastError());
}

void cgrid_sub(hipStream_t st, const Box3& b, gust_array4* y, const gust_array4* x, double sign)
{
    accumulate_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, dview<double>(y), dview<const double>(x),
                                                                 sign < 0 ? 1 : 0);
    GUST_HIP(hipGetLastError());
}

KConst closure_consts(const gust_closure_options& o)
{
    KConst c;
    c.kappa = o.kappa;
    c.Cmu0 = o.Cmu0;
    c.Cmu0_pow3 = o.Cmu0 * o.Cmu0 * o.Cmu0;
    c.inv_Cb_sq = 1.0 / (o.Cb * o.Cb);
    c.Rt_crit = o.Rt_crit;
    c.Rt_min = o.Rt_min;
    c.l_g_max = o.max_geom_lscale;
    c.tke_floor = o.tke_floor > 0 ? o.tke_floor : DBL_EPSILON;
    c.n2_eps = DBL_EPSILON;
    c.nu_mol = o.nu_molecular;
    c.alpha_mol = o.alpha_molecular;
    c.inv_sigma = 1.0 / o.sigma_k;
    return c;
}

void check_closure_options(const char* op, const gust_closure_options* o)
{
    if (!o) { GUST_FAIL("gust: %s: null closure options", op); }
    auto pos = [&](double x, const char* key) {
        if (std::isnan(x) || !(x > 0 && std::isfinite(x))) { GUST_FAIL("gust: %s: turb.%s must be finite and > 0", op, key); }
    };
    pos(o->Cmu0, "Cmu0");
    pos(o->Cb, "Cb");
    pos(o->max_geom_lscale, "max_geom_lscale");
    pos(o->sigma_k, "sigma_k");
    pos(o->kappa, "kappa");
    if (std::isnan(o->Rt_crit) || !(o->Rt_crit <= 0)) { GUST_FAIL("gust: %s: turb.Rt_crit must be <= 0", op); }
    if (std::isnan(o->Rt_min) || !(o->Rt_min < o->Rt_crit && o->Rt_min > -3.6)) {
        GUST_FAIL("gust: %s: turb.Rt_min must be < turb.Rt_crit and > -3.6", op);
    }
    if (std::isnan(o->tke_floor) || !(o->tke_floor == 0 || (o->tke_floor >= DBL_EPSILON && std::isfinite(o->tke_floor)))) {
        GUST_FAIL("gust: %s: turb.tke_floor must be 0 or >= machine epsilon", op);
    }
    if (std::isnan(o->nu_molecular) || !(o->nu_molecular >= 0) || std::isnan(o->alpha_molecular) ||
        !(o->alpha_molecular >= 0)) {
        GUST_FAIL("gust: %s: turb.nu_molecular and turb.alpha_molecular must be >= 0", op);
    }
}

void check_surface_options(const char* op, const gust_surface_options* o)
{
    if (!o) { GUST_FAIL("gust: %s: null surface-layer options", op); }
    for (const double x : {o->z0, o->Cmu0, o->kappa}) {
        if (std::isnan(x) || !(x > 0 && std::isfinite(x))) {
            GUST_FAIL("gust: %s: surface layer z0, Cmu0 and kappa must be finite and > 0", op);
        }
    }
}

void check_state(const char* op, const gust_geometry& g, const gust_state* q)
{
    if (!q) { GUST_FAIL("gust: %s: null state", op); }
    const char* names[6] = {"u", "v", "w", "p", "theta", "tke"};
    const int locs[6] = {GUST_LOC_XFACE, GUST_LOC_YFACE, GUST_LOC_ZFACE, GUST_LOC_CELL, GUST_LOC_CELL, GUST_LOC_CELL};
    for (int n = 0; n < 6; ++n) { check_field(op, names[n], q->f[n], g, locs[n], 2); }
}

void check_mesh(const char* op, const gust_geometry& g, const gust_mesh* m)
{
    if (!m) { GUST_FAIL("gust: %s: null mesh", op); }
    check_nodal_heights(op, m->z_nd, g);
    check_field(op, "detJ", m->detJ, g, GUST_LOC_CELL, 1);
    check_field(op, "ax", m->ax, g, GUST_LOC_XFACE, 1);
    check_field(op, "ay", m->ay, g, GUST_LOC_YFACE, 1);
    check_field(op, "az", m->az, g, GUST_LOC_ZFACE, 1);
}

void check_physics(const char* op, const gust_geometry& g, const gust_physics* p)
{
    if (!p) { GUST_FAIL("gust: %s: null physics", op); }
    if (p->use_closure) {
        check_closure_options(op, &p->closure);
    } else {
        check_field(op, "nu", p->nu, g, GUST_LOC_CELL, 1);
        check_field(op, "gamma_theta", p->gamma_theta, g, GUST_LOC_CELL, 1);
        check_field(op, "gamma_tke", p->gamma_tke, g, GUST_LOC_CELL, 1);
    }
    if (p->use_bc) {
        const char* keys[6] = {"xlo", "ylo", "zlo", "xhi", "yhi", "zhi"};
        for (const int o : {0, 1, 3, 4}) {
            const int t = p->face[o];
            const bool per = g.periodic[o % 3] != 0;
            if (per != (t == GUST_FACE_PERIODIC)) {
                GUST_FAIL("gust: %s: bc.%s must be periodic exactly where the geometry is", op, keys[o]);
            }
            if (!per && !(t == GUST_FACE_INFLOW || t == GUST_FACE_OUTFLOW || t == GUST_FACE_SLIP_WALL ||
                          t == GUST_FACE_NO_SLIP_WALL)) {
                GUST_FAIL("gust: %s: bc.%s must be inflow, outflow, slip_wall or no_slip_wall", op, keys[o]);
            }
            if (t == GUST_FACE_INFLOW) {
                const std::vector<std::string> err = profile::problems(p->inflow);
                if (!err.empty()) { GUST_FAIL("gust: %s: bc.%s = inflow: %s", op, keys[o], err.front().c_str()); }
            }
        }
        {
            const gust_damping& dm = p->damping;
            if (std::isnan(dm.rayleigh_zdamp) || !(dm.rayleigh_zdamp >= 0 && dm.rayleigh_zdamp < dm.ztop) ||
                std::isnan(dm.rayleigh_dampcoef) || !(dm.rayleigh_dampcoef >= 0 && std::isfinite(dm.rayleigh_dampcoef))) {
                if (dm.rayleigh_zdamp != 0 || dm.rayleigh_dampcoef != 0) {
                    GUST_FAIL("gust: %s: bc.rayleigh_zdamp must be in [0, the top level) and bc.rayleigh_dampcoef "
                              "finite and >= 0 (1/s)", op);
                }
            }
            if (dm.rayleigh_zdamp > 0 && !(dm.rayleigh_dampcoef > 0)) {
                GUST_FAIL("gust: %s: bc.rayleigh_zdamp > 0 needs bc.rayleigh_dampcoef > 0", op);
            }
            if (std::isnan(dm.sponge_width) || !(dm.sponge_width >= 0) || std::isnan(dm.sponge_strength) ||
                !(dm.sponge_strength >= 0)) {
                GUST_FAIL("gust: %s: bc.sponge_width and bc.sponge_strength must be >= 0", op);
            }
            if (dm.rayleigh_zdamp > 0 || dm.sponge_width > 0) {
                const std::vector<std::string> err = profile::problems(p->inflow);
                if (!err.empty()) { GUST_FAIL("gust: %s: the damping layers' profile: %s", op, err.front().c_str()); }
            }
        }
        const int zlo = p->face[2], zhi = p->face[5];
        if (!(zlo == GUST_FACE_SURFACE_LAYER || zlo == GUST_FACE_SLIP_WALL || zlo == GUST_FACE_NO_SLIP_WALL)) {
            GUST_FAIL("gust: %s: bc.zlo must be surface_layer, slip_wall or no_slip_wall", op);
        }
        if (!(zhi == GUST_FACE_SLIP_WALL || zhi == GUST_FACE_NO_SLIP_WALL || zhi == GUST_FACE_GEOSTROPHIC)) {
            GUST_FAIL("gust: %s: bc.zhi must be slip_wall, no_slip_wall or geostrophic", op);
        }
        if (zhi == GUST_FACE_GEOSTROPHIC && !(p->forcing.coriolis && p->forcing.geostrophic)) {
            GUST_FAIL("gust: %s: bc.zhi = geostrophic needs Coriolis with geostrophic forcing", op);
        }
        if (zlo == GUST_FACE_SURFACE_LAYER) { check_surface_options(op, &p->surface); }
        if (!std::isfinite(p->zhi_theta_grad)) { GUST_FAIL("gust: %s: bc.zhi_theta_grad must be finite", op); }
    }
    if (p->n_theta_ref < 1 || p->n_theta_ref > 64) {
        GUST_FAIL("gust: %s: the reference profile needs 1 to 64 points", op);
    }
    for (int m = 0; m < p->n_theta_ref; ++m) {
        if (!std::isfinite(p->theta_ref_z[m]) || !(p->theta_ref[m] > 0 && std::isfinite(p->theta_ref[m]))) {
            GUST_FAIL("gust: %s: reference profile point %d: z must be finite and theta finite and > 0", op, m);
        }
        if (m > 0 && !(p->theta_ref_z[m] > p->theta_ref_z[m - 1])) {
            GUST_FAIL("gust: %s: reference profile: heights must increase strictly (point %d)", op, m);
        }
    }
    const gust_forcing& f = p->forcing;
    if (std::isnan(f.gravity) || !(f.gravity >= 0 && std::isfinite(f.gravity)) || std::isnan(f.theta0) ||
        !(f.theta0 > 0 && std::isfinite(f.theta0))) {
        GUST_FAIL("gust: %s: buoyancy gravity must be finite and >= 0 and theta0 finite and > 0", op);
    }
    if (f.coriolis && (std::isnan(f.latitude) || !(f.latitude >= -90 && f.latitude <= 90) ||
                       !(f.rotation_period > 0 && std::isfinite(f.rotation_period)))) {
        GUST_FAIL("gust: %s: coriolis latitude must be in [-90, 90] and rotation_period finite and > 0", op);
    }
    for (int d = 0; d < 3; ++d) {
        if (!std::isfinite(f.pressure_gradient[d])) { GUST_FAIL("gust: %s: the force must be finite", op); }
    }
    if (p->momentum_limiter < 0 || p->momentum_limiter > 2 || p->theta_limiter < 0 || p->theta_limiter > 2) {
        GUST_FAIL("gust: %s: unknown limiter", op);
    }
}

bool wall_on(const gust_physics& p) { return p.use_bc && p.face[2] == GUST_FACE_SURFACE_LAYER; }

SurfOpt surf_opt(const gust_surface_options& o)
{
    return SurfOpt{o.z0, o.kappa, 1.0 / (o.Cmu0 * o.Cmu0), o.denominator_floor == 1 ? 1 : 0};
}

void launch_heights(hipStream_t st, const gust_geometry& g, const gust_array4* z, gust_array4* dist)
{
    const IBox cells = grow(loc_box(g, GUST_LOC_CELL), 1, 1, 0);
    const Box3 b = box3(cells);
    height_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, dview<const double>(z), dview<double>(dist), g.lo[2]);
    IBox plane = cells;
    plane.hi[2] = plane.lo[2];
    const Box3 pb = box3(plane);
    copy_vertical_kernel<<<blocks_for(pb.npts()), kThreads, 0, st>>>(pb, dview<double>(dist), g.lo[2], g.hi[2]);
    GUST_HIP(hipGetLastError());
}

void launch_closure(hipStream_t st, const gust_geometry& g, const gust_state& q, const gust_array4* z,
                    const gust_array4* dist, double z0, const gust_closure_options& o, double gravity, double theta0,
                    const gust_closure_fields& out)
{
    const IBox cells = grow(loc_box(g, GUST_LOC_CELL), 1, 1, 0);
    const Box3 b = box3(cells);
    ClosureOut co{dview<double>(out.nu),  dview<double>(out.gamma_theta), dview<double>(out.gamma_tke),
                  dview<double>(out.Ps),  dview<double>(out.Pb),          dview<double>(out.eps),
                  dview<double>(out.dSdk)};
    closure_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(
        b, dview<const double>(q.f[0]), dview<const double>(q.f[1]), dview<const double>(q.f[2]),
        dview<const double>(q.f[4]), dview<const double>(q.f[5]), dview<const double>(z), dview<const double>(dist), z0,
        closure_consts(o), gravity / theta0, 1.0 / g.dx[0], 1.0 / g.dx[1], 1.0 / g.dx[2], co, g.lo[0], g.hi[0], g.lo[1],
        g.hi[1]);
    IBox plane = cells;
    plane.hi[2] = plane.lo[2];
    const Box3 pb = box3(plane);
    for (gust_array4* a : {out.nu, out.gamma_theta, out.gamma_tke}) {
        copy_vertical_kernel<<<blocks_for(pb.npts()), kThreads, 0, st>>>(pb, dview<double>(a), g.lo[2], g.hi[2]);
    }
    GUST_HIP(hipGetLastError());
}

void launch_surface(hipStream_t st, const gust_geometry& g, const gust_array4* u, const gust_array4* v,
                    const gust_array4* z, const gust_surface_options& o, gust_array4* ss)
{
    const Box3 cols = box3(slab_box(g));
    // The count of bad columns is read with the caller's other checks (the
    // low word of a slot that starts at zero).
    DeferredUse dq(st);
    const double z0 = o.z0;
    auto* bad = reinterpret_cast<unsigned*>(dq->zero_slot([z0](unsigned long long n) {
        if (n) {
            GUST_FAIL("gust: compute_surface_state: no positive, finite friction velocity at %llu columns (surface "
                      "layer z0 = %g)",
                      n, z0);
        }
    }));
    surface_kernel<<<blocks_for(cols.npts()), kThreads, 0, st>>>(cols, dview<const double>(u), dview<const double>(v),
                                                                 dview<const double>(z), surf_opt(o), 1.0 / g.dx[0],
                                                                 1.0 / g.dx[1], ss_view(ss), g.lo[2], bad);
    GUST_HIP(hipGetLastError());
    dq.finish();
}

void check_ss(const char* op, const gust_geometry& g, const gust_array4* ss)
{
    if (!covers(ss, slab_box(g), GUST_SS_N)) {
        GUST_FAIL("gust: %s: the surface state must cover the ground columns and one lateral ghost column with %d "
                  "components",
                  op, int(GUST_SS_N));
    }
}

bool damping_on(const gust_physics& p) { return p.damping.rayleigh_zdamp > 0 || p.damping.sponge_width > 0; }

DampCoef damp_coef(const gust_physics& p)
{
    const gust_damping& d = p.damping;
    DampCoef c{};
    c.zd = d.rayleigh_zdamp;
    c.rc = d.rayleigh_dampcoef;
    c.ztop = d.ztop;
    c.sw = d.sponge_width;
    c.ss = d.sponge_strength;
    c.rayleigh = c.zd > 0 ? 1 : 0;
    c.sponge = c.sw > 0 ? 1 : 0;
    c.inv_w = c.sponge ? 1 / c.sw : 0.0;
    for (int k = 0; k < 2; ++k) {
        c.plo[k] = d.prob_lo[k];
        c.phi[k] = d.prob_hi[k];
    }
    for (int dd = 0; dd < 2; ++dd) {
        for (int side = 0; side < 2; ++side) {
            const int t = p.face[side == 0 ? dd : dd + 3];
            c.lateral[2 * dd + side] = (t == GUST_FACE_INFLOW || t == GUST_FACE_OUTFLOW) ? 1 : 0;
        }
    }
    return c;
}

template <class F>
void dispatch_loc(int loc, F&& f)
{
    if (loc == 1) { f(std::integral_constant<int, 1>{}); }
    else if (loc == 2) { f(std::integral_constant<int, 2>{}); }
    else if (loc == 4) { f(std::integral_constant<int, 4>{}); }
    else { f(std::integral_constant<int, 0>{}); }
}

IBox view_box(const gust_array4* a)
{
    IBox b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = a->begin[d];
        b.hi[d] = a->end[d] - 1;
    }
    return b;
}

GgMap gg_map(const gust_geometry& g)
{
    return GgMap{g.lo[0], g.hi[0], g.lo[1], g.hi[1], g.periodic[0] ? 1 : 0, g.periodic[1] ? 1 : 0,
                 g.hi[0] - g.lo[0] + 1, g.hi[1] - g.lo[1] + 1};
}

// The fills along z over every lateral column of the field's fab (the ghost
// slab of the z faces spans the lateral ghosts, BoundaryConditions.cpp:83-104).
template <int T>
// homogeneous: the fill of a perturbation (BoundaryConditions.cpp:427, 510):
// every prescribed value and gradient 0, the wall law's gradient too.
FzField fz_field(const gust_geometry& g, gust_array4* a, int f, const gust_physics& p, bool homogeneous = false)
{
    FzField F{};
    for (int side = 0; side < 2; ++side) {
        const bool lo = side == 0;
        const int t = lo ? p.face[2] : p.face[5];
        FaceFill& x = lo ? F.lo : F.hi;
        x.rule = rule_for(f, t, 2, lo ? 1 : 0);
        x.V = 0;
        if (t == GUST_FACE_GEOSTROPHIC) { x.V = f == F_U ? p.forcing.ug : f == F_V ? p.forcing.vg : 0.0; }
        if (t == GUST_FACE_NO_SLIP_WALL && f <= F_W) { x.V = p.wall_velocity[lo ? 2 : 5][f]; }
        x.G = p.zhi_theta_grad;
        x.wall = (t == GUST_FACE_SURFACE_LAYER) ? 1 : 0;
        x.comp = f == F_U ? 0 : f == F_V ? 1 : 2;
        if (homogeneous) {
            x.V = 0;
            x.G = 0;
            x.wall = 0;
        }
    }
    IBox cols = view_box(a);
    cols.lo[2] = cols.hi[2] = g.lo[2];
    F.a = dview<double>(a);
    F.cols = box3(cols);
    F.loc = T;
    F.edge_lo = g.lo[2];
    F.edge_hi = sc::nodal(T, 2) ? g.hi[2] + 1 : g.hi[2];
    return F;
}

void launch_fill_z(hipStream_t st, const gust_geometry& g, const FzSet& set, int n, const gust_physics& p,
                   const gust_array4* z, const gust_array4* ss)
{
    int64_t most = 1;
    for (int q = 0; q < n; ++q) { most = std::max<int64_t>(most, set.f[q].cols.npts()); }
    SSV sv = ss ? ss_view(ss) : SSV{nullptr, 0, 0, 0, 0};
    fill_z_kernel<<<dim3(blocks_for(most), n), kThreads, 0, st>>>(set, node_z(z), sv, surf_opt(p.surface), gg_map(g));
    GUST_HIP(hipGetLastError());
}

// The fills along x (d = 0) or y (d = 1) over the ghost slab's plane: the
// whole fab along x for d = 1, the domain (or the fab where periodic) along
// the later directions.
template <int T>
SideField side_field(const gust_geometry& g, gust_array4* a, int f, int d, const gust_physics& p,
                     bool homogeneous = false)
{
    SideField F{};
    for (int side = 0; side < 2; ++side) {
        const bool lo = side == 0;
        const int o = lo ? d : d + 3;
        const int t = p.face[o];
        SideFill& x = lo ? F.lo : F.hi;
        x.rule = rule_for(f, t, d, lo ? 1 : 0);
        x.V = 0;
        if (t == GUST_FACE_NO_SLIP_WALL && f <= F_W) { x.V = p.wall_velocity[o][f]; }
        if (homogeneous) {
            x.V = 0;
            if (x.rule == R_PROFILE) { x.rule = R_ZERO; }
        }
    }
    const IBox fab = view_box(a);
    IBox plane;
    for (int e = 0; e < 3; ++e) {
        if (e == d) {
            plane.lo[e] = plane.hi[e] = g.lo[e];
        } else if (e < d || (e < 2 && g.periodic[e])) {
            plane.lo[e] = fab.lo[e];
            plane.hi[e] = fab.hi[e];
        } else {
            plane.lo[e] = g.lo[e];
            plane.hi[e] = g.hi[e] + (sc::nodal(T, e) ? 1 : 0);
        }
    }
    F.a = dview<double>(a);
    F.plane = box3(plane);
    F.loc = T;
    F.f = f;
    F.edge_lo = g.lo[d];
    F.edge_hi = sc::nodal(T, d) ? g.hi[d] + 1 : g.hi[d];
    return F;
}

void launch_fill_sides(hipStream_t st, const gust_geometry& g, const gust_state& q, int d, const gust_physics& p,
                       const gust_array4* z, const profile::ProfileView& prof, bool homogeneous = false)
{
    SideSet set{};
    set.f[0] = side_field<sc::XF>(g, q.f[0], F_U, d, p, homogeneous);
    set.f[1] = side_field<sc::YF>(g, q.f[1], F_V, d, p, homogeneous);
    set.f[2] = side_field<sc::ZF>(g, q.f[2], F_W, d, p, homogeneous);
    for (int n = 3; n < 6; ++n) { set.f[n] = side_field<sc::CC>(g, q.f[n], n, d, p, homogeneous); }
    int64_t most = 1;
    for (int n = 0; n < 6; ++n) { most = std::max<int64_t>(most, set.f[n].plane.npts()); }
    fill_side_kernel<<<dim3(blocks_for(most), 6), kThreads, 0, st>>>(set, d, node_z(z), prof, g.lo[2]);
    GUST_HIP(hipGetLastError());
}

// FillBoundary of n fields (locations locs) in one launch.
void launch_fill_boundary(hipStream_t st, const gust_geometry& g, gust_array4* const* a, const int* locs, int n)
{
    if (!g.periodic[0] && !g.periodic[1]) { return; }
    FbSet set{};
    int64_t most = 1;
    for (int q = 0; q < n; ++q) {
        const IBox vb = loc_box(g, locs[q]);
        FbField& F = set.f[q];
        F.a = dview<double>(a[q]);
        F.fab = box3(view_box(a[q]));
        F.vlo0 = vb.lo[0]; F.vhi0 = vb.hi[0];
        F.vlo1 = vb.lo[1]; F.vhi1 = vb.hi[1];
        F.vlo2 = vb.lo[2]; F.vhi2 = vb.hi[2];
        most = std::max<int64_t>(most, F.fab.npts());
    }
    fill_boundary_kernel<<<dim3(blocks_for(most), n), kThreads, 0, st>>>(
        set, g.periodic[0] ? 1 : 0, g.periodic[1] ? 1 : 0, g.hi[0] - g.lo[0] + 1, g.hi[1] - g.lo[1] + 1);
    GUST_HIP(hipGetLastError());
}

// fill_no_flux_w's columns: those whose u, v points and nodes exist
// (BoundaryConditions.cpp:766-778).
IBox no_flux_columns(const gust_array4* w, const gust_array4* u, const gust_array4* v, const gust_array4* z, int klo)
{
    IBox c;
    c.lo[0] = std::max({w->begin[0], u->begin[0], v->begin[0], z->begin[0]});
    c.hi[0] = std::min({w->end[0] - 1, u->end[0] - 2, v->end[0] - 1, z->end[0] - 2});
    c.lo[1] = std::max({w->begin[1], u->begin[1], v->begin[1], z->begin[1]});
    c.hi[1] = std::min({w->end[1] - 1, u->end[1] - 1, v->end[1] - 2, z->end[1] - 2});
    c.lo[2] = c.hi[2] = klo;
    return c;
}

// fill_ghosts on validated arguments (BoundaryConditions.cpp:597-666); ss
// receives the surface state. homogeneous: the fill of a perturbation
// (every prescribed value, profile and gradient 0, no surface state; the
// ground w's no-flux value is linear and kept).
void fill_ghosts_impl(hipStream_t st, const gust_geometry& g, const gust_state& q, const gust_array4* z,
                      const gust_physics& p, gust_array4* ss, bool homogeneous = false)
{
    const int locs[6] = {GUST_LOC_XFACE, GUST_LOC_YFACE, GUST_LOC_ZFACE, GUST_LOC_CELL, GUST_LOC_CELL, GUST_LOC_CELL};
    launch_fill_boundary(st, g, q.f, locs, 6);
    const bool wall = wall_on(p) && !homogeneous;
    bool inflow = false;
    for (int o = 0; o < 6; ++o) { inflow = inflow || p.face[o] == GUST_FACE_INFLOW; }
    inflow = inflow && !homogeneous;
    const profile::ProfileView prof = inflow ? profile::device_view(st, p.inflow) : profile::ProfileView{};
    for (int d = 0; d < 2; ++d) {
        if (!g.periodic[d]) { launch_fill_sides(st, g, q, d, p, z, prof, homogeneous); }
    }
    if (wall) { launch_surface(st, g, q.f[0], q.f[1], z, p.surface, ss); }
    const gust_array4* sv = wall ? ss : nullptr;
    // The z fills of all six fields: each reads only its own field (and the
    // surface state), so one launch does the reference's per-field passes.
    FzSet fz{};
    fz.f[0] = fz_field<sc::XF>(g, q.f[0], F_U, p, homogeneous);
    fz.f[1] = fz_field<sc::YF>(g, q.f[1], F_V, p, homogeneous);
    fz.f[2] = fz_field<sc::ZF>(g, q.f[2], F_W, p, homogeneous);
    for (int n = 3; n < 6; ++n) { fz.f[n] = fz_field<sc::CC>(g, q.f[n], n, p, homogeneous); }
    launch_fill_z(st, g, fz, 6, p, z, sv);
    // w on the ground and top from u and v (their z ghosts already filled).
    // The reference's FillBoundary of w after it is subsumed by the final one:
    // nothing in between changes w.
    const Box3 b = box3(no_flux_columns(q.f[2], q.f[0], q.f[1], z, g.lo[2]));
    no_flux_w_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(
        b, dview<double>(q.f[2]), dview<const double>(q.f[0]), dview<const double>(q.f[1]), dview<const double>(z),
        g.lo[2], g.hi[2] + 1, dirichlet_normal(p.face[2]) ? 1 : 0, dirichlet_normal(p.face[5]) ? 1 : 0,
        1.0 / g.dx[0], 1.0 / g.dx[1]);
    GUST_HIP(hipGetLastError());
    launch_fill_boundary(st, g, q.f, locs, 6);
}

Rot rotation(const gust_forcing& f)
{
    Rot r{};
    r.on = f.coriolis ? 1 : 0;
    r.three_d = f.coriolis_3d ? 1 : 0;
    const double pi = 3.14159265358979323846;
    const double phi = f.latitude * pi / 180;
    const double fc = 4 * pi / f.rotation_period;
    r.fs = f.coriolis ? fc * std::sin(phi) : 0.0;
    r.fcos = (f.coriolis && f.coriolis_3d) ? fc * std::cos(phi) : 0.0;
    r.ug = f.geostrophic ? f.ug : 0.0;
    r.vg = f.geostrophic ? f.vg : 0.0;
    for (int d = 0; d < 3; ++d) { r.pg[d] = f.pressure_gradient[d]; }
    return r;
}

IBox cv_box(const gust_geometry& g, int loc, int d)
{
    IBox b = loc_box(g, loc);
    if ((loc >> d) & 1) {
        b.lo[d] = g.lo[d] - 1;
        b.hi[d] = g.hi[d] + 1;
    } else {
        b.hi[d] = g.hi[d] + 1;
    }
    return b;
}

}   // namespace

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_height_above_ground(const gust_geometry* geom, const gust_array4* z_nd, gust_array4* dist, void* stream)
{
    return guarded([&] {
        const char* op = "height_above_ground";
        check_geom(op, geom);
        check_nodal_heights(op, z_nd, *geom);
        check_field(op, "dist", dist, *geom, GUST_LOC_CELL, 1);
        launch_heights(static_cast<hipStream_t>(stream), *geom, z_nd, dist);
    });
}

int gust_closure_evaluate(const gust_geometry* geom, const gust_state* q, const gust_array4* z_nd,
                          const gust_array4* dist, double z0, const gust_closure_options* opt, double gravity,
                          double theta0, gust_closure_fields* out, void* stream)
{
    return guarded([&] {
        const char* op = "Closure::evaluate";
        check_geom(op, geom);
        check_state(op, *geom, q);
        check_nodal_heights(op, z_nd, *geom);
        check_field(op, "dist", dist, *geom, GUST_LOC_CELL, 1);
        check_closure_options(op, opt);
        if (std::isnan(z0) || !(z0 >= 0 && std::isfinite(z0))) { GUST_FAIL("gust: %s: z0 must be finite and >= 0", op); }
        if (!out) { GUST_FAIL("gust: %s: null fields", op); }
        check_field(op, "nu", out->nu, *geom, GUST_LOC_CELL, 1);
        check_field(op, "gamma_theta", out->gamma_theta, *geom, GUST_LOC_CELL, 1);
        check_field(op, "gamma_tke", out->gamma_tke, *geom, GUST_LOC_CELL, 1);
        for (gust_array4* a : {out->Ps, out->Pb, out->eps, out->dSdk}) { check_field(op, "a source", a, *geom, GUST_LOC_CELL, 0); }
        launch_closure(static_cast<hipStream_t>(stream), *geom, *q, z_nd, dist, z0, *opt, gravity, theta0, *out);
    });
}

int gust_surface_state(const gust_geometry* geom, const gust_array4* u, const gust_array4* v,
                       const gust_array4* z_nd, const gust_surface_options* opt, gust_array4* ss, void* stream)
{
    return guarded([&] {
        const char* op = "compute_surface_state";
        check_geom(op, geom);
        check_field(op, "u", u, *geom, GUST_LOC_XFACE, 1);
        check_field(op, "v", v, *geom, GUST_LOC_YFACE, 1);
        check_nodal_heights(op, z_nd, *geom);
        check_surface_options(op, opt);
        check_ss(op, *geom, ss);
        if (geom->hi[2] - geom->lo[2] + 1 < 2) { GUST_FAIL("gust: %s: the mesh needs at least two layers", op); }
        launch_surface(static_cast<hipStream_t>(stream), *geom, u, v, z_nd, *opt, ss);
    });
}

int gust_fill_ghosts(const gust_geometry* geom, gust_state* q, const gust_mesh* mesh, const gust_physics* phys,
                     gust_array4* ss, void* stream)
{
    return guarded([&] {
        const char* op = "fill_ghosts";
        check_geom(op, geom);
        check_state(op, *geom, q);
        check_mesh(op, *geom, mesh);
        check_physics(op, *geom, phys);
        if (!phys->use_bc) { GUST_FAIL("gust: %s: needs the boundary conditions", op); }
        hipStream_t st = static_cast<hipStream_t>(stream);
        gust_array4 local;
        if (wall_on(*phys)) {
            if (ss) {
                check_ss(op, *geom, ss);
            } else {
                local = scratch_view(st, kSlotSS, slab_box(*geom), GUST_SS_N);
                ss = &local;
            }
        }
        fill_ghosts_impl(st, *geom, *q, mesh->z_nd, *phys, ss);
    });
}

int gust_fill_ghosts_homogeneous(const gust_geometry* geom, gust_state* q, const gust_mesh* mesh,
                                 const gust_physics* phys, void* stream)
{
    return guarded([&] {
        const char* op = "fill_ghosts";
        check_geom(op, geom);
        check_state(op, *geom, q);
        check_mesh(op, *geom, mesh);
        check_physics(op, *geom, phys);
        if (!phys->use_bc) { GUST_FAIL("gust: %s: needs the boundary conditions", op); }
        fill_ghosts_impl(static_cast<hipStream_t>(stream), *geom, *q, mesh->z_nd, *phys, nullptr, true);
    });
}

int gust_residual(const gust_geometry* geom, const gust_state* q, const gust_mesh* mesh, const gust_physics* phys,
                  gust_residuals* R, gust_closure_fields* cf, gust_array4* ss, void* stream)
{
    return guarded([&] {
        const char* op = "residual";
        check_geom(op, geom);
        const gust_geometry& g = *geom;
        check_state(op, g, q);
        check_mesh(op, g, mesh);
        check_physics(op, g, phys);
        const gust_physics& p = *phys;
        if (!R) { GUST_FAIL("gust: %s: null residual fields", op); }
        const int locs[6] = {GUST_LOC_XFACE, GUST_LOC_YFACE, GUST_LOC_ZFACE, GUST_LOC_CELL, GUST_LOC_CELL, GUST_LOC_CELL};
        gust_array4* Rw[6];
        for (int n = 0; n < 6; ++n) {
            Rw[n] = const_cast<gust_array4*>(R->R[n]);
            check_field(op, "R", Rw[n], g, locs[n], 0);
        }
        hipStream_t st = static_cast<hipStream_t>(stream);
        const IBox cells = loc_box(g, GUST_LOC_CELL);
        const bool wall = wall_on(p);

        // the surface state
        gust_array4 ss_local;
        if (wall) {
            if (ss) {
                check_ss(op, g, ss);
            } else {
                ss_local = scratch_view(st, kSlotSS, slab_box(g), GUST_SS_N);
                ss = &ss_local;
            }
            launch_surface(st, g, q->f[0], q->f[1], mesh->z_nd, p.surface, ss);
        }
        // the closure's fields
        gust_closure_fields cl_local;
        gust_array4 cl_views[7];
        const gust_array4 *nu = p.nu, *gth = p.gamma_theta, *gk = p.gamma_tke;
        if (p.use_closure) {
            if (!cf) {
                const IBox c1 = grow(cells, 1, 1, 1);
                cl_views[0] = scratch_view(st, kSlotNu, c1);
                cl_views[1] = scratch_view(st, kSlotGth, c1);
                cl_views[2] = scratch_view(st, kSlotGk, c1);
                cl_views[3] = scratch_view(st, kSlotPs, cells);
                cl_views[4] = scratch_view(st, kSlotPb, cells);
                cl_views[5] = scratch_view(st, kSlotEps, cells);
                cl_views[6] = scratch_view(st, kSlotDs, cells);
                cl_local = gust_closure_fields{&cl_views[0], &cl_views[1], &cl_views[2], &cl_views[3],
                                               &cl_views[4], &cl_views[5], &cl_views[6]};
                cf = &cl_local;
            } else {
                check_field(op, "nu", cf->nu, g, GUST_LOC_CELL, 1);
                check_field(op, "gamma_theta", cf->gamma_theta, g, GUST_LOC_CELL, 1);
                check_field(op, "gamma_tke", cf->gamma_tke, g, GUST_LOC_CELL, 1);
                for (gust_array4* a : {cf->Ps, cf->Pb, cf->eps, cf->dSdk}) { check_field(op, "a source", a, g, GUST_LOC_CELL, 0); }
            }
            gust_array4 dist = scratch_view(st, kSlotDist, grow(cells, 1, 1, 1));
            launch_heights(st, g, mesh->z_nd, &dist);
            const double z0 = wall ? p.surface.z0 : 0.0;
            launch_closure(st, g, *q, mesh->z_nd, &dist, z0, p.closure, p.forcing.gravity, p.forcing.theta0, *cf);
            nu = cf->nu;
            gth = cf->gamma_theta;
            gk = cf->gamma_tke;
        }
        // transport fluxes with one ghost
        gust_array4 F[3];
        for (int d = 0; d < 3; ++d) { F[d] = scratch_view(st, kSlotF + d, grow(loc_box(g, 1 << d), 1, 1, 1)); }
        check_rc(gust_transport_fluxes(geom, q->f[0], q->f[1], q->f[2], mesh->z_nd, mesh->ax, mesh->ay, &F[0], &F[1],
                                       &F[2], stream));
        // momentum: C, then - D with the wall model's ground fluxes, + G p
        gust_array4 G[3][3], Gt[3], Gk[3], tmp[3];
        for (int c = 0; c < 3; ++c) {
            const int loc = 1 << c;
            check_rc(gust_convection(loc, geom, q->f[c], &F[0], &F[1], &F[2], mesh->detJ, p.momentum_limiter,
                                     p.momentum_eps, 0, Rw[c], stream));
            for (int d = 0; d < 3; ++d) { G[c][d] = scratch_view(st, kSlotG + 3 * c + d, cv_box(g, loc, d)); }
            check_rc(gust_momentum_diffusive_fluxes(c, geom, q->f[0], q->f[1], q->f[2], nu, mesh->z_nd, p.face_mean,
                                                    &G[c][0], &G[c][1], &G[c][2], stream));
            tmp[c] = scratch_view(st, kSlotTmp + c, loc_box(g, loc));
        }
        for (int d = 0; d < 3; ++d) {
            Gt[d] = scratch_view(st, kSlotG + 9 + d, cv_box(g, GUST_LOC_CELL, d));
            Gk[d] = scratch_view(st, kSlotG + 12 + d, cv_box(g, GUST_LOC_CELL, d));
        }
        check_rc(gust_scalar_diffusive_fluxes(geom, q->f[4], gth, mesh->z_nd, p.face_mean, &Gt[0], &Gt[1], &Gt[2],
                                              stream));
        check_rc(gust_scalar_diffusive_fluxes(geom, q->f[5], gk, mesh->z_nd, p.face_mean, &Gk[0], &Gk[1], &Gk[2],
                                              stream));
        if (wall) {
            IBox cols = grow(loc_box(g, GUST_LOC_CELL), 0, 0, 0);
            cols.hi[0] += 1;
            cols.hi[1] += 1;
            cols.hi[2] = cols.lo[2];
            const Box3 b = box3(cols);
            ground_flux_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, ss_view(ss), dview<double>(&G[0][2]),
                                                                          dview<double>(&G[1][2]), dview<double>(&Gt[2]),
                                                                          dview<double>(&Gk[2]), g.lo[2], g.hi[0],
                                                                          g.hi[1]);
            GUST_HIP(hipGetLastError());
        }
        for (int c = 0; c < 3; ++c) {
            check_rc(gust_flux_divergence(1 << c, geom, &G[c][0], &G[c][1], &G[c][2], mesh->detJ, &tmp[c], stream));
        }
        cgrid_sub3(st, g, Rw, tmp, -1.0);
        check_rc(gust_pressure_gradient(geom, q->f[3], mesh->z_nd, mesh->detJ, mesh->ax, mesh->ay, &tmp[0], &tmp[1],
                                        &tmp[2], stream));
        cgrid_sub3(st, g, Rw, tmp, 1.0);
        // forcing
        {
            gust_array4 tr = scratch_view(st, kSlotTref, grow(cells, 1, 1, 1));
            const Sounding prof = heights::sounding(p.n_theta_ref, p.theta_ref_z, p.theta_ref);
            const Box3 tb = box3(grow(cells, 1, 1, 1));
            theta_ref_kernel<<<blocks_for(tb.npts()), kThreads, 0, st>>>(tb, dview<const double>(mesh->z_nd), prof,
                                                                         dview<double>(&tr));
            const Rot rot = rotation(p.forcing);
            const Box3 bw = box3(loc_box(g, GUST_LOC_ZFACE));
            force_w_kernel<<<blocks_for(bw.npts()), kThreads, 0, st>>>(bw, dview<double>(Rw[2]),
                                                                       dview<const double>(q->f[4]),
                                                                       dview<const double>(&tr),
                                                                       dview<const double>(q->f[0]),
                                                                       p.forcing.gravity / p.forcing.theta0, rot);
            const Box3 bu = box3(loc_box(g, GUST_LOC_XFACE));
            force_u_kernel<<<blocks_for(bu.npts()), kThreads, 0, st>>>(bu, dview<double>(Rw[0]),
                                                                       dview<const double>(q->f[1]),
                                                                       dview<const double>(q->f[2]), rot);
            const Box3 bv = box3(loc_box(g, GUST_LOC_YFACE));
            force_v_kernel<<<blocks_for(bv.npts()), kThreads, 0, st>>>(bv, dview<double>(Rw[1]),
                                                                       dview<const double>(q->f[0]), rot);
            GUST_HIP(hipGetLastError());
        }
        // continuity
        check_rc(gust_velocity_divergence(geom, q->f[0], q->f[1], q->f[2], mesh->z_nd, mesh->ax, mesh->ay, mesh->az,
                                          mesh->detJ, Rw[3], stream));
        // theta: C - D
        gust_array4 d = scratch_view(st, kSlotTmp + 3, cells);
        const Box3 bc = box3(cells);
        check_rc(gust_convection(GUST_LOC_CELL, geom, q->f[4], &F[0], &F[1], &F[2], mesh->detJ, p.theta_limiter,
                                 p.theta_eps, 0, Rw[4], stream));
        check_rc(gust_flux_divergence(GUST_LOC_CELL, geom, &Gt[0], &Gt[1], &Gt[2], mesh->detJ, &d, stream));
        cgrid_sub(st, bc, Rw[4], &d, -1.0);
        // k: bounded upwind C - D - S_k, the wall-k rows
        check_rc(gust_convection(GUST_LOC_CELL, geom, q->f[5], &F[0], &F[1], &F[2], mesh->detJ, GUST_LIMITER_UPWIND,
                                 1e-5, 1, Rw[5], stream));
        check_rc(gust_flux_divergence(GUST_LOC_CELL, geom, &Gk[0], &Gk[1], &Gk[2], mesh->detJ, &d, stream));
        cgrid_sub(st, bc, Rw[5], &d, -1.0);
        if (p.use_closure) {
            k_sources_kernel<<<blocks_for(bc.npts()), kThreads, 0, st>>>(bc, dview<double>(Rw[5]),
                                                                         dview<const double>(cf->Ps),
                                                                         dview<const double>(cf->Pb),
                                                                         dview<const double>(cf->eps));
        }
        IBox ground = cells;
        ground.hi[2] = ground.lo[2];
        const Box3 bg = box3(ground);
        if (wall) {
            wall_k_kernel<<<blocks_for(bg.npts()), kThreads, 0, st>>>(bg, ss_view(ss), dview<const double>(q->f[5]),
                                                                      dview<double>(Rw[5]), g.lo[2],
                                                                      p.surface.wall_k == 0 ? 1 : 0);
        }
        if (p.use_bc) {
            bool inflow = damping_on(p);
            for (int o = 0; o < 6; ++o) { inflow = inflow || p.face[o] == GUST_FACE_INFLOW; }
            const profile::ProfileView prof = inflow ? profile::device_view(st, p.inflow) : profile::ProfileView{};
            if (damping_on(p)) {
                const DampCoef dcf = damp_coef(p);
                const int ff[5] = {F_U, F_V, F_W, F_THETA, F_TKE};
                const int qn[5] = {0, 1, 2, 4, 5};
                for (int n = 0; n < 5; ++n) {
                    const int loc = n < 3 ? (1 << n) : GUST_LOC_CELL;
                    const Box3 b = box3(loc_box(g, loc));
                    dispatch_loc(loc, [&](auto tc) {
                        constexpr int T = decltype(tc)::value;
                        damping_kernel<T><<<blocks_for(b.npts()), kThreads, 0, st>>>(
                            b, dview<double>(Rw[qn[n]]), dview<const double>(q->f[qn[n]]), dcf, ff[n],
                            p.damping.rayleigh_fields[n], p.damping.rayleigh_above_ground, node_z(mesh->z_nd), prof,
                            g.lo[2], g.dx[0], g.dx[1]);
                    });
                    GUST_HIP(hipGetLastError());
                }
            }
            for (int d = 0; d < 2; ++d) {
                if (g.periodic[d]) { continue; }
                const int tl = p.face[d], th = p.face[d + 3];
                const int loc = 1 << d;
                IBox plane = loc_box(g, loc);
                plane.hi[d] = plane.lo[d];
                const Box3 b = box3(plane);
                auto launch = [&](auto tc) {
                    constexpr int T = decltype(tc)::value;
                    side_rows_kernel<T><<<blocks_for(b.npts()), kThreads, 0, st>>>(
                        b, dview<double>(Rw[d]), dview<const double>(q->f[d]), d, g.lo[d], g.hi[d] + 1,
                        dirichlet_normal(tl) ? 1 : 0, dirichlet_normal(th) ? 1 : 0, tl == GUST_FACE_INFLOW ? 1 : 0,
                        th == GUST_FACE_INFLOW ? 1 : 0, node_z(mesh->z_nd), prof, g.lo[2]);
                };
                if (d == 0) { launch(std::integral_constant<int, sc::XF>{}); }
                else { launch(std::integral_constant<int, sc::YF>{}); }
                GUST_HIP(hipGetLastError());
            }
            w_rows_kernel<<<blocks_for(bg.npts()), kThreads, 0, st>>>(
                bg, dview<double>(Rw[2]), dview<const double>(q->f[2]), dview<const double>(q->f[0]),
                dview<const double>(q->f[1]), dview<const double>(mesh->z_nd), g.lo[2], g.hi[2] + 1,
                dirichlet_normal(p.face[2]) ? 1 : 0, dirichlet_normal(p.face[5]) ? 1 : 0, 1.0 / g.dx[0], 1.0 / g.dx[1]);
        }
        // OverrideSync of the face residuals
        for (int axis = 0; axis < 2; ++axis) {
            if (!g.periodic[axis]) { continue; }
            IBox plane = loc_box(g, 1 << axis);
            plane.hi[axis] = plane.lo[axis];
            const Box3 b = box3(plane);
            const int n = g.hi[axis] - g.lo[axis] + 1;
            seam_copy_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, dview<double>(Rw[axis]), axis, n);
        }
        GUST_HIP(hipGetLastError());
    });
}

int gust_implicit_rows(int32_t eq, const gust_geometry* geom, const gust_faces* F, const gust_mesh* mesh,
                       const gust_physics* phys, const gust_closure_fields* cf, const gust_array4* ss,
                       gust_array4* A, void* stream)
{
    return guarded([&] {
        const char* op = "implicit_rows";
        check_geom(op, geom);
        const gust_geometry& g = *geom;
        check_mesh(op, g, mesh);
        check_physics(op, g, phys);
        const gust_physics& p = *phys;
        int e;
        switch (eq) {
        case GUST_EQ_U: e = 0; break;
        case GUST_EQ_V: e = 1; break;
        case GUST_EQ_W: e = 2; break;
        case GUST_EQ_THETA: e = 3; break;
        case GUST_EQ_TKE: e = 4; break;
        default: GUST_FAIL("gust: %s: unknown equation %d", op, int(eq));
        }
        if (!F) { GUST_FAIL("gust: %s: null transport fluxes", op); }
        const int loc = e < 3 ? (1 << e) : GUST_LOC_CELL;
        const gust_array4* gdiff;
        if (p.use_closure) {
            if (!cf) { GUST_FAIL("gust: %s: with a closure the rows need its fields at the state (Closure::evaluate)", op); }
            gdiff = e == 3 ? cf->gamma_theta : e == 4 ? cf->gamma_tke : cf->nu;
        } else {
            gdiff = e == 3 ? p.gamma_theta : e == 4 ? p.gamma_tke : p.nu;
        }
        const bool wall = wall_on(p) && e != 2;
        if (wall) {
            if (!ss) { GUST_FAIL("gust: %s: with a wall model the first cells' rows need the surface state", op); }
            check_ss(op, g, ss);
        }
        hipStream_t st = static_cast<hipStream_t>(stream);
        check_rc(gust_implicit_coefficients(loc, geom, F->c[0], F->c[1], F->c[2], gdiff, mesh->z_nd, mesh->detJ,
                                            p.face_mean, A, stream));
        if (p.use_closure && e == 4) {
            const Box3 b = box3(loc_box(g, loc));
            add_dsdk_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, a_view(A), dview<const double>(cf->dSdk));
        }
        if (!p.use_bc && !wall) { return; }
        const int f = e == 0 ? F_U : e == 1 ? F_V : e == 2 ? F_W : e == 3 ? F_THETA : F_TKE;
        const IBox vb = loc_box(g, loc);
        if (wall) {
            IBox cols = vb;
            cols.hi[2] = cols.lo[2];
            const Box3 b = box3(cols);
            auto launch = [&](auto tc) {
                constexpr int T = decltype(tc)::value;
                wall_fold_kernel<T><<<blocks_for(b.npts()), kThreads, 0, st>>>(
                    b, a_view(A), ss_view(ss), dview<const double>(mesh->detJ), g.lo[2], e,
                    p.surface.wall_k == 0 ? 1 : 0, 1.0 / g.dx[2]);
            };
            if (loc == 1) { launch(std::integral_constant<int, 1>{}); }
            else if (loc == 2) { launch(std::integral_constant<int, 2>{}); }
            else if (loc == 4) { launch(std::integral_constant<int, 4>{}); }
            else { launch(std::integral_constant<int, 0>{}); }
            GUST_HIP(hipGetLastError());
        }
        if (!p.use_bc) { return; }
        if (damping_on(p)) {
            const Box3 b = box3(vb);
            const DampCoef dcf = damp_coef(p);
            dispatch_loc(loc, [&](auto tc) {
                constexpr int T = decltype(tc)::value;
                damping_diag_kernel<T><<<blocks_for(b.npts()), kThreads, 0, st>>>(
                    b, a_view(A), dcf, p.damping.rayleigh_fields[e], p.damping.rayleigh_above_ground,
                    node_z(mesh->z_nd), g.lo[2], g.dx[0], g.dx[1]);
            });
            GUST_HIP(hipGetLastError());
        }
        for (int d = 0; d < 3; ++d) {
            if (d < 2 && g.periodic[d]) { continue; }
            const bool nodal_d = ((loc >> d) & 1) != 0;
            Fold fl[2];
            for (int side = 0; side < 2; ++side) {
                const bool lo = side == 0;
                const int rule = rule_for(f, p.face[lo ? d : d + 3], d, lo ? 1 : 0);
                Fold& x = fl[side];
                x.dirichlet_pt = nodal_d && (rule == R_VALUE || rule == R_PROFILE);
                x.to_other = 0;
                x.fold = 0;
                if (nodal_d) { x.to_other = 1; }
                else if (rule == R_EVEN) { x.fold = -1; }
                else if (rule == R_ODD || rule == R_VALUE) { x.fold = 1; }
                else if (rule == R_GRADIENT) { x.to_other = 1; }
            }
            IBox plane = vb;
            plane.hi[d] = plane.lo[d];
            const Box3 b = box3(plane);
            bc_fold_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, a_view(A), d, vb.lo[d], vb.hi[d], fl[0],
                                                                     fl[1]);
            GUST_HIP(hipGetLastError());
        }
    });
}

}   // extern "C"
