// transport.hip -- convection and diffusion over the C-grid's control volumes
// for gfx950.
//
// Reference behaviour (GUST):
//   Source/CVFluxes.cpp:16-80        cv_flux_box, flux_divergence
//   Source/Limiter_K.H:23-69         limited_slope, muscl_face
//   Source/Convection.cpp:16-155     convective fluxes, convection, bounded form
//   Source/Diffusion.cpp:10-119      scalar and momentum diffusive fluxes
//   Source/Diffusion_K.H:22-100      face means, area_dot, stress_row
//   Source/Stencil_K.H:111-127       grad_at
//
// Each operator exists twice: the reference's two-pass form (the fluxes G on
// the control-volume faces, then flux_divergence), and a fused kernel that
// forms a point's six face fluxes in registers and differences them at once.
// The fused form reads the inputs once and writes only the result; a face's
// flux is evaluated by both points that share it, with the same expression,
// so both forms give the same bits (and the reference's, in FP64).
#include "cgrid.hpp"
#include "stencil.hpp"

#include <cfloat>
#include <cmath>

namespace gust_amd {

namespace {

using namespace cgrid;
namespace sc = stencil_cg;

enum { LIM_VAN_LEER = 0, LIM_VAN_ALBADA = 1, LIM_UPWIND = 2 };

template <int L>
__device__ __forceinline__ double limited_slope(double a, double b, double eps2)
{
    if constexpr (L == LIM_VAN_LEER) {
        const double ab = a * b;
        return (ab > 0) ? 2.0 * ab / (a + b) : 0.0;
    } else if constexpr (L == LIM_VAN_ALBADA) {
        return ((b * b + eps2) * a + (a * a + eps2) * b) / (a * a + b * b + 2.0 * eps2);
    } else {
        return 0.0;
    }
}

template <int L>
__device__ __forceinline__ double muscl_face(double tr, double qm, double q0, double q1, double qp, double eps2)
{
    if constexpr (L == LIM_UPWIND) {
        return (tr >= 0) ? q0 : q1;
    } else {
        return (tr >= 0) ? q0 + 0.5 * limited_slope<L>(q0 - qm, q1 - q0, eps2)
                         : q1 - 0.5 * limited_slope<L>(q1 - q0, qp - q1, eps2);
    }
}

__device__ __forceinline__ double log_mean(double a, double b)
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

template <int S>
__device__ __forceinline__ double face_gamma(const CV& g, int i, int j, int k, int logz)
{
    if (!logz || !sc::nodal(S, 2)) { return sc::interp<sc::CC, S>(g, i, j, k); }
    constexpr int lx = sc::lo_off(sc::CC, S, 0), ly = sc::lo_off(sc::CC, S, 1);
    constexpr int nx = sc::npts(sc::CC, S, 0), ny = sc::npts(sc::CC, S, 1);
    double sum = 0;
#pragma unroll
    for (int b = 0; b < ny; ++b)
#pragma unroll
        for (int a = 0; a < nx; ++a) { sum += log_mean(g(i + lx + a, j + ly + b, k - 1), g(i + lx + a, j + ly + b, k)); }
    return sum * (1.0 / double(nx * ny));
}

template <int TQ, int TS>
__device__ __forceinline__ void grad_at(const CV& q, int i, int j, int k, const sc::Metric& m,
                                        const double (&dxinv)[3], double (&g)[3])
{
    const double qx = sc::deriv<TQ, TS, 0>(q, i, j, k) * dxinv[0];
    const double qy = sc::deriv<TQ, TS, 1>(q, i, j, k) * dxinv[1];
    const double qz = sc::deriv<TQ, TS, 2>(q, i, j, k) * dxinv[2] / m.zz;
    g[0] = qx - m.zx * qz;
    g[1] = qy - m.zy * qz;
    g[2] = qz;
}

template <int D>
__device__ __forceinline__ double area_dot(const sc::Metric& m, const double (&v)[3])
{
    if constexpr (D == 0) { return m.zz * v[0]; }
    else if constexpr (D == 1) { return m.zz * v[1]; }
    else { return v[2] - m.zx * v[0] - m.zy * v[1]; }
}

// ---- face fluxes ----------------------------------------------------------

struct ConvArgs {
    CV q, f[3];
    double eps2;
};

template <int T, int D, int L>
__device__ __forceinline__ double conv_flux(const ConvArgs& c, int i, int j, int k)
{
    constexpr int S = T ^ (1 << D);
    constexpr int l = sc::lo_off(T, S, D);
    constexpr int ex = D == 0, ey = D == 1, ez = D == 2;
    const double tr = sc::interp<1 << D, S>(c.f[D], i, j, k);
    const int i0 = i + l * ex, j0 = j + l * ey, k0 = k + l * ez;
    return tr * muscl_face<L>(tr, c.q(i0 - ex, j0 - ey, k0 - ez), c.q(i0, j0, k0), c.q(i0 + ex, j0 + ey, k0 + ez),
                              c.q(i0 + 2 * ex, j0 + 2 * ey, k0 + 2 * ez), c.eps2);
}

struct DiffArgs {
    CV q[3];   // scalar: q[0]; momentum: u, v, w
    CV g, z;
    double dxinv[3];
    int logz;
};

template <int D>
__device__ __forceinline__ double scalar_flux(const DiffArgs& a, int i, int j, int k)
{
    constexpr int S = 1 << D;
    const sc::Metric m = sc::metric_at<S>(a.z, i, j, k, a.dxinv);
    double gr[3];
    grad_at<sc::CC, S>(a.q[0], i, j, k, m, a.dxinv, gr);
    return face_gamma<S>(a.g, i, j, k, a.logz) * area_dot<D>(m, gr);
}

template <int C, int D>
__device__ __forceinline__ double momentum_flux(const DiffArgs& a, int i, int j, int k)
{
    constexpr int T = sc::face(C);
    constexpr int S = T ^ (1 << D);
    const sc::Metric m = sc::metric_at<S>(a.z, i, j, k, a.dxinv);
    double gu[3], gv[3], gw[3];
    grad_at<sc::XF, S>(a.q[0], i, j, k, m, a.dxinv, gu);
    grad_at<sc::YF, S>(a.q[1], i, j, k, m, a.dxinv, gv);
    grad_at<sc::ZF, S>(a.q[2], i, j, k, m, a.dxinv, gw);
    const double* gc = C == 0 ? gu : C == 1 ? gv : gw;
    const double tau[3] = {gc[0] + gu[C], gc[1] + gv[C], gc[2] + gw[C]};
    return face_gamma<S>(a.g, i, j, k, a.logz) * area_dot<D>(m, tau);
}

// Flux functors: op(D, i, j, k) for a field at location T.
template <int T, int L>
struct ConvOp {
    ConvArgs a;
    template <int D>
    __device__ __forceinline__ double flux(int i, int j, int k) const { return conv_flux<T, D, L>(a, i, j, k); }
};

struct ScalarDiffOp {
    DiffArgs a;
    template <int D>
    __device__ __forceinline__ double flux(int i, int j, int k) const { return scalar_flux<D>(a, i, j, k); }
};

template <int C>
struct MomentumDiffOp {
    DiffArgs a;
    template <int D>
    __device__ __forceinline__ double flux(int i, int j, int k) const { return momentum_flux<C, D>(a, i, j, k); }
};

// ---- kernels ---------------------------------------------------------------

// G[D] on its cv_flux_box.
template <int D, class Op>
__global__ __launch_bounds__(kThreads) void flux_kernel(Box3 b, Op op, V g)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        g(i, j, k) = op.template flux<D>(i, j, k);
    }
}

// out = (1/J) sum_D (G_D(hi) - G_D(lo)) dxinv_D with G from the functor.
// bounded: out -= q div(F) / J (Source/Convection.cpp:60-84).
template <int T, class Op, bool kBounded>
__global__ __launch_bounds__(kThreads) void fused_kernel(Box3 b, Op op, CV J, V out, double dx0, double dx1,
                                                          double dx2, ConvArgs bt)
{
    constexpr int lx = sc::flux_lo(T, 0), ly = sc::flux_lo(T, 1), lz = sc::flux_lo(T, 2);
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double Jc = sc::interp<sc::CC, T>(J, i, j, k);
        double o = ((op.template flux<0>(i + lx + 1, j, k) - op.template flux<0>(i + lx, j, k)) * dx0 +
                    (op.template flux<1>(i, j + ly + 1, k) - op.template flux<1>(i, j + ly, k)) * dx1 +
                    (op.template flux<2>(i, j, k + lz + 1) - op.template flux<2>(i, j, k + lz)) * dx2) /
                   Jc;
        if constexpr (kBounded) {
            constexpr int SX = T ^ sc::XF, SY = T ^ sc::YF, SZ = T ^ sc::ZF;
            const double div =
                (sc::interp<sc::XF, SX>(bt.f[0], i + lx + 1, j, k) - sc::interp<sc::XF, SX>(bt.f[0], i + lx, j, k)) *
                    dx0 +
                (sc::interp<sc::YF, SY>(bt.f[1], i, j + ly + 1, k) - sc::interp<sc::YF, SY>(bt.f[1], i, j + ly, k)) *
                    dx1 +
                (sc::interp<sc::ZF, SZ>(bt.f[2], i, j, k + lz + 1) - sc::interp<sc::ZF, SZ>(bt.f[2], i, j, k + lz)) *
                    dx2;
            o -= bt.q(i, j, k) * div / Jc;
        }
        out(i, j, k) = o;
    }
}

template <int T>
__global__ __launch_bounds__(kThreads) void divergence_kernel(Box3 b, CV g0, CV g1, CV g2, CV J, V out, double dx0,
                                                               double dx1, double dx2)
{
    constexpr int lx = sc::flux_lo(T, 0), ly = sc::flux_lo(T, 1), lz = sc::flux_lo(T, 2);
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        out(i, j, k) = ((g0(i + lx + 1, j, k) - g0(i + lx, j, k)) * dx0 + (g1(i, j + ly + 1, k) - g1(i, j + ly, k)) * dx1 +
                        (g2(i, j, k + lz + 1) - g2(i, j, k + lz)) * dx2) /
                       sc::interp<sc::CC, T>(J, i, j, k);
    }
}

// ---- host helpers -------------------------------------------------------------

IBox cv_flux_box(const gust_geometry& g, int loc, int d)
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

void check_cv_fluxes(const char* op, const gust_geometry& g, int loc, gust_array4* const G[3])
{
    for (int d = 0; d < 3; ++d) {
        if (!covers(G[d], cv_flux_box(g, loc, d), 1)) {
            GUST_FAIL("gust: %s: flux G[%d] must be on the control-volume faces of the field, on the mesh's layout, "
                      "with one ghost cell",
                      op, d);
        }
    }
}

template <class Op>
void launch_fluxes(const gust_geometry& g, int loc, const Op& op, gust_array4* const G[3], hipStream_t st)
{
    const Box3 b0 = box3(cv_flux_box(g, loc, 0)), b1 = box3(cv_flux_box(g, loc, 1)), b2 = box3(cv_flux_box(g, loc, 2));
    flux_kernel<0><<<blocks_for(b0.npts()), kThreads, 0, st>>>(b0, op, dview<double>(G[0]));
    flux_kernel<1><<<blocks_for(b1.npts()), kThreads, 0, st>>>(b1, op, dview<double>(G[1]));
    flux_kernel<2><<<blocks_for(b2.npts()), kThreads, 0, st>>>(b2, op, dview<double>(G[2]));
    GUST_HIP(hipGetLastError());
}

template <int T, class Op, bool kB>
void launch_fused(const gust_geometry& g, const Op& op, const gust_array4* J, gust_array4* out, const ConvArgs& bt,
                  hipStream_t st)
{
    const Box3 b = box3(loc_box(g, T));
    fused_kernel<T, Op, kB><<<blocks_for(b.npts()), kThreads, 0, st>>>(b, op, dview<const double>(J),
                                                                       dview<double>(out), 1.0 / g.dx[0],
                                                                       1.0 / g.dx[1], 1.0 / g.dx[2], bt);
    GUST_HIP(hipGetLastError());
}

void check_convection_options(const char* op, int limiter, double eps)
{
    if (limiter < 0 || limiter > 2) { GUST_FAIL("gust: %s: unknown limiter %d", op, limiter); }
    if (std::isnan(eps) || !(eps >= std::sqrt(DBL_MIN) && eps <= std::sqrt(DBL_MAX))) {
        GUST_FAIL("gust: %s: convection smooth_eps = %f: must be > 0 with smooth_eps^2 a finite, normal number", op,
                  eps);
    }
}

ConvArgs conv_args(const gust_array4* q, const gust_array4* Fx, const gust_array4* Fy, const gust_array4* Fz,
                   double eps)
{
    ConvArgs a;
    a.q = dview<const double>(q);
    a.f[0] = dview<const double>(Fx);
    a.f[1] = dview<const double>(Fy);
    a.f[2] = dview<const double>(Fz);
    a.eps2 = eps * eps;
    return a;
}

// Dispatch on (loc, limiter) to f<T, L>().
template <class F>
void by_loc_limiter(int loc, int limiter, F&& f)
{
    auto with_t = [&](auto tc) {
        constexpr int T = decltype(tc)::value;
        switch (limiter) {
        case LIM_VAN_LEER: f(std::integral_constant<int, T>{}, std::integral_constant<int, LIM_VAN_LEER>{}); break;
        case LIM_VAN_ALBADA: f(std::integral_constant<int, T>{}, std::integral_constant<int, LIM_VAN_ALBADA>{}); break;
        default: f(std::integral_constant<int, T>{}, std::integral_constant<int, LIM_UPWIND>{}); break;
        }
    };
    switch (loc) {
    case 0: with_t(std::integral_constant<int, 0>{}); break;
    case 1: with_t(std::integral_constant<int, 1>{}); break;
    case 2: with_t(std::integral_constant<int, 2>{}); break;
    default: with_t(std::integral_constant<int, 4>{}); break;
    }
}

DiffArgs diff_args(const gust_array4* q0, const gust_array4* q1, const gust_array4* q2, const gust_array4* g,
                   const gust_array4* z, const gust_geometry& geo, int logz)
{
    DiffArgs a;
    a.q[0] = dview<const double>(q0);
    a.q[1] = dview<const double>(q1 ? q1 : q0);
    a.q[2] = dview<const double>(q2 ? q2 : q0);
    a.g = dview<const double>(g);
    a.z = dview<const double>(z);
    for (int d = 0; d < 3; ++d) { a.dxinv[d] = 1.0 / geo.dx[d]; }
    a.logz = logz;
    return a;
}

void check_vel(const char* op, const gust_geometry& g, const gust_array4* u, const gust_array4* v,
               const gust_array4* w, int ngrow)
{
    check_field(op, "vel[0]", u, g, GUST_LOC_XFACE, ngrow);
    check_field(op, "vel[1]", v, g, GUST_LOC_YFACE, ngrow);
    check_field(op, "vel[2]", w, g, GUST_LOC_ZFACE, ngrow);
}

}   // namespace

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_convective_fluxes(int32_t loc, const gust_geometry* geom, const gust_array4* q, const gust_array4* Fx,
                           const gust_array4* Fy, const gust_array4* Fz, int32_t limiter, double smooth_eps,
                           gust_array4* Gx, gust_array4* Gy, gust_array4* Gz, void* stream)
{
    return guarded([&] {
        const char* op = "convective_fluxes";
        check_geom(op, geom);
        if (!valid_loc(loc)) { GUST_FAIL("gust: %s: unknown location %d", op, loc); }
        check_convection_options(op, limiter, smooth_eps);
        check_field(op, "q", q, *geom, loc, 2);
        check_field(op, "F[0]", Fx, *geom, GUST_LOC_XFACE, 1);
        check_field(op, "F[1]", Fy, *geom, GUST_LOC_YFACE, 1);
        check_field(op, "F[2]", Fz, *geom, GUST_LOC_ZFACE, 1);
        gust_array4* G[3] = {Gx, Gy, Gz};
        check_cv_fluxes(op, *geom, loc, G);
        const ConvArgs a = conv_args(q, Fx, Fy, Fz, smooth_eps);
        by_loc_limiter(loc, limiter, [&](auto tc, auto lc) {
            ConvOp<decltype(tc)::value, decltype(lc)::value> o{a};
            launch_fluxes(*geom, loc, o, G, (hipStream_t)stream);
        });
    });
}

int gust_flux_divergence(int32_t loc, const gust_geometry* geom, const gust_array4* Gx, const gust_array4* Gy,
                         const gust_array4* Gz, const gust_array4* detJ, gust_array4* out, void* stream)
{
    return guarded([&] {
        const char* op = "flux_divergence";
        check_geom(op, geom);
        if (!valid_loc(loc)) { GUST_FAIL("gust: %s: unknown location %d", op, loc); }
        check_field(op, "out", out, *geom, loc, 0);
        gust_array4* G[3] = {const_cast<gust_array4*>(Gx), const_cast<gust_array4*>(Gy), const_cast<gust_array4*>(Gz)};
        check_cv_fluxes(op, *geom, loc, G);
        check_field(op, "detJ", detJ, *geom, GUST_LOC_CELL, 1);
        const Box3 b = box3(loc_box(*geom, loc));
        auto go = [&](auto kern) {
            kern<<<blocks_for(b.npts()), kThreads, 0, (hipStream_t)stream>>>(
                b, dview<const double>(Gx), dview<const double>(Gy), dview<const double>(Gz),
                dview<const double>(detJ), dview<double>(out), 1.0 / geom->dx[0], 1.0 / geom->dx[1],
                1.0 / geom->dx[2]);
        };
        switch (loc) {
        case 0: go(divergence_kernel<0>); break;
        case 1: go(divergence_kernel<1>); break;
        case 2: go(divergence_kernel<2>); break;
        default: go(divergence_kernel<4>); break;
        }
        GUST_HIP(hipGetLastError());
    });
}

int gust_convection(int32_t loc, const gust_geometry* geom, const gust_array4* q, const gust_array4* Fx,
                    const gust_array4* Fy, const gust_array4* Fz, const gust_array4* detJ, int32_t limiter,
                    double smooth_eps, int32_t bounded, gust_array4* out, void* stream)
{
    return guarded([&] {
        const char* op = "convection";
        check_geom(op, geom);
        if (!valid_loc(loc)) { GUST_FAIL("gust: %s: unknown location %d", op, loc); }
        check_convection_options(op, limiter, smooth_eps);
        check_field(op, "q", q, *geom, loc, 2);
        check_field(op, "F[0]", Fx, *geom, GUST_LOC_XFACE, 1);
        check_field(op, "F[1]", Fy, *geom, GUST_LOC_YFACE, 1);
        check_field(op, "F[2]", Fz, *geom, GUST_LOC_ZFACE, 1);
        check_field(op, "detJ", detJ, *geom, GUST_LOC_CELL, 1);
        check_field(op, "out", out, *geom, loc, 0);
        const ConvArgs a = conv_args(q, Fx, Fy, Fz, smooth_eps);
        by_loc_limiter(loc, limiter, [&](auto tc, auto lc) {
            constexpr int T = decltype(tc)::value;
            ConvOp<T, decltype(lc)::value> o{a};
            if (bounded) { launch_fused<T, decltype(o), true>(*geom, o, detJ, out, a, (hipStream_t)stream); }
            else { launch_fused<T, decltype(o), false>(*geom, o, detJ, out, a, (hipStream_t)stream); }
        });
    });
}

int gust_scalar_diffusive_fluxes(const gust_geometry* geom, const gust_array4* q, const gust_array4* gamma,
                                 const gust_array4* z_nd, int32_t face_mean, gust_array4* Gx, gust_array4* Gy,
                                 gust_array4* Gz, void* stream)
{
    return guarded([&] {
        const char* op = "scalar_diffusive_fluxes";
        check_geom(op, geom);
        check_field(op, "q", q, *geom, GUST_LOC_CELL, 1);
        check_field(op, "gamma", gamma, *geom, GUST_LOC_CELL, 1);
        check_nodal_heights(op, z_nd, *geom);
        gust_array4* G[3] = {Gx, Gy, Gz};
        check_cv_fluxes(op, *geom, GUST_LOC_CELL, G);
        ScalarDiffOp o{diff_args(q, nullptr, nullptr, gamma, z_nd, *geom, face_mean ? 1 : 0)};
        launch_fluxes(*geom, GUST_LOC_CELL, o, G, (hipStream_t)stream);
    });
}

int gust_scalar_diffusion(const gust_geometry* geom, const gust_array4* q, const gust_array4* gamma,
                          const gust_array4* z_nd, const gust_array4* detJ, int32_t face_mean, gust_array4* out,
                          void* stream)
{
    return guarded([&] {
        const char* op = "scalar_diffusion";
        check_geom(op, geom);
        check_field(op, "q", q, *geom, GUST_LOC_CELL, 1);
        check_field(op, "gamma", gamma, *geom, GUST_LOC_CELL, 1);
        check_nodal_heights(op, z_nd, *geom);
        check_field(op, "detJ", detJ, *geom, GUST_LOC_CELL, 0);
        check_field(op, "out", out, *geom, GUST_LOC_CELL, 0);
        ScalarDiffOp o{diff_args(q, nullptr, nullptr, gamma, z_nd, *geom, face_mean ? 1 : 0)};
        launch_fused<0, ScalarDiffOp, false>(*geom, o, detJ, out, ConvArgs{}, (hipStream_t)stream);
    });
}

int gust_momentum_diffusive_fluxes(int32_t c, const gust_geometry* geom, const gust_array4* u, const gust_array4* v,
                                   const gust_array4* w, const gust_array4* nu, const gust_array4* z_nd,
                                   int32_t face_mean, gust_array4* Gx, gust_array4* Gy, gust_array4* Gz,
                                   void* stream)
{
    return guarded([&] {
        const char* op = "momentum_diffusive_fluxes";
        check_geom(op, geom);
        if (c < 0 || c > 2) { GUST_FAIL("gust: %s: the component must be 0, 1 or 2", op); }
        check_vel(op, *geom, u, v, w, 2);
        check_field(op, "nu", nu, *geom, GUST_LOC_CELL, 1);
        check_nodal_heights(op, z_nd, *geom);
        gust_array4* G[3] = {Gx, Gy, Gz};
        check_cv_fluxes(op, *geom, 1 << c, G);
        const DiffArgs a = diff_args(u, v, w, nu, z_nd, *geom, face_mean ? 1 : 0);
        hipStream_t st = (hipStream_t)stream;
        if (c == 0) { launch_fluxes(*geom, 1, MomentumDiffOp<0>{a}, G, st); }
        else if (c == 1) { launch_fluxes(*geom, 2, MomentumDiffOp<1>{a}, G, st); }
        else { launch_fluxes(*geom, 4, MomentumDiffOp<2>{a}, G, st); }
    });
}

int gust_momentum_diffusion(const gust_geometry* geom, const gust_array4* u, const gust_array4* v,
                            const gust_array4* w, const gust_array4* nu, const gust_array4* z_nd,
                            const gust_array4* detJ, int32_t face_mean, gust_array4* out_u, gust_array4* out_v,
                            gust_array4* out_w, void* stream)
{
    return guarded([&] {
        const char* op = "momentum_diffusion";
        check_geom(op, geom);
        check_vel(op, *geom, u, v, w, 2);
        check_field(op, "nu", nu, *geom, GUST_LOC_CELL, 1);
        check_nodal_heights(op, z_nd, *geom);
        check_field(op, "detJ", detJ, *geom, GUST_LOC_CELL, 1);
        check_field(op, "out[0]", out_u, *geom, GUST_LOC_XFACE, 0);
        check_field(op, "out[1]", out_v, *geom, GUST_LOC_YFACE, 0);
        check_field(op, "out[2]", out_w, *geom, GUST_LOC_ZFACE, 0);
        const DiffArgs a = diff_args(u, v, w, nu, z_nd, *geom, face_mean ? 1 : 0);
        hipStream_t st = (hipStream_t)stream;
        launch_fused<1, MomentumDiffOp<0>, false>(*geom, MomentumDiffOp<0>{a}, detJ, out_u, ConvArgs{}, st);
        launch_fused<2, MomentumDiffOp<1>, false>(*geom, MomentumDiffOp<1>{a}, detJ, out_v, ConvArgs{}, st);
        launch_fused<4, MomentumDiffOp<2>, false>(*geom, MomentumDiffOp<2>{a}, detJ, out_w, ConvArgs{}, st);
    });
}

}   // extern "C"
