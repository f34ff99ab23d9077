// implicit.hip -- the first-order implicit operator A1 for gfx950.
//
// Reference behaviour (GUST):
//   Source/ImplicitOperator.H:1-75, ImplicitOperator.cpp:13-135
//   Source/Diffusion_K.H:22-64 (the logarithmic face mean)
//
// One thread per point of the field's location writes the point's seven
// coefficients (component planes of the caller's A). The six face weights
// are evaluated from the metric, the face diffusivity and the transport flux
// interpolated to the control-volume face, in the reference's order, so FP64
// coefficients with the arithmetic face mean are its bits (the logarithmic
// mean calls the device log, within an ulp of the host's).
#include "cgrid.hpp"
#include "stencil.hpp"

#include <cfloat>

namespace gust_amd {

namespace {

using namespace cgrid;
namespace sc = stencil_cg;

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

template <int T, int C, int D>
__device__ __forceinline__ double face_weight(int side, int i, int j, int k, const CV& f, const CV& g, const CV& z,
                                              const double (&dxinv)[3], int logz)
{
    constexpr int S = T ^ (1 << D);
    const sc::Metric m = sc::metric_at<S>(z, i, j, k, dxinv);
    const double gs = face_gamma<S>(g, i, j, k, logz);
    const double tr = sc::interp<sc::face(D), S>(f, i, j, k);
    double dif;
    if constexpr (D < 2) {
        constexpr double own = (C == D) ? 2.0 : 1.0;
        dif = gs * m.zz * own * dxinv[D];
    } else {
        double nc2 = 0;
        if constexpr (C == 0) { nc2 = m.zx * m.zx; }
        if constexpr (C == 1) { nc2 = m.zy * m.zy; }
        if constexpr (C == 2) { nc2 = 1; }
        dif = gs * (1 + m.zx * m.zx + m.zy * m.zy + nc2) / m.zz * dxinv[2];
    }
    const double inflow = (side == 0) ? sc::dmax(tr, 0.0) : sc::dmax(-tr, 0.0);
    return (dif + inflow) * dxinv[D];
}

template <int C>
__global__ __launch_bounds__(kThreads) void coefficients_kernel(Box3 b, CV fx, CV fy, CV fz, CV g, CV z, CV J,
                                                                 V A, int64_t nstride, double dx0, double dx1,
                                                                 double dx2, int logz)
{
    constexpr int T = (C < 0) ? sc::CC : sc::face(C);
    constexpr int lx = sc::flux_lo(T, 0), ly = sc::flux_lo(T, 1), lz = sc::flux_lo(T, 2);
    const double dxinv[3] = {dx0, dx1, dx2};
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double Jinv = 1.0 / sc::interp<sc::CC, T>(J, i, j, k);
        const double xm = face_weight<T, C, 0>(0, i + lx, j, k, fx, g, z, dxinv, logz) * Jinv;
        const double xp = face_weight<T, C, 0>(1, i + lx + 1, j, k, fx, g, z, dxinv, logz) * Jinv;
        const double ym = face_weight<T, C, 1>(0, i, j + ly, k, fy, g, z, dxinv, logz) * Jinv;
        const double yp = face_weight<T, C, 1>(1, i, j + ly + 1, k, fy, g, z, dxinv, logz) * Jinv;
        const double zm = face_weight<T, C, 2>(0, i, j, k + lz, fz, g, z, dxinv, logz) * Jinv;
        const double zp = face_weight<T, C, 2>(1, i, j, k + lz + 1, fz, g, z, dxinv, logz) * Jinv;
        double* a = &A(i, j, k);
        a[1 * nstride] = xm;
        a[2 * nstride] = xp;
        a[3 * nstride] = ym;
        a[4 * nstride] = yp;
        a[5 * nstride] = zm;
        a[6 * nstride] = zp;
        a[0] = ((xm + xp) + (ym + yp)) + (zm + zp);
    }
}

__global__ __launch_bounds__(kThreads) void apply_kernel(Box3 b, CV A, int64_t nstride, CV x, V o)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double* a = &A(i, j, k);
        o(i, j, k) = a[0] * x(i, j, k) - (a[1 * nstride] * x(i - 1, j, k) + a[2 * nstride] * x(i + 1, j, k)) -
                     (a[3 * nstride] * x(i, j - 1, k) + a[4 * nstride] * x(i, j + 1, k)) -
                     (a[5 * nstride] * x(i, j, k - 1) + a[6 * nstride] * x(i, j, k + 1));
    }
}

}   // namespace

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_implicit_coefficients(int32_t loc, const gust_geometry* geom, const gust_array4* Fx, const gust_array4* Fy,
                               const gust_array4* Fz, const gust_array4* g, const gust_array4* z_nd,
                               const gust_array4* detJ, int32_t face_mean, gust_array4* A, void* stream)
{
    return guarded([&] {
        const char* op = "implicit_coefficients";
        check_geom(op, geom);
        if (!valid_loc(loc)) { GUST_FAIL("gust: implicit_coefficients: unknown location %d", loc); }
        if (face_mean != 0 && face_mean != 1) { GUST_FAIL("gust: implicit_coefficients: unknown face mean"); }
        check_field(op, "F[0]", Fx, *geom, GUST_LOC_XFACE, 1);
        check_field(op, "F[1]", Fy, *geom, GUST_LOC_YFACE, 1);
        check_field(op, "F[2]", Fz, *geom, GUST_LOC_ZFACE, 1);
        check_field(op, "g", g, *geom, GUST_LOC_CELL, 1);
        check_nodal_heights(op, z_nd, *geom);
        check_field(op, "detJ", detJ, *geom, GUST_LOC_CELL, 1);
        check_field(op, "A", A, *geom, loc, 0);
        if (A->ncomp != GUST_N_COEFS) { GUST_FAIL("gust: implicit_coefficients: A needs 7 components"); }
        const Box3 b = box3(loc_box(*geom, loc));
        const double d0 = 1.0 / geom->dx[0], d1 = 1.0 / geom->dx[1], d2 = 1.0 / geom->dx[2];
        auto args = [&](auto kern) {
            kern<<<blocks_for(b.npts()), kThreads, 0, (hipStream_t)stream>>>(
                b, dview<const double>(Fx), dview<const double>(Fy), dview<const double>(Fz), dview<const double>(g),
                dview<const double>(z_nd), dview<const double>(detJ), dview<double>(A), A->nstride, d0, d1, d2,
                face_mean);
        };
        switch (loc) {
        case GUST_LOC_CELL: args(coefficients_kernel<-1>); break;
        case GUST_LOC_XFACE: args(coefficients_kernel<0>); break;
        case GUST_LOC_YFACE: args(coefficients_kernel<1>); break;
        default: args(coefficients_kernel<2>); break;
        }
        GUST_HIP(hipGetLastError());
    });
}

int gust_apply_implicit(int32_t loc, const gust_geometry* geom, const gust_array4* A, const gust_array4* q,
                        gust_array4* out, void* stream)
{
    return guarded([&] {
        const char* op = "apply_implicit";
        check_geom(op, geom);
        if (!valid_loc(loc)) { GUST_FAIL("gust: apply_implicit: unknown location %d", loc); }
        check_field(op, "A", A, *geom, loc, 0);
        if (A->ncomp != GUST_N_COEFS) { GUST_FAIL("gust: apply_implicit: A needs 7 components"); }
        check_field(op, "q", q, *geom, loc, 1);
        check_field(op, "out", out, *geom, loc, 0);
        const Box3 b = box3(loc_box(*geom, loc));
        apply_kernel<<<blocks_for(b.npts()), kThreads, 0, (hipStream_t)stream>>>(
            b, dview<const double>(A), A->nstride, dview<const double>(q), dview<double>(out));
        GUST_HIP(hipGetLastError());
    });
}

}   // extern "C"
