// massflux.hip -- the terrain-following C-grid mass-flux operators for gfx950.
//
// Reference behaviour (GUST):
//   Source/MassFlux_K.H:19-39      slope_x, slope_y (z-face slopes from its nodes)
//   Source/MassFlux.cpp:29-39      slope_dp (one-sided d_zeta p on the ground face)
//   Source/MassFlux.cpp:51-74      omega_from_w
//   Source/MassFlux.cpp:76-99      w_from_omega
//   Source/MassFlux.cpp:101-127    transport_fluxes
//   Source/MassFlux.cpp:129-136    velocity_divergence (Omega, then face_divergence)
//   Source/MassFlux.cpp:138-179    pressure_gradient
//   Source/Stencil_K.H:49-99       the averages and differences they use
//
// Design: one thread per output point, i fastest (coalesced rows); every
// operand is read where the reference reads it, from the caller's views with
// the ghosts the caller filled. Expressions keep the reference's order
// (sums from 0 in loop order, then the 1/n scale), so FP64 results are the
// CPU build's bits. velocity_divergence is one kernel: the two z-face Omegas
// of a cell are formed in registers from the same expression
// omega_from_w evaluates, then combined as face_divergence combines them.
#include "cgrid.hpp"

namespace gust_amd {

namespace {

using namespace cgrid;

template <int D>
__device__ __forceinline__ double slope_dp(const CV& z, const CV& p, int i, int j, int k, int klo, double dxinv,
                                           double dyinv, double dzinv)
{
    const double s = D == 0 ? slope_x(z, i, j, k, dxinv) : slope_y(z, i, j, k, dyinv);
    const double dp = (k == klo) ? 3.0 * p(i, j, k + 1) - 2.0 * p(i, j, k) - p(i, j, k + 2) : p(i, j, k) - p(i, j, k - 1);
    return s * dp * dzinv;
}

__global__ __launch_bounds__(kThreads) void omega_kernel(Box3 b, CV u, CV v, CV w, CV z, V om, double dxinv,
                                                          double dyinv)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        om(i, j, k) = omega_at(u, v, w, z, i, j, k, dxinv, dyinv);
    }
}

__global__ __launch_bounds__(kThreads) void w_kernel(Box3 b, CV u, CV v, CV om, CV z, V w, double dxinv,
                                                      double dyinv)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        w(i, j, k) = om(i, j, k) + slope_x(z, i, j, k, dxinv) * u_at_zface(u, i, j, k) +
                     slope_y(z, i, j, k, dyinv) * v_at_zface(v, i, j, k);
    }
}

// F_x = a_x u and F_y = a_y v in one launch (blockIdx.y: the direction).
struct AreaFlux2 {
    Box3 b[2];
    CV a[2], q[2];
    V f[2];
};

__global__ __launch_bounds__(kThreads) void area_flux2_kernel(AreaFlux2 s)
{
    const int d = blockIdx.y;
    const Box3& b = s.b[d];
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        s.f[d](i, j, k) = s.a[d](i, j, k) * s.q[d](i, j, k);
    }
}

__global__ __launch_bounds__(kThreads) void divergence_kernel(Box3 b, CV u, CV v, CV w, CV z, CV ax, CV ay, CV az,
                                                               CV J, V div, double dxinv, double dyinv,
                                                               double dzinv)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double olo = omega_at(u, v, w, z, i, j, k, dxinv, dyinv);
        const double ohi = omega_at(u, v, w, z, i, j, k + 1, dxinv, dyinv);
        div(i, j, k) = ((ax(i + 1, j, k) * u(i + 1, j, k) - ax(i, j, k) * u(i, j, k)) * dxinv +
                        (ay(i, j + 1, k) * v(i, j + 1, k) - ay(i, j, k) * v(i, j, k)) * dyinv +
                        (az(i, j, k + 1) * ohi - az(i, j, k) * olo) * dzinv) /
                       J(i, j, k);
    }
}

template <int D>
__global__ __launch_bounds__(kThreads) void grad_h_kernel(Box3 b, CV p, CV z, CV J, CV a, V g, int klo, double dxinv,
                                                           double dyinv, double dzinv)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const int di = D == 0, dj = D == 1;
        const double cross =
            0.25 * ((slope_dp<D>(z, p, i - di, j - dj, k, klo, dxinv, dyinv, dzinv) +
                     slope_dp<D>(z, p, i, j, k, klo, dxinv, dyinv, dzinv)) +
                    (slope_dp<D>(z, p, i - di, j - dj, k + 1, klo, dxinv, dyinv, dzinv) +
                     slope_dp<D>(z, p, i, j, k + 1, klo, dxinv, dyinv, dzinv)));
        g(i, j, k) = (a(i, j, k) * (p(i, j, k) - p(i - di, j - dj, k)) * (D == 0 ? dxinv : dyinv) - cross) /
                     J_at_face(J, i, j, k, D);
    }
}

__global__ __launch_bounds__(kThreads) void grad_z_kernel(Box3 b, CV p, CV J, V g, double dzinv)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        g(i, j, k) = (p(i, j, k) - p(i, j, k - 1)) * dzinv / J_at_face(J, i, j, k, 2);
    }
}

}   // namespace

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_omega_from_w(const gust_geometry* geom, const gust_array4* u, const gust_array4* v, const gust_array4* w,
                      const gust_array4* z_nd, gust_array4* omega, int32_t ngrow, void* stream)
{
    return guarded([&] {
        const char* op = "omega_from_w";
        check_geom(op, geom);
        if (ngrow < 0 || ngrow > 1) { GUST_FAIL("gust: omega_from_w: ngrow must be 0 or 1"); }
        check_field(op, "u", u, *geom, GUST_LOC_XFACE, ngrow + 1);
        check_field(op, "v", v, *geom, GUST_LOC_YFACE, ngrow + 1);
        check_field(op, "w", w, *geom, GUST_LOC_ZFACE, ngrow);
        check_field(op, "omega", omega, *geom, GUST_LOC_ZFACE, ngrow);
        check_nodal_heights(op, z_nd, *geom);
        const IBox ob = grow(loc_box(*geom, GUST_LOC_ZFACE), ngrow, ngrow, ngrow);
        const Box3 b = box3(ob);
        omega_kernel<<<blocks_for(b.npts()), kThreads, 0, (hipStream_t)stream>>>(
            b, dview<const double>(u), dview<const double>(v), dview<const double>(w), dview<const double>(z_nd),
            dview<double>(omega), 1.0 / geom->dx[0], 1.0 / geom->dx[1]);
        GUST_HIP(hipGetLastError());
    });
}

int gust_w_from_omega(const gust_geometry* geom, const gust_array4* u, const gust_array4* v,
                      const gust_array4* omega, const gust_array4* z_nd, gust_array4* w, void* stream)
{
    return guarded([&] {
        const char* op = "w_from_omega";
        check_geom(op, geom);
        check_field(op, "u", u, *geom, GUST_LOC_XFACE, 1);
        check_field(op, "v", v, *geom, GUST_LOC_YFACE, 1);
        check_field(op, "omega", omega, *geom, GUST_LOC_ZFACE, 0);
        check_field(op, "w", w, *geom, GUST_LOC_ZFACE, 0);
        check_nodal_heights(op, z_nd, *geom);
        const Box3 b = box3(loc_box(*geom, GUST_LOC_ZFACE));
        w_kernel<<<blocks_for(b.npts()), kThreads, 0, (hipStream_t)stream>>>(
            b, dview<const double>(u), dview<const double>(v), dview<const double>(omega),
            dview<const double>(z_nd), dview<double>(w), 1.0 / geom->dx[0], 1.0 / geom->dx[1]);
        GUST_HIP(hipGetLastError());
    });
}

int gust_transport_fluxes(const gust_geometry* geom, const gust_array4* u, const gust_array4* v,
                          const gust_array4* w, const gust_array4* z_nd, const gust_array4* ax,
                          const gust_array4* ay, gust_array4* Fx, gust_array4* Fy, gust_array4* Fz, void* stream)
{
    return guarded([&] {
        const char* op = "transport_fluxes";
        check_geom(op, geom);
        check_field(op, "u", u, *geom, GUST_LOC_XFACE, 2);
        check_field(op, "v", v, *geom, GUST_LOC_YFACE, 2);
        check_field(op, "w", w, *geom, GUST_LOC_ZFACE, 1);
        check_field(op, "F", Fx, *geom, GUST_LOC_XFACE, 1);
        check_field(op, "F", Fy, *geom, GUST_LOC_YFACE, 1);
        check_field(op, "F", Fz, *geom, GUST_LOC_ZFACE, 1);
        check_field(op, "ax", ax, *geom, GUST_LOC_XFACE, 1);
        check_field(op, "ay", ay, *geom, GUST_LOC_YFACE, 1);
        check_nodal_heights(op, z_nd, *geom);
        hipStream_t st = (hipStream_t)stream;
        const Box3 bx = box3(grow(loc_box(*geom, GUST_LOC_XFACE), 1, 1, 1));
        const Box3 by = box3(grow(loc_box(*geom, GUST_LOC_YFACE), 1, 1, 1));
        const Box3 bz = box3(grow(loc_box(*geom, GUST_LOC_ZFACE), 1, 1, 1));
        AreaFlux2 af{{bx, by}, {dview<const double>(ax), dview<const double>(ay)},
                     {dview<const double>(u), dview<const double>(v)}, {dview<double>(Fx), dview<double>(Fy)}};
        area_flux2_kernel<<<dim3(blocks_for(std::max(bx.npts(), by.npts())), 2), kThreads, 0, st>>>(af);
        omega_kernel<<<blocks_for(bz.npts()), kThreads, 0, st>>>(
            bz, dview<const double>(u), dview<const double>(v), dview<const double>(w), dview<const double>(z_nd),
            dview<double>(Fz), 1.0 / geom->dx[0], 1.0 / geom->dx[1]);
        GUST_HIP(hipGetLastError());
    });
}

int gust_velocity_divergence(const gust_geometry* geom, const gust_array4* u, const gust_array4* v,
                             const gust_array4* w, const gust_array4* z_nd, const gust_array4* ax,
                             const gust_array4* ay, const gust_array4* az, const gust_array4* detJ,
                             gust_array4* div, void* stream)
{
    return guarded([&] {
        const char* op = "velocity_divergence";
        check_geom(op, geom);
        check_field(op, "u", u, *geom, GUST_LOC_XFACE, 1);
        check_field(op, "v", v, *geom, GUST_LOC_YFACE, 1);
        check_field(op, "w", w, *geom, GUST_LOC_ZFACE, 0);
        check_field(op, "ax", ax, *geom, GUST_LOC_XFACE, 0);
        check_field(op, "ay", ay, *geom, GUST_LOC_YFACE, 0);
        check_field(op, "az", az, *geom, GUST_LOC_ZFACE, 0);
        check_field(op, "detJ", detJ, *geom, GUST_LOC_CELL, 0);
        check_field(op, "div", div, *geom, GUST_LOC_CELL, 0);
        check_nodal_heights(op, z_nd, *geom);
        const Box3 b = box3(loc_box(*geom, GUST_LOC_CELL));
        divergence_kernel<<<blocks_for(b.npts()), kThreads, 0, (hipStream_t)stream>>>(
            b, dview<const double>(u), dview<const double>(v), dview<const double>(w), dview<const double>(z_nd),
            dview<const double>(ax), dview<const double>(ay), dview<const double>(az), dview<const double>(detJ),
            dview<double>(div), 1.0 / geom->dx[0], 1.0 / geom->dx[1], 1.0 / geom->dx[2]);
        GUST_HIP(hipGetLastError());
    });
}

int gust_pressure_gradient(const gust_geometry* geom, const gust_array4* p, const gust_array4* z_nd,
                           const gust_array4* detJ, const gust_array4* ax, const gust_array4* ay, gust_array4* gx,
                           gust_array4* gy, gust_array4* gz, void* stream)
{
    return guarded([&] {
        const char* op = "pressure_gradient";
        check_geom(op, geom);
        check_field(op, "p", p, *geom, GUST_LOC_CELL, 1);
        check_nodal_heights(op, z_nd, *geom);
        check_field(op, "detJ", detJ, *geom, GUST_LOC_CELL, 1);
        check_field(op, "ax", ax, *geom, GUST_LOC_XFACE, 0);
        check_field(op, "ay", ay, *geom, GUST_LOC_YFACE, 0);
        check_field(op, "Gp", gx, *geom, GUST_LOC_XFACE, 0);
        check_field(op, "Gp", gy, *geom, GUST_LOC_YFACE, 0);
        check_field(op, "Gp", gz, *geom, GUST_LOC_ZFACE, 0);
        if (geom->hi[2] - geom->lo[2] + 1 < 2) { GUST_FAIL("gust: pressure_gradient: the domain needs nz >= 2"); }
        hipStream_t st = (hipStream_t)stream;
        const double dxinv = 1.0 / geom->dx[0], dyinv = 1.0 / geom->dx[1], dzinv = 1.0 / geom->dx[2];
        const int klo = geom->lo[2];
        const Box3 bx = box3(loc_box(*geom, GUST_LOC_XFACE));
        const Box3 by = box3(loc_box(*geom, GUST_LOC_YFACE));
        const Box3 bz = box3(loc_box(*geom, GUST_LOC_ZFACE));
        grad_h_kernel<0><<<blocks_for(bx.npts()), kThreads, 0, st>>>(bx, dview<const double>(p),
                                                                     dview<const double>(z_nd),
                                                                     dview<const double>(detJ),
                                                                     dview<const double>(ax), dview<double>(gx), klo,
                                                                     dxinv, dyinv, dzinv);
        grad_h_kernel<1><<<blocks_for(by.npts()), kThreads, 0, st>>>(by, dview<const double>(p),
                                                                     dview<const double>(z_nd),
                                                                     dview<const double>(detJ),
                                                                     dview<const double>(ay), dview<double>(gy), klo,
                                                                     dxinv, dyinv, dzinv);
        grad_z_kernel<<<blocks_for(bz.npts()), kThreads, 0, st>>>(bz, dview<const double>(p),
                                                                  dview<const double>(detJ), dview<double>(gz),
                                                                  dzinv);
        GUST_HIP(hipGetLastError());
    });
}

}   // extern "C"
