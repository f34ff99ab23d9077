// pseudotime.hip -- the local pseudo-time step and its face means for gfx950.
//
// Reference behaviour (GUST):
//   Source/PseudoTime.H:8-38, PseudoTime.cpp:12-51   local_dtau
//   Source/PseudoTime.cpp:53-61                      ser_cfl (host scalar)
//   Source/Predictor.cpp:30-77                       face_dtau
//
// local_dtau is one kernel per cell: the cell's two z-face Omegas are formed
// in registers (the expression omega_from_w evaluates), so no Omega field is
// written. face_dtau checks dtau_cc (finite, > 0) with a flag reduction and
// then forms the three face fields in one launch each, periodic images read
// by index wrap instead of a ghost fill.
#include "cgrid.hpp"
#include "deferred.hpp"

#include <cmath>

namespace gust_amd {

namespace {

using namespace cgrid;

constexpr double kMinRate = 1.0e-12;   // Source/PseudoTime.H:27

__global__ __launch_bounds__(kThreads) void local_dtau_kernel(Box3 b, CV u, CV v, CV w, CV z, CV J, CV nu, int has_nu,
                                                               double cfl, V dtau, double dxinv, double dyinv,
                                                               double dzinv, double diff_h)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double uc = cell_mean(u, i, j, k, 0), vc = cell_mean(v, i, j, k, 1);
        double s = 0.0;
        s += omega_at(u, v, w, z, i, j, k, dxinv, dyinv);
        s += omega_at(u, v, w, z, i, j, k + 1, dxinv, dyinv);
        const double oc = s * 0.5;
        double rate = fabs(uc) * dxinv + fabs(vc) * dyinv + fabs(oc) * dzinv / J(i, j, k);
        if (has_nu) { rate += nu(i, j, k) * diff_h; }
        dtau(i, j, k) = cfl / (rate < kMinRate ? kMinRate : rate);
    }
}

__global__ __launch_bounds__(kThreads) void positive_flag_kernel(Box3 b, CV d, unsigned int* flag)
{
    bool bad = false;
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double x = d(i, j, k);
        if (!(x > 0) || !(x < INFINITY)) { bad = true; }
    }
    if (bad) { atomicOr(flag, 1u); }
}

// The face means of one direction D (Source/Predictor.cpp:56-75).
template <int D>
__global__ __launch_bounds__(kThreads) void face_dtau_kernel(Box3 b, CV d, V f, int lo, int hi, int n, int per)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const int q = D == 0 ? i : D == 1 ? j : k;
        const bool at_lo = !per && q == lo, at_hi = !per && q == hi + 1;
        int qm = q - 1, qc = q;
        if (per) {   // periodic images (FillBoundary)
            if (qm < lo) { qm += n; }
            if (qc > hi) { qc -= n; }
        }
        auto val = [&](int qq) { return D == 0 ? d(qq, j, k) : D == 1 ? d(i, qq, k) : d(i, j, qq); };
        f(i, j, k) = at_lo ? val(q) : at_hi ? val(q - 1) : 0.5 * (val(qm) + val(qc));
    }
}

}   // namespace

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_local_dtau(const gust_geometry* geom, const gust_array4* u, const gust_array4* v, const gust_array4* w,
                    const gust_array4* z_nd, const gust_array4* detJ, const gust_array4* nu, double cfl,
                    gust_array4* dtau, void* stream)
{
    return guarded([&] {
        const char* op = "local_dtau";
        check_geom(op, geom);
        if (std::isnan(cfl) || !std::isfinite(cfl) || !(cfl > 0)) {
            GUST_FAIL("gust: local_dtau: cfl must be finite and > 0 (got %f)", cfl);
        }
        check_field(op, "u", u, *geom, GUST_LOC_XFACE, 1);
        check_field(op, "v", v, *geom, GUST_LOC_YFACE, 1);
        check_field(op, "w", w, *geom, GUST_LOC_ZFACE, 0);
        check_nodal_heights(op, z_nd, *geom);
        check_field(op, "detJ", detJ, *geom, GUST_LOC_CELL, 0);
        check_field(op, "dtau", dtau, *geom, GUST_LOC_CELL, 0);
        if (nu) { check_field(op, "nu", nu, *geom, GUST_LOC_CELL, 0); }
        const double dxinv = 1.0 / geom->dx[0], dyinv = 1.0 / geom->dx[1], dzinv = 1.0 / geom->dx[2];
        const double diff_h = 2 * (dxinv * dxinv + dyinv * dyinv);
        const Box3 b = box3(loc_box(*geom, GUST_LOC_CELL));
        local_dtau_kernel<<<blocks_for(b.npts()), kThreads, 0, (hipStream_t)stream>>>(
            b, dview<const double>(u), dview<const double>(v), dview<const double>(w), dview<const double>(z_nd),
            dview<const double>(detJ), nu ? dview<const double>(nu) : dview<const double>(detJ), nu ? 1 : 0, cfl,
            dview<double>(dtau), dxinv, dyinv, dzinv, diff_h);
        GUST_HIP(hipGetLastError());
    });
}

int gust_ser_cfl(double cfl, double r, double cfl0, double cfl_max, double p, double growth, double* out)
{
    return guarded([&] {
        if (!out) { GUST_FAIL("gust: ser_cfl: null output"); }
        const double gc = growth * cfl;
        const double capped = gc < cfl_max ? gc : cfl_max;   // std::min(cfl_max, g cfl)
        if (p == 0) {
            *out = cfl0 < capped ? cfl0 : capped;
            return;
        }
        const double r_min = std::pow(cfl0 / cfl_max, 1.0 / p);
        if (!(r > r_min)) {
            *out = capped;
            return;
        }
        const double law = cfl0 * std::pow(r, -p);
        *out = law < capped ? law : capped;
    });
}

int gust_face_dtau(const gust_geometry* geom, const gust_array4* dtau_cc, gust_array4* dtau_x, gust_array4* dtau_y,
                   gust_array4* dtau_z, void* stream)
{
    return guarded([&] {
        const char* op = "face_dtau";
        check_geom(op, geom);
        check_field(op, "dtau_cc", dtau_cc, *geom, GUST_LOC_CELL, 0);
        check_field(op, "dtau_x", dtau_x, *geom, GUST_LOC_XFACE, 0);
        check_field(op, "dtau_y", dtau_y, *geom, GUST_LOC_YFACE, 0);
        check_field(op, "dtau_z", dtau_z, *geom, GUST_LOC_ZFACE, 0);
        hipStream_t st = (hipStream_t)stream;
        const Box3 bc = box3(loc_box(*geom, GUST_LOC_CELL));
        // the positivity flag is read with the caller's other checks
        DeferredUse dq(st);
        auto* flag = reinterpret_cast<unsigned int*>(dq->zero_slot([](unsigned long long f) {
            if (f) { GUST_FAIL("gust: face_dtau: dtau_cc must be finite and > 0 everywhere"); }
        }));
        positive_flag_kernel<<<blocks_for(bc.npts()), kThreads, 0, st>>>(bc, dview<const double>(dtau_cc), flag);
        const CV d = dview<const double>(dtau_cc);
        const int nx = geom->hi[0] - geom->lo[0] + 1, ny = geom->hi[1] - geom->lo[1] + 1;
        const int nz = geom->hi[2] - geom->lo[2] + 1;
        const Box3 bx = box3(loc_box(*geom, GUST_LOC_XFACE));
        const Box3 by = box3(loc_box(*geom, GUST_LOC_YFACE));
        const Box3 bz = box3(loc_box(*geom, GUST_LOC_ZFACE));
        face_dtau_kernel<0><<<blocks_for(bx.npts()), kThreads, 0, st>>>(bx, d, dview<double>(dtau_x), geom->lo[0],
                                                                        geom->hi[0], nx, geom->periodic[0] ? 1 : 0);
        face_dtau_kernel<1><<<blocks_for(by.npts()), kThreads, 0, st>>>(by, d, dview<double>(dtau_y), geom->lo[1],
                                                                        geom->hi[1], ny, geom->periodic[1] ? 1 : 0);
        face_dtau_kernel<2><<<blocks_for(bz.npts()), kThreads, 0, st>>>(bz, d, dview<double>(dtau_z), geom->lo[2],
                                                                        geom->hi[2], nz, 0);
        GUST_HIP(hipGetLastError());
        dq.finish();
    });
}

}   // extern "C"
