// simplec.hip -- SIMPLEC steps on the terrain-following C-grid for gfx950.
//
// Reference behaviour (GUST):
//   Source/Simplec.cpp:87-112     face_coefficients (beta = 1/d, keep)
//   Source/Simplec.cpp:114-130    max_abs_masked
//   Source/Simplec.cpp:134-220    pressure_correction
//   Source/Simplec.H:105-123      SimplecWorkspace, PressureCorrectionResult
//
// The correction is a fixed sequence of memory-bound kernels around the
// Poisson solve (poisson.hip) on one stream. Scratch fields (rhs, phi, Omega,
// beta, keep, the fluxes) live in the per-(device, stream) workspace and are
// kept between calls, as the reference keeps its SimplecWorkspace. Maxima
// are order-independent (atomic max over the bits of |x|), so every result
// is reproducible; they return to the host where the reference reduces too.
#include "cgrid.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

namespace gust_amd {

namespace {

using namespace cgrid;

// ---------------------------------------------------------------------------
// Box kernels.

__global__ __launch_bounds__(kThreads) void nonneg_flag_kernel(Box3 b, CV d, unsigned int* flag)
{
    bool bad = false;
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double x = d(i, j, k);
        if (!(x >= 0) || !(x < INFINITY)) { bad = true; }
    }
    if (bad) { atomicOr(flag, 1u); }
}

__global__ __launch_bounds__(kThreads) void beta_keep_kernel(Box3 b, CV d, V beta, V keep)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const bool on = d(i, j, k) > 0;
        beta(i, j, k) = on ? 1.0 / d(i, j, k) : 1.0;
        keep(i, j, k) = on ? 1.0 : 0.0;
    }
}

// max |a (* m)| over the box, as ordered bits into *out (atomic max).
__global__ __launch_bounds__(kThreads) void max_abs_kernel(Box3 b, CV a, CV m, int masked,
                                                            unsigned long long* out)
{
    unsigned long long best = 0;
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double x = masked ? a(i, j, k) * m(i, j, k) : a(i, j, k);
        const unsigned long long bits = dev::abs_bits(x);
        best = bits > best ? bits : best;
    }
    best = dev::wave_max_u64(best);
    if ((threadIdx.x & 63) == 0 && best) { atomicMax(out, best); }
}

// min a over the box (a > 0), as ordered bits into *out (atomic min).
__global__ __launch_bounds__(kThreads) void min_pos_kernel(Box3 b, CV a, unsigned long long* out)
{
    unsigned long long best = ~0ull;
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const unsigned long long bits = dev::abs_bits(a(i, j, k));
        best = bits < best ? bits : best;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(best, off, 64);
        best = o < best ? o : best;
    }
    if ((threadIdx.x & 63) == 0) { atomicMin(out, best); }
}

// y op= x (op 0: y += a x, 1: y *= x, 2: y -= x, 3: y = 0).
__global__ __launch_bounds__(kThreads) void axpy_kernel(Box3 b, V y, CV x, double a, int op)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        if (op == 0) { y(i, j, k) += a * x(i, j, k); }
        else if (op == 1) { y(i, j, k) *= x(i, j, k); }
        else if (op == 2) { y(i, j, k) -= x(i, j, k); }
        else { y(i, j, k) = 0.0; }
    }
}

// MultiFab::Add on the valid box: y = y + x.
__global__ __launch_bounds__(kThreads) void add_kernel(Box3 b, V y, CV x)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        y(i, j, k) = y(i, j, k) + x(i, j, k);
    }
}

// OverrideSync of a field nodal along a periodic axis: the high seam copy
// takes the low one (the low box owns it; one box per tile).
__global__ __launch_bounds__(kThreads) void seam_kernel(Box3 b, V f, int axis, int n)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);   // b: the low seam plane
        if (axis == 0) { f(i + n, j, k) = f(i, j, k); } else { f(i, j + n, k) = f(i, j, k); }
    }
}

// FillBoundary with one ghost layer: every ghost point whose periodic image
// (x and/or y) is a valid point takes its value.
__global__ __launch_bounds__(kThreads) void fill_periodic_kernel(Box3 grown, V f, int vlo0, int vhi0, int vlo1,
                                                                  int vhi1, int vlo2, int vhi2, int perx, int pery,
                                                                  int nx, int ny)
{
    GUST_GRID_LOOP(grown)
    {
        int i, j, k;
        grown.at(t, i, j, k);
        const bool inside = i >= vlo0 && i <= vhi0 && j >= vlo1 && j <= vhi1 && k >= vlo2 && k <= vhi2;
        if (inside) { continue; }
        int a = i, c = j;
        if (perx) { if (a < vlo0) { a += nx; } else if (a > vhi0) { a -= nx; } }
        if (pery) { if (c < vlo1) { c += ny; } else if (c > vhi1) { c -= ny; } }
        if (a >= vlo0 && a <= vhi0 && c >= vlo1 && c <= vhi1 && k >= vlo2 && k <= vhi2 && (a != i || c != j)) {
            f(i, j, k) = f(a, c, k);
        }
    }
}

// ---------------------------------------------------------------------------
// Host helpers.

struct Scratch {
    double* p;
    gust_array4 v;
};

// A compact device field over box b (grown by g) from the workspace.
Scratch scratch(hipStream_t st, int slot, const IBox& valid, int g)
{
    const IBox b = grow(valid, g, g, g);
    const int64_t n = b.npts();
    Scratch s;
    s.p = static_cast<double*>(Workspace::get(st, slot, size_t(n) * sizeof(double)));
    s.v.p = s.p;
    s.v.jstride = b.len(0);
    s.v.kstride = int64_t(b.len(0)) * b.len(1);
    s.v.nstride = n;
    for (int d = 0; d < 3; ++d) {
        s.v.begin[d] = b.lo[d];
        s.v.end[d] = b.hi[d] + 1;
    }
    s.v.ncomp = 1;
    return s;
}

class Reducer {
public:
    explicit Reducer(hipStream_t st) : st_(st)
    {
        dev_ = static_cast<unsigned long long*>(Workspace::get(st, 50, 16 * sizeof(unsigned long long)));
        host_ = static_cast<unsigned long long*>(Workspace::get_host_pinned(50, 16 * sizeof(unsigned long long)));
    }
    double max_abs(const IBox& box, const gust_array4* a, const gust_array4* mask = nullptr)
    {
        const Box3 b = box3(box);
        GUST_HIP(hipMemsetAsync(dev_, 0, sizeof(unsigned long long), st_));
        max_abs_kernel<<<blocks_for(b.npts()), kThreads, 0, st_>>>(b, dview<const double>(a),
                                                                   dview<const double>(mask ? mask : a),
                                                                   mask ? 1 : 0, dev_);
        return fetch();
    }
    double min_pos(const IBox& box, const gust_array4* a)
    {
        const Box3 b = box3(box);
        GUST_HIP(hipMemsetAsync(dev_, 0xff, sizeof(unsigned long long), st_));
        min_pos_kernel<<<blocks_for(b.npts()), kThreads, 0, st_>>>(b, dview<const double>(a), dev_);
        return fetch();
    }
    bool flag_bad_nonneg(const IBox& box, const gust_array4* a)
    {
        const Box3 b = box3(box);
        GUST_HIP(hipMemsetAsync(dev_, 0, sizeof(unsigned long long), st_));
        nonneg_flag_kernel<<<blocks_for(b.npts()), kThreads, 0, st_>>>(b, dview<const double>(a),
                                                                       reinterpret_cast<unsigned int*>(dev_));
        GUST_HIP(hipGetLastError());
        GUST_HIP(hipMemcpyAsync(host_, dev_, sizeof(unsigned long long), hipMemcpyDeviceToHost, st_));
        GUST_HIP(hipStreamSynchronize(st_));
        return (host_[0] & 0xffffffffull) != 0;
    }

private:
    double fetch()
    {
        GUST_HIP(hipGetLastError());
        GUST_HIP(hipMemcpyAsync(host_, dev_, sizeof(unsigned long long), hipMemcpyDeviceToHost, st_));
        GUST_HIP(hipStreamSynchronize(st_));
        double d;
        std::memcpy(&d, host_, sizeof(double));
        return d;
    }
    hipStream_t st_;
    unsigned long long* dev_;
    unsigned long long* host_;
};

void box_op(hipStream_t st, const IBox& box, const gust_array4* y, const gust_array4* x, double a, int op)
{
    const Box3 b = box3(box);
    axpy_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, dview<double>(y), dview<const double>(x ? x : y), a,
                                                           op);
    GUST_HIP(hipGetLastError());
}

void box_add(hipStream_t st, const IBox& box, const gust_array4* y, const gust_array4* x)
{
    const Box3 b = box3(box);
    add_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, dview<double>(y), dview<const double>(x));
    GUST_HIP(hipGetLastError());
}

// OverrideSync of a field at loc on a tile's single box.
void override_sync(hipStream_t st, const gust_geometry& g, int loc, const gust_array4* f)
{
    for (int axis = 0; axis < 2; ++axis) {
        if (!g.periodic[axis] || !((loc >> axis) & 1)) { continue; }
        IBox plane = loc_box(g, loc);
        plane.hi[axis] = plane.lo[axis];
        const Box3 b = box3(plane);
        const int n = g.hi[axis] - g.lo[axis] + 1;
        seam_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, dview<double>(f), axis, n);
        GUST_HIP(hipGetLastError());
    }
}

// FillBoundary (periodic images, one ghost layer) of a field at loc.
void fill_boundary(hipStream_t st, const gust_geometry& g, int loc, const gust_array4* f)
{
    if (!g.periodic[0] && !g.periodic[1]) { return; }
    const IBox vb = loc_box(g, loc);
    const Box3 gb = box3(grow(vb, 1, 1, 1));
    const int nx = g.hi[0] - g.lo[0] + 1, ny = g.hi[1] - g.lo[1] + 1;
    fill_periodic_kernel<<<blocks_for(gb.npts()), kThreads, 0, st>>>(gb, dview<double>(f), vb.lo[0], vb.hi[0],
                                                                     vb.lo[1], vb.hi[1], vb.lo[2], vb.hi[2],
                                                                     g.periodic[0] ? 1 : 0, g.periodic[1] ? 1 : 0,
                                                                     nx, ny);
    GUST_HIP(hipGetLastError());
}

// Propagate a failure of another entry point (its message is already set).
void check_rc(int rc)
{
    if (rc != 0) { throw Fail{rc}; }
}

}   // namespace

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_pressure_correction(gust_poisson* solver, const gust_geometry* geom, gust_array4* u, gust_array4* v,
                             gust_array4* w, gust_array4* p, const gust_array4* z_nd, const gust_array4* ax,
                             const gust_array4* ay, const gust_array4* az, const gust_array4* detJ,
                             const gust_array4* d_u, const gust_array4* d_v, const gust_array4* d_w,
                             const int32_t* bc, double alpha_p, const gust_array4* source,
                             int32_t project_net_source, gust_pressure_correction_result* result, void* stream)
{
    return guarded([&] {
        const char* op = "pressure_correction";
        hipStream_t st = (hipStream_t)stream;
        check_geom(op, geom);
        if (!solver) { GUST_FAIL("gust: pressure_correction: the Poisson solver is null"); }
        if (std::isnan(alpha_p) || !(alpha_p > 0 && alpha_p <= 1)) {
            GUST_FAIL("gust: pressure_correction: alpha_p must be > 0 and <= 1");
        }
        const gust_geometry& g = *geom;
        check_field(op, "d_u", d_u, g, GUST_LOC_XFACE, 0);
        check_field(op, "d_v", d_v, g, GUST_LOC_YFACE, 0);
        check_field(op, "d_w", d_w, g, GUST_LOC_ZFACE, 0);
        check_field(op, "u", u, g, GUST_LOC_XFACE, 1);
        check_field(op, "v", v, g, GUST_LOC_YFACE, 1);
        check_field(op, "w", w, g, GUST_LOC_ZFACE, 1);
        check_field(op, "p", p, g, GUST_LOC_CELL, 1);
        if (source) { check_field(op, "source", source, g, GUST_LOC_CELL, 0); }
        check_field(op, "detJ", detJ, g, GUST_LOC_CELL, 0);
        check_nodal_heights(op, z_nd, g);
        const IBox cells = loc_box(g, GUST_LOC_CELL);
        const IBox fb[3] = {loc_box(g, GUST_LOC_XFACE), loc_box(g, GUST_LOC_YFACE), loc_box(g, GUST_LOC_ZFACE)};
        const gust_array4* dd[3] = {d_u, d_v, d_w};
        const char* dname[3] = {"d_u", "d_v", "d_w"};
        Scratch rhs = scratch(st, 30, cells, 0), phi = scratch(st, 31, cells, 1), omega = scratch(st, 32, fb[2], 0),
                div = scratch(st, 33, cells, 0);
        Scratch beta[3], keep[3], F[3];
        Reducer red(st);
        // face_coefficients (Simplec.cpp:87-112)
        for (int d = 0; d < 3; ++d) {
            beta[d] = scratch(st, 34 + d, fb[d], 0);
            keep[d] = scratch(st, 37 + d, fb[d], 0);
            F[d] = scratch(st, 40 + d, fb[d], 0);
            if (red.flag_bad_nonneg(fb[d], dd[d])) {
                GUST_FAIL("gust: pressure_correction: the SIMPLEC denominator %s must be finite and >= 0 everywhere",
                          dname[d]);
            }
            const Box3 b = box3(fb[d]);
            beta_keep_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, dview<const double>(dd[d]),
                                                                        dview<double>(&beta[d].v),
                                                                        dview<double>(&keep[d].v));
            GUST_HIP(hipGetLastError());
        }
        gust_pressure_correction_result res{};
        check_rc(gust_velocity_divergence(geom, u, v, w, z_nd, ax, ay, az, detJ, &rhs.v, stream));
        res.divergence_before = red.max_abs(cells, &rhs.v);
        if (source) { box_op(st, cells, &rhs.v, source, 0.0, 2); }
        const double speed = std::max({red.max_abs(fb[0], u), red.max_abs(fb[1], v), red.max_abs(fb[2], w)});
        const double dz_min = red.min_pos(cells, detJ) * g.dx[2];
        const double inv_h = 1.0 / g.dx[0] + 1.0 / g.dx[1] + 1.0 / dz_min;
        gust_face_coeffs coeffs{&beta[0].v, &beta[1].v, &beta[2].v, nullptr, 1.0};
        box_op(st, grow(cells, 1, 1, 1), &phi.v, nullptr, 0.0, 3);
        check_rc(gust_poisson_solve(solver, &rhs.v, &phi.v, &coeffs, bc, &F[0].v, &F[1].v, &F[2].v, &res.poisson,
                                    stream));
        if (!project_net_source && res.poisson.singular &&
            std::abs(res.poisson.removed_mean) > 1.0e-8 * speed * inv_h) {
            GUST_FAIL("gust: pressure_correction: the predicted velocity has a net source the boundaries cannot "
                      "carry: J-weighted mean divergence %s 1/s against the scale %s 1/s (a boundary condition or "
                      "the predictor is inconsistent)",
                      std::to_string(res.poisson.removed_mean).c_str(), std::to_string(speed * inv_h).c_str());
        }
        // p += alpha_p phi; u, v, Omega += the masked fluxes; w follows.
        box_op(st, cells, p, &phi.v, alpha_p, 0);
        res.max_dp = alpha_p * red.max_abs(cells, &phi.v);
        res.max_du = red.max_abs(fb[0], &F[0].v, &keep[0].v);
        res.max_dv = red.max_abs(fb[1], &F[1].v, &keep[1].v);
        res.max_dw = red.max_abs(fb[2], &F[2].v, &keep[2].v);
        for (int d = 0; d < 3; ++d) { box_op(st, fb[d], &F[d].v, &keep[d].v, 0.0, 1); }
        check_rc(gust_omega_from_w(geom, u, v, w, z_nd, &omega.v, 0, stream));
        box_add(st, fb[0], u, &F[0].v);
        box_add(st, fb[1], v, &F[1].v);
        box_add(st, fb[2], &omega.v, &F[2].v);
        override_sync(st, g, GUST_LOC_XFACE, u);
        override_sync(st, g, GUST_LOC_YFACE, v);
        override_sync(st, g, GUST_LOC_ZFACE, &omega.v);
        fill_boundary(st, g, GUST_LOC_XFACE, u);
        fill_boundary(st, g, GUST_LOC_YFACE, v);
        check_rc(gust_w_from_omega(geom, u, v, &omega.v, z_nd, w, stream));
        override_sync(st, g, GUST_LOC_ZFACE, w);
        fill_boundary(st, g, GUST_LOC_ZFACE, w);
        check_rc(gust_velocity_divergence(geom, u, v, w, z_nd, ax, ay, az, detJ, &div.v, stream));
        res.divergence_after = red.max_abs(cells, &div.v);
        if (result) { *result = res; }
    });
}

}   // extern "C"
