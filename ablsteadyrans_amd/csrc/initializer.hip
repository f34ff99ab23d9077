// initializer.hip -- the mass-consistent initial state on a tile: the inflow
// profile by height above the ground (log law, power law or a table), the
// first guess with a no-flux ground, the optional O'Brien column correction
// and the anisotropic projection
//
//   div(Sigma grad lambda) = div(u0),  Sigma = diag(ah^2, ah^2, av^2),
//   (u, v, Omega) = (u0, v0, Omega0) + F,  F = -Sigma grad lambda,
//
// solved by the terrain Poisson solver (the multigrid-preconditioned GMRES
// handles the split z-face coefficient directly, so no back-end switch).
//
//   Source/InflowProfile.H:137-200      table_interp, View::eval
//   Source/InflowProfile.cpp:14-47, 178-205  option checks, the view's constants
//   Source/Initializer.cpp:25-53         fill_from_profile
//   Source/Initializer.cpp:109-195       check_init_options, initializer_poisson_bc
//   Source/Initializer.cpp:197-267       obrien_omega (here one thread per column)
//   Source/Initializer.cpp:269-418       first_guess_state, initialize_state
//
// Every point value is formed with the reference's operations in its order
// (contraction off); the device's log, exp and pow may differ from the
// host's in the last bit.
#include "cgrid.hpp"
#include "heights.hpp"
#include "profile.hpp"
#include "reduce.hpp"
#include "stencil.hpp"

#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

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
using heights::NodeZ;

// Workspace slots of this file.
constexpr int kSlotRhs = 111, kSlotLam = 112, kSlotF = 113 /* 3 slots */, kSlotOm = 116,
              kSlotCoef = 117 /* 4 slots */, kSlotDh = 121, kSlotZero = 122, kSlotWorst = 123, kSlotWorstHost = 124;

using profile::dmaxd;
using profile::profile_eval;
using profile::ProfileView;

__global__ __launch_bounds__(kThreads) void eval_kernel(ProfileView p, int n, const double* z, double* u, double* v,
                                                        double* th, double* k, double* nut)
{
    for (int t = blockIdx.x * kThreads + threadIdx.x; t < n; t += gridDim.x * kThreads) {
        double a, b, c, d, e;
        profile_eval(p, z[t], a, b, c, d, e);
        if (u) { u[t] = a; }
        if (v) { v[t] = b; }
        if (th) { th[t] = c; }
        if (k) { k[t] = d; }
        if (nut) { nut[t] = e; }
    }
}

// One profile component at every point of b, at the point's height above
// its column's ground floored at h_min; k decays above zi > 0
// (Initializer.cpp:25-53).
template <int T>
__global__ __launch_bounds__(kThreads) void fill_profile_kernel(Box3 b, V a, NodeZ z, ProfileView p, int comp,
                                                                double h_min, double zi, double zinv, int klo)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double h = dmaxd(heights::height_above_ground<T>(z, i, j, k, klo), h_min);
        double u, v, th, kk, nut;
        profile_eval(p, h, u, v, th, kk, nut);
        double val = (comp == 0) ? u : (comp == 1) ? v : (comp == 2) ? th : kk;
        if (comp == 3 && zi > 0) {
            const double arg = dmaxd(h - zi, 0.0) * zinv;
            val *= exp(-arg);
        }
        a(i, j, k) = val;
    }
}

__global__ __launch_bounds__(kThreads) void set_kernel(Box3 b, V a, double val)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        a(i, j, k) = val;
    }
}

__global__ __launch_bounds__(kThreads) void add_kernel(Box3 b, V y, CV x)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        y(i, j, k) += x(i, j, k);
    }
}

// FillBoundary(periodicity) on one tile: a ghost point whose periodic image
// is a valid point takes its value.
struct PeriodicFill {
    int vlo[3], vhi[3], n[2], per[2];
};

__global__ __launch_bounds__(kThreads) void periodic_fill_kernel(Box3 b, V a, PeriodicFill f)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const bool ghost = i < f.vlo[0] || i > f.vhi[0] || j < f.vlo[1] || j > f.vhi[1] || k < f.vlo[2] ||
                           k > f.vhi[2];
        if (!ghost) { continue; }
        int si = i, sj = j;
        if (f.per[0]) { si += (i < f.vlo[0]) ? f.n[0] : (i > f.vhi[0] ? -f.n[0] : 0); }
        if (f.per[1]) { sj += (j < f.vlo[1]) ? f.n[1] : (j > f.vhi[1] ? -f.n[1] : 0); }
        if (si < f.vlo[0] || si > f.vhi[0] || sj < f.vlo[1] || sj > f.vhi[1] || k < f.vlo[2] || k > f.vhi[2]) {
            continue;
        }
        a(i, j, k) = a(si, sj, k);
    }
}

// O'Brien: one thread per column. a_z Omega_{k+1/2} = a_z Omega_{k-1/2} -
// J_k Dh_k dzeta from Omega = 0 at the ground, then the top residual E
// removed with (height above the ground / depth)^2 (Initializer.cpp:218-261).
__global__ __launch_bounds__(kThreads) void obrien_kernel(int lo0, int lo1, int nx, int ny, int klo, int nz, CV dh,
                                                          CV J, CV az, CV zn, V om, double dzeta,
                                                          unsigned long long* worst)
{
    const int ncol = nx * ny;
    unsigned long long wmax = 0;
    for (int t = blockIdx.x * kThreads + threadIdx.x; t < ncol; t += gridDim.x * kThreads) {
        const int i = lo0 + t % nx, j = lo1 + t / nx;
        om(i, j, klo) = 0.0;
        for (int k = klo; k < klo + nz; ++k) {
            om(i, j, k + 1) = (az(i, j, k) * om(i, j, k) - J(i, j, k) * dh(i, j, k) * dzeta) / az(i, j, k + 1);
        }
        const double E = om(i, j, klo + nz);
        const double aE = fabs(E);
        if (aE == aE) { wmax = dev::abs_bits(aE) > wmax ? dev::abs_bits(aE) : wmax; }
        auto zf = [&](int k) { return 0.25 * (zn(i, j, k) + zn(i + 1, j, k) + zn(i, j + 1, k) + zn(i + 1, j + 1, k)); };
        const double zg = zf(klo), depth = zf(klo + nz) - zg;
        for (int k = klo + 1; k <= klo + nz; ++k) {
            const double frac = (zf(k) - zg) / depth;
            om(i, j, k) -= frac * frac * E;
        }
        om(i, j, klo + nz) = 0.0;
    }
    wmax = dev::block_max_u64<kThreads>(wmax);
    if (threadIdx.x == 0 && wmax) { atomicMax(worst, wmax); }
}

// ---- host side -------------------------------------------------------------------

void check_rc(int rc)
{
    if (rc != 0) { throw Fail{rc}; }
}

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

IBox view_box(const gust_array4* a)
{
    IBox b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = a->begin[d];
        b.hi[d] = a->end[d] - 1;
    }
    return b;
}

void set_box(hipStream_t st, gust_array4* a, const IBox& b, double val)
{
    const Box3 bb = box3(b);
    set_kernel<<<blocks_for(bb.npts()), kThreads, 0, st>>>(bb, dview<double>(a), val);
    GUST_HIP(hipGetLastError());
}

void add_box(hipStream_t st, gust_array4* y, const gust_array4* x, const IBox& b)
{
    const Box3 bb = box3(b);
    add_kernel<<<blocks_for(bb.npts()), kThreads, 0, st>>>(bb, dview<double>(y), dview<const double>(x));
    GUST_HIP(hipGetLastError());
}

void periodic_fill(hipStream_t st, const gust_geometry& g, gust_array4* a, int loc)
{
    if (!g.periodic[0] && !g.periodic[1]) { return; }
    const IBox vb = loc_box(g, loc);
    PeriodicFill f;
    for (int d = 0; d < 3; ++d) {
        f.vlo[d] = vb.lo[d];
        f.vhi[d] = vb.hi[d];
    }
    for (int d = 0; d < 2; ++d) {
        f.n[d] = g.hi[d] - g.lo[d] + 1;
        f.per[d] = g.periodic[d] ? 1 : 0;
    }
    const Box3 b = box3(view_box(a));
    periodic_fill_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, dview<double>(a), f);
    GUST_HIP(hipGetLastError());
}

std::vector<std::string> init_problems(const gust_init_options& o)
{
    std::vector<std::string> err;
    auto bad = [&](double x, const char* key, const char* range, bool ok) {
        if (std::isnan(x) || !std::isfinite(x) || !ok) {
            err.push_back(std::string("init.") + key + " = " + std::to_string(x) + ": must be " + range);
        }
    };
    if (o.first_guess != 0 && o.first_guess != 1) { err.push_back("init.first_guess: must be none (0) or profile (1)"); }
    if (o.projection != 0 && o.projection != 1) { err.push_back("init.projection must be 0 or 1"); }
    if (o.obrien != 0 && o.obrien != 1) { err.push_back("init.obrien must be 0 or 1"); }
    bad(o.z_agl_min, "z_agl_min", "finite and >= 0 (m)", o.z_agl_min >= 0);
    bad(o.alpha_h, "alpha_h", "finite and > 0", o.alpha_h > 0);
    bad(o.alpha_v, "alpha_v", "finite and > 0", o.alpha_v > 0);
    bad(o.k_decay_height, "k_decay_height", "finite and >= 0 (m; 0: no decay)", o.k_decay_height >= 0);
    return err;
}

void fail_first(const char* op, const std::vector<std::string>& err)
{
    if (!err.empty()) { GUST_FAIL("gust: %s: %s", op, err.front().c_str()); }
}

template <int T>
void fill_profile(hipStream_t st, const gust_geometry& g, gust_array4* a, const gust_array4* z, const ProfileView& p,
                  int comp, double h_min, double zi)
{
    const double zinv = 1.0 / std::max(zi, DBL_MIN);
    const Box3 b = box3(view_box(a));
    fill_profile_kernel<T><<<blocks_for(b.npts()), kThreads, 0, st>>>(b, dview<double>(a), heights::node_z(z), p,
                                                                       comp, h_min, zi, zinv, g.lo[2]);
    GUST_HIP(hipGetLastError());
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

// obrien_omega on validated arguments; returns max |Omega_top| before the removal.
double obrien_impl(hipStream_t st, const gust_geometry& g, const gust_mesh& mesh, const gust_array4* u,
                   const gust_array4* v, gust_array4* omega)
{
    const IBox cells = loc_box(g, GUST_LOC_CELL);
    gust_array4 dh = scratch_view(st, kSlotDh, cells);
    gust_array4 zero = scratch_view(st, kSlotZero, loc_box(g, GUST_LOC_ZFACE));
    set_box(st, &zero, loc_box(g, GUST_LOC_ZFACE), 0.0);
    const gust_geometry* geom = &g;
    check_rc(gust_face_divergence(geom, mesh.ax, mesh.ay, mesh.az, mesh.detJ, u, v, &zero, &dh, st));
    auto* worst = static_cast<unsigned long long*>(Workspace::get(st, kSlotWorst, sizeof(unsigned long long)));
    // a fill kernel, not a memset: the call may sit in a captured graph
    gust_array4 wv = scratch_view(st, kSlotWorst, IBox{{0, 0, 0}, {0, 0, 0}});
    set_box(st, &wv, IBox{{0, 0, 0}, {0, 0, 0}}, 0.0);
    const int nx = cells.len(0), ny = cells.len(1), nz = cells.len(2);
    obrien_kernel<<<blocks_for(int64_t(nx) * ny), kThreads, 0, st>>>(
        g.lo[0], g.lo[1], nx, ny, g.lo[2], nz, dview<const double>(&dh), dview<const double>(mesh.detJ),
        dview<const double>(mesh.az), dview<const double>(mesh.z_nd), dview<double>(omega), g.dx[2], worst);
    GUST_HIP(hipGetLastError());
    auto* h = static_cast<unsigned long long*>(Workspace::get_host_pinned(kSlotWorstHost, sizeof(unsigned long long)));
    GUST_HIP(hipMemcpyAsync(h, worst, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    GUST_HIP(hipStreamSynchronize(st));
    double r;
    const unsigned long long bits = *h;
    std::memcpy(&r, &bits, sizeof r);
    return r;
}

// first_guess_state on validated arguments.
double first_guess_impl(hipStream_t st, const gust_geometry& g, const gust_mesh& mesh, const gust_inflow_profile& prof,
                        const gust_init_options& opt, gust_state& q)
{
    if (opt.first_guess == 0) { return 0.0; }
    const ProfileView pv = profile::device_view(st, prof);
    const gust_array4* z = mesh.z_nd;
    // 1. u, v from the profile at every point; w = 0 but on the ground faces,
    // which carry no flux (Omega_0 = 0 there).
    fill_profile<sc::XF>(st, g, q.f[0], z, pv, 0, opt.z_agl_min, 0.0);
    fill_profile<sc::YF>(st, g, q.f[1], z, pv, 1, opt.z_agl_min, 0.0);
    set_box(st, q.f[2], view_box(q.f[2]), 0.0);
    const IBox zf = loc_box(g, GUST_LOC_ZFACE);
    gust_array4 om = scratch_view(st, kSlotOm, zf);
    const gust_geometry* geom = &g;
    check_rc(gust_omega_from_w(geom, q.f[0], q.f[1], q.f[2], z, &om, 0, st));
    IBox ground = zf;
    ground.hi[2] = ground.lo[2];
    set_box(st, &om, ground, 0.0);
    check_rc(gust_w_from_omega(geom, q.f[0], q.f[1], &om, z, q.f[2], st));
    // 2. k from the profile with the decay; theta the sounding at the
    // absolute height (one ghost cell, the rest 0) or the table's by height
    // above the ground; p' = 0.
    fill_profile<sc::CC>(st, g, q.f[5], z, pv, 3, opt.z_agl_min, opt.k_decay_height);
    if (opt.column_theta && prof.kind == GUST_INFLOW_TABLE) {
        fill_profile<sc::CC>(st, g, q.f[4], z, pv, 2, opt.z_agl_min, 0.0);
    } else {
        set_box(st, q.f[4], view_box(q.f[4]), 0.0);
        const heights::Sounding s = heights::sounding(prof.n_theta_ref, prof.theta_ref_z, prof.theta_ref);
        const Box3 tb = box3(grow(loc_box(g, GUST_LOC_CELL), 1, 1, 1));
        heights::theta_ref_kernel<<<blocks_for(tb.npts()), kThreads, 0, st>>>(tb, dview<const double>(z), s,
                                                                               dview<double>(q.f[4]));
        GUST_HIP(hipGetLastError());
    }
    set_box(st, q.f[3], view_box(q.f[3]), 0.0);
    // 3. O'Brien: w0 from the column continuity with the top residual removed.
    double residual = 0.0;
    if (opt.obrien) {
        residual = obrien_impl(st, g, mesh, q.f[0], q.f[1], &om);
        check_rc(gust_w_from_omega(geom, q.f[0], q.f[1], &om, z, q.f[2], st));
    }
    GUST_HIP(hipStreamSynchronize(st));
    return residual;
}

void check_common(const char* op, const gust_geometry* geom, const gust_mesh* mesh, const gust_inflow_profile* prof,
                  const gust_init_options* opt, const gust_state* q)
{
    check_geom(op, geom);
    if (!prof || !opt) { GUST_FAIL("gust: %s: null profile or options", op); }
    fail_first(op, init_problems(*opt));
    fail_first(op, profile::problems(*prof));
    check_mesh(op, *geom, mesh);
    check_state(op, *geom, q);
}

}   // namespace

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_inflow_default_profile(gust_inflow_profile* p)
{
    return guarded([&] {
        if (!p) { GUST_FAIL("gust: inflow_default_profile: null output"); }
        *p = gust_inflow_profile{};
        p->kind = GUST_INFLOW_LOG_LAW;
        p->ustar = 0.4;
        p->z0 = 0.1;
        p->speed_ref = 10.0;
        p->z_ref = 10.0;
        p->alpha = 0.14;
        p->direction = 0.0;
        p->Cmu0 = 0.5562;
        p->kappa = 0.41;
        p->unbounded = 0;
        p->l_max = 30.0;
        p->n_theta_ref = 1;
        p->theta_ref_z[0] = 0.0;
        p->theta_ref[0] = 300.0;
    });
}

int gust_check_inflow_profile(const gust_inflow_profile* p)
{
    return guarded([&] {
        if (!p) { GUST_FAIL("gust: InflowProfile: null profile"); }
        fail_first("InflowProfile", profile::problems(*p));
    });
}

int gust_inflow_eval(const gust_inflow_profile* p, int32_t n, const double* z, double* u, double* v, double* theta,
                     double* k, double* nut, void* stream)
{
    return guarded([&] {
        const char* op = "InflowProfile::eval";
        if (!p) { GUST_FAIL("gust: %s: null profile", op); }
        fail_first(op, profile::problems(*p));
        if (n < 0 || (n > 0 && !z)) { GUST_FAIL("gust: %s: n >= 0 heights needed", op); }
        if (n == 0) { return; }
        auto st = static_cast<hipStream_t>(stream);
        const ProfileView pv = profile::device_view(st, *p);
        eval_kernel<<<blocks_for(n), kThreads, 0, st>>>(pv, n, z, u, v, theta, k, nut);
        GUST_HIP(hipGetLastError());
    });
}

int gust_init_default_options(gust_init_options* o)
{
    return guarded([&] {
        if (!o) { GUST_FAIL("gust: init_default_options: null output"); }
        *o = gust_init_options{};
        o->first_guess = 0;
        o->projection = 1;
        o->z_agl_min = 0.0;
        o->alpha_h = 1.0;
        o->alpha_v = 1.0;
        o->k_decay_height = 0.0;
        o->obrien = 0;
        o->column_theta = 0;
    });
}

int gust_check_init_options(const gust_init_options* o)
{
    return guarded([&] {
        if (!o) { GUST_FAIL("gust: check_init_options: null options"); }
        fail_first("check_init_options", init_problems(*o));
    });
}

int gust_initializer_poisson_bc(const gust_geometry* geom, const int32_t* faces, int32_t* pbc)
{
    return guarded([&] {
        const char* op = "initializer_poisson_bc";
        check_geom(op, geom);
        if (!faces || !pbc) { GUST_FAIL("gust: %s: null face types", op); }
        for (int o = 0; o < 6; ++o) {
            const int d = o % 3;
            if (d < 2 && geom->periodic[d]) {
                pbc[o] = GUST_PF_PERIODIC;
            } else if (faces[o] == GUST_FACE_INFLOW || faces[o] == GUST_FACE_OUTFLOW) {
                pbc[o] = GUST_PF_DIRICHLET;
            } else {
                pbc[o] = GUST_PF_NEUMANN;
            }
        }
    });
}

int gust_obrien_omega(const gust_geometry* geom, const gust_mesh* mesh, const gust_array4* u, const gust_array4* v,
                      gust_array4* omega, double* top_residual, void* stream)
{
    return guarded([&] {
        const char* op = "obrien_omega";
        check_geom(op, geom);
        check_mesh(op, *geom, mesh);
        check_field(op, "u", u, *geom, GUST_LOC_XFACE, 0);
        check_field(op, "v", v, *geom, GUST_LOC_YFACE, 0);
        check_field(op, "omega", omega, *geom, GUST_LOC_ZFACE, 0);
        const double r = obrien_impl(static_cast<hipStream_t>(stream), *geom, *mesh, u, v, omega);
        if (top_residual) { *top_residual = r; }
    });
}

int gust_first_guess_state(const gust_geometry* geom, const gust_mesh* mesh, const gust_inflow_profile* prof,
                           const gust_init_options* opt, gust_state* q, double* obrien_residual, void* stream)
{
    return guarded([&] {
        const char* op = "first_guess_state";
        check_common(op, geom, mesh, prof, opt, q);
        const double r = first_guess_impl(static_cast<hipStream_t>(stream), *geom, *mesh, *prof, *opt, *q);
        if (obrien_residual) { *obrien_residual = r; }
    });
}

int gust_initialize_state(const gust_geometry* geom, const gust_mesh* mesh, const gust_inflow_profile* prof,
                          const gust_init_options* opt, const int32_t* faces, gust_poisson* solver, gust_state* q,
                          gust_array4* lambda, gust_init_result* result, void* stream)
{
    return guarded([&] {
        const char* op = "initialize_state";
        check_common(op, geom, mesh, prof, opt, q);
        if (!faces || !result) { GUST_FAIL("gust: %s: null face types or result", op); }
        if (opt->first_guess != 0 && opt->projection && !solver) { GUST_FAIL("gust: %s: null Poisson solver", op); }
        const gust_geometry& g = *geom;
        const IBox cells = loc_box(g, GUST_LOC_CELL);
        if (lambda) { check_field(op, "lambda", lambda, g, GUST_LOC_CELL, 1); }
        auto st = static_cast<hipStream_t>(stream);
        *result = gust_init_result{};
        result->obrien_top_residual = first_guess_impl(st, g, *mesh, *prof, *opt, *q);
        if (opt->first_guess == 0) { return; }
        const gust_array4* z = mesh->z_nd;
        gust_array4 rhs = scratch_view(st, kSlotRhs, cells);
        check_rc(gust_velocity_divergence(geom, q->f[0], q->f[1], q->f[2], z, mesh->ax, mesh->ay, mesh->az,
                                          mesh->detJ, &rhs, st));
        result->divergence_before = std::sqrt(sum_squares(st, &rhs, cells));
        if (!opt->projection) {
            // the first guess as it is: the stage's pressure correction
            // removes the divergence (reported as before = after)
            result->divergence_after = result->divergence_before;
            return;
        }
        // 4. The correction, applied to (u, v, Omega); w follows.
        gust_array4 lam_local;
        gust_array4* lam = lambda;
        if (!lam) {
            lam_local = scratch_view(st, kSlotLam, grow(cells, 1, 1, 1));
            lam = &lam_local;
        }
        set_box(st, lam, view_box(lam), 0.0);
        int32_t pbc[6];
        check_rc(gust_initializer_poisson_bc(geom, faces, pbc));
        gust_array4 F[3];
        for (int d = 0; d < 3; ++d) { F[d] = scratch_view(st, kSlotF + d, loc_box(g, 1 << d)); }
        gust_face_coeffs coeffs{};
        const gust_face_coeffs* cp = nullptr;
        gust_array4 beta[4];
        if (!(opt->alpha_h == 1 && opt->alpha_v == 1)) {
            // Sigma: x, y and the cross part of the z faces carry ah^2, the z
            // part av^2; without a split when av = ah
            const double ah2 = opt->alpha_h * opt->alpha_h, av2 = opt->alpha_v * opt->alpha_v;
            const int locs[4] = {GUST_LOC_XFACE, GUST_LOC_YFACE, GUST_LOC_ZFACE, GUST_LOC_ZFACE};
            const double vals[4] = {ah2, ah2, av2, ah2};
            for (int c = 0; c < 4; ++c) {
                beta[c] = scratch_view(st, kSlotCoef + c, loc_box(g, locs[c]));
                set_box(st, &beta[c], loc_box(g, locs[c]), vals[c]);
            }
            coeffs.x = &beta[0];
            coeffs.y = &beta[1];
            coeffs.z = &beta[2];
            coeffs.z_cross = (opt->alpha_v != opt->alpha_h) ? &beta[3] : nullptr;
            coeffs.precond_ratio = av2 / ah2;
            cp = &coeffs;
        }
        check_rc(gust_poisson_solve(solver, &rhs, lam, cp, pbc, &F[0], &F[1], &F[2], &result->poisson, st));
        const IBox zf = loc_box(g, GUST_LOC_ZFACE);
        gust_array4 om = scratch_view(st, kSlotOm, zf);
        check_rc(gust_omega_from_w(geom, q->f[0], q->f[1], q->f[2], z, &om, 0, st));
        add_box(st, q->f[0], &F[0], loc_box(g, GUST_LOC_XFACE));
        add_box(st, q->f[1], &F[1], loc_box(g, GUST_LOC_YFACE));
        add_box(st, &om, &F[2], zf);
        // u and v with their ghosts: the periodic images; outside the
        // domain the first-guess values stay
        periodic_fill(st, g, q->f[0], GUST_LOC_XFACE);
        periodic_fill(st, g, q->f[1], GUST_LOC_YFACE);
        check_rc(gust_w_from_omega(geom, q->f[0], q->f[1], &om, z, q->f[2], st));
        check_rc(gust_velocity_divergence(geom, q->f[0], q->f[1], q->f[2], z, mesh->ax, mesh->ay, mesh->az,
                                          mesh->detJ, &rhs, st));
        result->divergence_after = std::sqrt(sum_squares(st, &rhs, cells));
    });
}

}   // extern "C"
