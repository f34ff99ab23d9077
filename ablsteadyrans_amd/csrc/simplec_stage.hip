// simplec_stage.hip -- one SIMPLEC iteration of the steady RANS equations on
// a tile and the steady driver loop around it (Source/Simplec.cpp:293-506,
// Source/Driver.cpp:325-492 with one stage). The stage owns its fields for
// its lifetime (allocated at creation, as the reference's workspace is kept
// after first use) and composes the library's entry points in the
// reference's order:
//
//   1. ghosts and the wall state at q;          2. R(q), the closure's fields, the norms;
//   3. the momentum predictor with the implicit rows at q's transport;
//   4. ghosts, the pressure correction;         5. theta's convection at the corrected
//      transport, the theta and k sweeps with the rows at the corrected transport.
#include "cgrid.hpp"
#include "deferred.hpp"
#include "simplec.hpp"
#include "stage.hpp"

#include <cfloat>
#include <cmath>
#include <exception>
#include <cstring>
#include <memory>
#include <vector>

namespace gust_amd {

namespace {

using cgrid::blocks_for;
using cgrid::Box3;
using cgrid::box3;
using cgrid::check_geom;
using cgrid::CV;
using cgrid::kThreads;
using cgrid::V;

// c1 -= c0; R += c1 (Source/Simplec.cpp:497-498).
__global__ __launch_bounds__(kThreads) void conv_update_kernel(Box3 b, V c1, CV c0, V r)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double d = c1(i, j, k) - c0(i, j, k);
        c1(i, j, k) = d;
        r(i, j, k) += d;
    }
}

// k >= floor on the valid cells, the count of raised cells
// (Source/Transfer.cpp:368-384).
__global__ __launch_bounds__(kThreads) void clip_floor_kernel(Box3 b, V k, double floor, unsigned long long* count)
{
    unsigned long long mine = 0;
    GUST_GRID_LOOP(b)
    {
        int i, j, kk;
        b.at(t, i, j, kk);
        if (k(i, j, kk) < floor) {
            k(i, j, kk) = floor;
            ++mine;
        }
    }
    if (mine) { atomicAdd(count, mine); }
}

// d = 1 - mask on one velocity's faces (the projection's denominators,
// Source/Transfer.cpp:388-399: d.mult(-1), d.plus(1)).
__global__ __launch_bounds__(kThreads) void one_minus_kernel(Box3 b, V d, CV m)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        d(i, j, k) = -1.0 * m(i, j, k) + 1.0;
    }
}

// dst = src over a box (both views cover it).
__global__ __launch_bounds__(kThreads) void copy_box_kernel(Box3 b, V dst, CV src)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        dst(i, j, k) = src(i, j, k);
    }
}

const int kLocs[6] = {GUST_LOC_XFACE, GUST_LOC_YFACE, GUST_LOC_ZFACE, GUST_LOC_CELL, GUST_LOC_CELL, GUST_LOC_CELL};

bool dirichlet_normal(int t)
{
    return t == GUST_FACE_INFLOW || t == GUST_FACE_SLIP_WALL || t == GUST_FACE_NO_SLIP_WALL ||
           t == GUST_FACE_SURFACE_LAYER || t == GUST_FACE_GEOSTROPHIC;
}

}   // namespace

namespace {

SimplecStage* as(gust_simplec* s) { return reinterpret_cast<SimplecStage*>(s); }
const SimplecStage* as(const gust_simplec* s) { return reinterpret_cast<const SimplecStage*>(s); }

void check_options(const char* op, const gust_simplec_options& o)
{
    if (std::isnan(o.pressure_reltol) || !(o.pressure_reltol > 0 && o.pressure_reltol < 1)) {
        GUST_FAIL("gust: %s: simplec.pressure_reltol must be in (0, 1)", op);
    }
    if (std::isnan(o.alpha_p) || !(o.alpha_p > 0 && o.alpha_p <= 1)) {
        GUST_FAIL("gust: %s: simplec.alpha_p must be in (0, 1]", op);
    }
    if (std::isnan(o.rotation_relax) || !(o.rotation_relax > 0 && o.rotation_relax <= 1)) {
        GUST_FAIL("gust: %s: simplec.rotation_relax must be in (0, 1]", op);
    }
    if (o.sweeps < 1) { GUST_FAIL("gust: %s: sweeps must be >= 1 (steady.zline_sweeps)", op); }
    if (o.sweep_precision != 0 && o.sweep_precision != 1) {
        GUST_FAIL("gust: %s: sweep_precision must be 0 (FP64) or 1 (FP32 sweeps)", op);
    }
    if (std::isnan(o.tke_floor) || !(o.tke_floor >= 0)) {
        GUST_FAIL("gust: %s: tke_floor must be >= 0 (turb.tke_floor)", op);
    }
    for (const double sc : {o.scales.velocity, o.scales.theta, o.scales.tke}) {
        if (std::isnan(sc) || !(sc > 0 && std::isfinite(sc))) {
            GUST_FAIL("gust: %s: the residual scales (jvp_scales) must be finite and > 0", op);
        }
    }
    if (!(o.height > 0 && std::isfinite(o.height))) { GUST_FAIL("gust: %s: the scale height must be > 0", op); }
}

double coriolis_f(const gust_forcing& f)
{
    if (!f.coriolis) { return 0; }
    const double pi = 3.14159265358979323846;
    return 4 * pi * std::sin(f.latitude * pi / 180) / f.rotation_period;
}

// One iteration (Source/Simplec.cpp:452-506).
void iterate_body(SimplecStage& s, gust_state* q, const gust_array4* dtau, gust_simplec_info* info, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    const gust_geometry* geom = &s.g;
    gust_array4* ssv = s.wall ? &s.ss.v : nullptr;
    gust_simplec_info local{};
    gust_simplec_info& inf = info ? *info : local;
    // 1-2. ghosts, the residual, the norms
    check_rc(gust_fill_ghosts(geom, q, &s.mesh, &s.phys, ssv, stream));
    gust_residuals R = s.residuals();
    check_rc(gust_residual(geom, q, &s.mesh, &s.phys, &R, s.phys.use_closure ? &s.cfv : nullptr, ssv, stream));
    check_rc(gust_stage_norms(geom, &R, q->f[0], q->f[1], q->f[2], q->f[5], ssv, &s.opt.scales, s.opt.height, s.res,
                              s.mon, &s.nmon, stream));
    // 3. the momentum predictor
    check_rc(gust_transport_fluxes(geom, q->f[0], q->f[1], q->f[2], &s.z, &s.ax, &s.ay, &s.F[0].v, &s.F[1].v,
                                   &s.F[2].v, stream));
    check_rc(gust_face_dtau(geom, dtau, &s.dtf[0].v, &s.dtf[1].v, &s.dtf[2].v, stream));
    gust_faces Ff{{&s.F[0].v, &s.F[1].v, &s.F[2].v}};
    const gust_closure_fields* cf = s.phys.use_closure ? &s.cfv : nullptr;
    const int eqs[3] = {GUST_EQ_U, GUST_EQ_V, GUST_EQ_W};
    for (int c = 0; c < 3; ++c) {
        check_rc(gust_implicit_rows(eqs[c], geom, &Ff, &s.mesh, &s.phys, cf, ssv, &s.A[c].v, stream));
    }
    const bool buoy = s.phys.forcing.gravity > 0;
    if (buoy) { check_rc(gust_implicit_rows(GUST_EQ_THETA, geom, &Ff, &s.mesh, &s.phys, cf, ssv, &s.At.v, stream)); }
    gust_faces vel{{q->f[0], q->f[1], q->f[2]}};
    gust_faces Rf{{&s.R[0].v, &s.R[1].v, &s.R[2].v}};
    gust_faces Af{{&s.A[0].v, &s.A[1].v, &s.A[2].v}};
    gust_faces Df{{&s.dtf[0].v, &s.dtf[1].v, &s.dtf[2].v}};
    gust_faces Mf{{&s.mask[0].v, &s.mask[1].v, &s.mask[2].v}};
    gust_faces Pf{{&s.ap[0].v, &s.ap[1].v, &s.ap[2].v}};
    gust_faces Qf{{&s.d[0].v, &s.d[1].v, &s.d[2].v}};
    gust_predictor_options po{};
    po.sweeps = s.opt.sweeps;
    po.rotation_relax = s.opt.rotation_relax;
    po.w_vnn_cap = s.opt.w_vnn_cap;
    po.w_vnn_slope_margin = s.opt.w_vnn_cap_slope_margin;
    po.coriolis_f = coriolis_f(s.phys.forcing);
    po.gravity = s.phys.forcing.gravity;
    po.theta0 = s.phys.forcing.theta0;
    po.sweep = s.opt.sweep;
    po.sweep_precision = s.opt.sweep_precision;
    check_rc(gust_momentum_predictor(geom, &vel, q->f[4], &s.z, &Rf, &s.R[4].v, &Af, buoy ? &s.At.v : nullptr, &Df,
                                     &Mf, &po, &Pf, &Qf, &inf.predictor, stream));
    // 4. the pressure correction on u* (its ghosts refilled; the wall state
    // of the iteration stays that of q)
    check_rc(gust_fill_ghosts(geom, q, &s.mesh, &s.phys, nullptr, stream));
    check_rc(gust_pressure_correction(s.solver, geom, q->f[0], q->f[1], q->f[2], q->f[3], &s.z, &s.ax, &s.ay, &s.az,
                                      &s.J, &s.d[0].v, &s.d[1].v, &s.d[2].v, s.pbc, s.opt.alpha_p, nullptr, 0,
                                      &inf.pressure, stream));
    // 5. theta's convection at the corrected transport, the scalar sweeps
    check_rc(gust_convection(GUST_LOC_CELL, geom, q->f[4], &s.F[0].v, &s.F[1].v, &s.F[2].v, &s.J,
                             s.phys.theta_limiter, s.phys.theta_eps, 0, &s.conv0.v, stream));
    check_rc(gust_fill_ghosts(geom, q, &s.mesh, &s.phys, nullptr, stream));
    check_rc(gust_transport_fluxes(geom, q->f[0], q->f[1], q->f[2], &s.z, &s.ax, &s.ay, &s.F[0].v, &s.F[1].v,
                                   &s.F[2].v, stream));
    check_rc(gust_convection(GUST_LOC_CELL, geom, q->f[4], &s.F[0].v, &s.F[1].v, &s.F[2].v, &s.J,
                             s.phys.theta_limiter, s.phys.theta_eps, 0, &s.conv1.v, stream));
    {
        const Box3 b = box3(loc_box(s.g, GUST_LOC_CELL));
        conv_update_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, dview<double>(&s.conv1.v),
                                                                      dview<const double>(&s.conv0.v),
                                                                      dview<double>(&s.R[4].v));
        GUST_HIP(hipGetLastError());
    }
    // k's rows do not depend on theta, so both sets of rows come first and
    // the two scalar sweeps share their line-solve launches
    check_rc(gust_implicit_rows(GUST_EQ_THETA, geom, &Ff, &s.mesh, &s.phys, cf, ssv, &s.At.v, stream));
    check_rc(gust_implicit_rows(GUST_EQ_TKE, geom, &Ff, &s.mesh, &s.phys, cf, ssv, &s.Ak.v, stream));
    scalar_sweeps_theta_tke(*geom, q->f[4], &s.R[4].v, &s.At.v, q->f[5], &s.R[5].v, &s.Ak.v, dtau, s.opt.sweeps,
                            s.opt.tke_floor, s.opt.sweep, &inf.theta, &inf.tke, st, s.opt.sweep_precision);
}

// One iteration with its checks and reported values read back together at
// its end (one copy, one wait); when a step fails, a check that failed before
// it is the error reported, as when each check is read at once.
void iterate(SimplecStage& s, gust_state* q, const gust_array4* dtau, gust_simplec_info* info, void* stream)
{
    gust_simplec_info local{};
    gust_simplec_info& inf = info ? *info : local;
    Deferred dq(static_cast<hipStream_t>(stream));
    Deferred::Scope scope(dq);
    try {
        iterate_body(s, q, dtau, &inf, stream);
    } catch (...) {
        const std::exception_ptr e = std::current_exception();
        dq.resolve();
        std::rethrow_exception(e);
    }
    dq.resolve();
}

// One iteration recorded into the stage's graph (not run): the stream is
// captured, the iteration's launches and its checks' copy become the graph.
// The caller has run an iteration on the same stream, state and dtau first,
// so every buffer and table the iteration uses exists (nothing allocates or
// uploads during the capture), and the pressure solver has fixed steps.
void capture(SimplecStage& s, gust_state* q, const gust_array4* dtau, hipStream_t st)
{
    s.drop_graph();
    if (st == nullptr) {
        GUST_FAIL("gust: SimplecStage::capture: the null stream cannot be captured; use a stream of your own");
    }
    if (s.opt.sweep.kind != GUST_SWEEP_ZJACOBI) {
        GUST_FAIL("gust: SimplecStage::capture: a captured iteration needs z-line Jacobi sweeps (steady.sweep = "
                  "zjacobi); the LU-SGS sweeps upload their plan at every call");
    }
    Deferred dq(st);
    Deferred::Scope scope(dq);
    dq.begin_recording();
    GUST_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    try {
        iterate_body(s, q, dtau, &s.ginfo, st);
        s.rec = dq.end_recording();
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g) { (void)hipGraphDestroy(g); }
        (void)dq.end_recording();
        throw;
    }
    GUST_HIP(hipStreamEndCapture(st, &s.graph));
    GUST_HIP(hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0));
}

// The captured iteration once: launch, wait, the checks' callbacks in order.
void replay(SimplecStage& s, gust_simplec_info* info, hipStream_t st)
{
    if (!s.exec) { GUST_FAIL("gust: SimplecStage: no captured iteration (capture first)"); }
    GUST_HIP(hipGraphLaunch(s.exec, st));
    GUST_HIP(hipStreamSynchronize(st));
    s.rec.replay();
    if (info) { *info = s.ginfo; }
}

}   // namespace

void check_rc(int rc)
{
    if (rc != 0) { throw Fail{rc}; }
}

void simplec_iterate(SimplecStage& s, gust_state* q, const gust_array4* dtau, gust_simplec_info* info, void* stream)
{
    iterate(s, q, dtau, info, stream);
}

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_simplec_default_options(gust_simplec_options* o)
{
    return guarded([&] {
        if (!o) { GUST_FAIL("gust: SimplecOptions: null output"); }
        std::memset(o, 0, sizeof *o);
        o->pressure_reltol = 1.0e-3;
        o->alpha_p = 1;
        o->rotation_relax = 0.3;
        o->w_vnn_cap = 0;
        o->w_vnn_cap_slope_margin = 4;
        o->sweeps = 2;
        o->sweep = gust_line_options{GUST_SWEEP_ZJACOBI, GUST_GHOSTS_LAGGED, 0, GUST_ORDER_SEQUENTIAL, 0};
        o->tke_floor = 0;
        o->scales = gust_jvp_scales{10, 10, 0.1, 100, std::sqrt(DBL_EPSILON)};
        o->height = 1;
        check_rc(gust_poisson_default_options(&o->poisson));
    });
}

int gust_simplec_create(const gust_geometry* geom, const gust_mesh* mesh, const gust_physics* phys,
                        const gust_simplec_options* opt, gust_simplec** out, void* stream)
{
    return guarded([&] {
        const char* op = "SimplecStage";
        check_geom(op, geom);
        if (!mesh || !phys || !opt || !out) { GUST_FAIL("gust: %s: null argument", op); }
        *out = nullptr;
        check_options(op, *opt);
        if (!phys->use_bc) { GUST_FAIL("gust: %s: the residual inputs need the boundary conditions (bc)", op); }
        if (!phys->use_closure && !phys->nu) {
            GUST_FAIL("gust: %s: without a closure the residual inputs need the diffusivities", op);
        }
        check_nodal_heights(op, mesh->z_nd, *geom);
        check_field(op, "detJ", mesh->detJ, *geom, GUST_LOC_CELL, 1);
        check_field(op, "ax", mesh->ax, *geom, GUST_LOC_XFACE, 1);
        check_field(op, "ay", mesh->ay, *geom, GUST_LOC_YFACE, 1);
        check_field(op, "az", mesh->az, *geom, GUST_LOC_ZFACE, 1);
        auto s = std::make_unique<SimplecStage>();
        s->g = *geom;
        s->z = *mesh->z_nd;
        s->J = *mesh->detJ;
        s->ax = *mesh->ax;
        s->ay = *mesh->ay;
        s->az = *mesh->az;
        s->mesh = gust_mesh{&s->z, &s->J, &s->ax, &s->ay, &s->az};
        s->phys = *phys;
        s->opt = *opt;
        s->wall = phys->face[2] == GUST_FACE_SURFACE_LAYER;
        const gust_geometry& g = *geom;
        const IBox cells = loc_box(g, GUST_LOC_CELL);
        for (int n = 0; n < 6; ++n) { s->R[n].alloc(loc_box(g, kLocs[n]), 1, hipStream_t(stream)); }
        for (int n = 0; n < 3; ++n) { s->cf[n].alloc(grow(cells, 1, 1, 1), 1, hipStream_t(stream)); }
        for (int n = 3; n < 7; ++n) { s->cf[n].alloc(cells, 1, hipStream_t(stream)); }
        s->cfv = gust_closure_fields{&s->cf[0].v, &s->cf[1].v, &s->cf[2].v, &s->cf[3].v,
                                     &s->cf[4].v, &s->cf[5].v, &s->cf[6].v};
        IBox slab = grow(cells, 1, 1, 0);
        slab.hi[2] = slab.lo[2];
        s->ss.alloc(slab, GUST_SS_N, hipStream_t(stream));
        for (int c = 0; c < 3; ++c) {
            const IBox fb = loc_box(g, 1 << c);
            s->F[c].alloc(grow(fb, 1, 1, 1), 1, hipStream_t(stream));
            s->dtf[c].alloc(fb, 1, hipStream_t(stream));
            s->A[c].alloc(fb, GUST_N_COEFS, hipStream_t(stream));
            s->mask[c].alloc(fb, 1, hipStream_t(stream));
            s->ap[c].alloc(fb, 1, hipStream_t(stream));
            s->d[c].alloc(fb, 1, hipStream_t(stream));
        }
        s->At.alloc(cells, GUST_N_COEFS, hipStream_t(stream));
        s->Ak.alloc(cells, GUST_N_COEFS, hipStream_t(stream));
        s->conv0.alloc(cells, 1, hipStream_t(stream));
        s->conv1.alloc(cells, 1, hipStream_t(stream));
        // The Dirichlet rows (BoundaryConditions::dirichlet_rows,
        // Source/BoundaryConditions.cpp:852-876): each velocity's face points
        // on a face of its own direction with a Dirichlet normal value
        // (inflow, the walls, the ground, the top).
        for (int c = 0; c < 3; ++c) {
            if (c < 2 && g.periodic[c]) { continue; }
            const IBox fb = loc_box(g, 1 << c);
            std::vector<double> m(size_t(fb.npts()), 0.0);
            bool any = false;
            for (int side = 0; side < 2; ++side) {
                if (!dirichlet_normal(phys->face[side == 0 ? c : c + 3])) { continue; }
                any = true;
                const int edge = side == 0 ? fb.lo[c] : fb.hi[c];
                for (int k = fb.lo[2]; k <= fb.hi[2]; ++k)
                    for (int j = fb.lo[1]; j <= fb.hi[1]; ++j)
                        for (int i = fb.lo[0]; i <= fb.hi[0]; ++i) {
                            const int pc = c == 0 ? i : c == 1 ? j : k;
                            if (pc != edge) { continue; }
                            m[size_t(i - fb.lo[0]) + size_t(fb.len(0)) * (size_t(j - fb.lo[1]) +
                                                                          size_t(fb.len(1)) * size_t(k - fb.lo[2]))] = 1.0;
                        }
            }
            if (any) {
                GUST_HIP(hipMemcpyAsync(s->mask[c].p, m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice,
                                        hipStream_t(stream)));
                GUST_HIP(hipStreamSynchronize(hipStream_t(stream)));   // m is a local
            }
        }
        // The Poisson controls with the correction's tolerance; the types of
        // the faces as ERF's get_fft_bc (Source/BoundaryConditions.cpp:300-320).
        gust_poisson_options po = opt->poisson;
        po.reltol = opt->pressure_reltol;
        check_rc(gust_poisson_create(geom, mesh->z_nd, mesh->detJ, mesh->ax, mesh->ay, mesh->az, &po, &s->solver,
                                     stream));
        for (int f = 0; f < 6; ++f) {
            const int d = f % 3;
            if (d < 2 && geom->periodic[d]) { s->pbc[f] = GUST_PF_PERIODIC; }
            else if (phys->face[f] == GUST_FACE_OUTFLOW) { s->pbc[f] = GUST_PF_DIRICHLET; }
            else { s->pbc[f] = GUST_PF_NEUMANN; }
        }
        *out = reinterpret_cast<gust_simplec*>(s.release());
    });
}

int gust_simplec_destroy(gust_simplec* s)
{
    delete as(s);
    return 0;
}

int gust_simplec_init(gust_simplec* s, gust_state* q, void* stream)
{
    return guarded([&] {
        if (!s) { GUST_FAIL("gust: SimplecStage::init: null stage"); }
        SimplecStage& st = *as(s);
        check_rc(gust_fill_ghosts(&st.g, q, &st.mesh, &st.phys, st.wall ? &st.ss.v : nullptr, stream));
    });
}

int gust_simplec_reset(gust_simplec* s, void* stream)
{
    return guarded([&] {
        if (!s) { GUST_FAIL("gust: SimplecStage::reset: null stage"); }
        SimplecStage& st = *as(s);
        hipStream_t h = static_cast<hipStream_t>(stream);
        for (DevField& f : st.cf) {
            if (f.p) { GUST_HIP(hipMemsetAsync(f.p, 0, size_t(f.v.nstride) * f.v.ncomp * sizeof(double), h)); }
        }
        GUST_HIP(hipStreamSynchronize(h));
    });
}

int gust_simplec_iterate(gust_simplec* s, gust_state* q, const gust_array4* dtau, gust_simplec_info* info,
                         void* stream)
{
    return guarded([&] {
        if (!s) { GUST_FAIL("gust: SimplecStage::iterate: null stage"); }
        SimplecStage& st = *as(s);
        check_field("SimplecStage::iterate", "dtau", dtau, st.g, GUST_LOC_CELL, 0);
        iterate(st, q, dtau, info, stream);
    });
}

int gust_simplec_norms(const gust_simplec* s, double res[6], double mon[3], int32_t* nmon)
{
    return guarded([&] {
        if (!s || !res || !mon || !nmon) { GUST_FAIL("gust: SimplecStage::residual_norms: null argument"); }
        const SimplecStage& st = *as(s);
        for (int n = 0; n < 6; ++n) { res[n] = st.res[n]; }
        for (int n = 0; n < 3; ++n) { mon[n] = st.mon[n]; }
        *nmon = st.nmon;
    });
}

int gust_simplec_nu(const gust_simplec* s, gust_array4* nu)
{
    return guarded([&] {
        if (!s || !nu) { GUST_FAIL("gust: SimplecStage::nu: null argument"); }
        const SimplecStage& st = *as(s);
        *nu = st.phys.use_closure ? st.cf[0].v : *st.phys.nu;
    });
}

// The steady loop (Source/Driver.cpp:403-468) with a level's seeding and
// start (Source/Driver.cpp:275-313) and, on a coarse level of grid
// sequencing, the loose stop after each step (Source/Sequencing.cpp:361-390).
static void solve_loop(const char* op, gust_simplec* s, gust_state* q, const gust_driver_options* opt,
                       const gust_level_control* lc, gust_level_result* out, void* stream)
{
    if (!s || !q || !opt || !out) { GUST_FAIL("gust: %s: null argument", op); }
    SimplecStage& st = *as(s);
    if (opt->max_iterations < 1) { GUST_FAIL("gust: %s: steady.max_iterations must be >= 1", op); }
    if (!(opt->cfl0 > 0 && std::isfinite(opt->cfl0)) || !(opt->cfl_max >= opt->cfl0 && std::isfinite(opt->cfl_max))) {
        GUST_FAIL("gust: %s: steady.cfl0 must be > 0 and steady.cfl_max >= steady.cfl0", op);
    }
    const char* names[6] = {"u", "v", "w", "p", "theta", "tke"};
    const char* mnames[3] = {"max_speed", "mean_tke", "mean_ustar"};
    const double mscales[3] = {st.opt.scales.velocity, st.opt.scales.tke, 0.1};
    const int nmon = st.wall ? 3 : 2;
    gust_steady* crit = nullptr;
    check_rc(gust_steady_create(&opt->criterion, 6, names, nmon, mnames, mscales, &crit));
    struct Holder {
        gust_steady* c;
        ~Holder() { gust_steady_destroy(c); }
    } hold_c{crit};
    int hold = opt->criterion.residual_ref_iterations;
    double cfl = opt->cfl0, r = 0;
    if (lc) {
        if (lc->seeded) { check_rc(gust_steady_seed_reference(crit, lc->seed_ref, 6)); }
        if (lc->release_hold) {
            bool seeded = false;
            for (int n = 0; n < 6 && lc->seeded; ++n) { seeded = seeded || lc->seed_ref[n] > 0; }
            if (!seeded) {
                GUST_FAIL("gust: Driver::release_reference_hold: the reference is not seeded (seed_reference first)");
            }
            hold = 0;
        }
        if (lc->start_cfl != 0) {
            if (hold > 0) { GUST_FAIL("gust: Driver::start_cfl: the reference hold is on"); }
            if (!std::isfinite(lc->start_cfl) || !(lc->start_cfl > 0) || lc->start_cfl > opt->cfl_max) {
                GUST_FAIL("gust: Driver::start_cfl: %.3e is not finite, > 0 and <= steady.cfl_max = %.3e",
                          lc->start_cfl, opt->cfl_max);
            }
            cfl = std::max(lc->start_cfl, opt->cfl0);
        }
    }
    const gust_sequencing_options* loose = lc ? lc->loose : nullptr;
    DevField dt;
    dt.alloc(loc_box(st.g, GUST_LOC_CELL), 1, hipStream_t(stream));
    gust_array4 nu = st.phys.use_closure ? st.cf[0].v : *st.phys.nu;   // zero before the first residual
    gust_steady_result v{};
    int it = 0;
    out->stop = GUST_LEVEL_MAX_ITERATIONS;
    out->reason[0] = '\0';
    check_rc(gust_simplec_init(s, q, stream));
    while (it < opt->max_iterations) {
        check_rc(gust_local_dtau(&st.g, q->f[0], q->f[1], q->f[2], &st.z, &st.J, &nu, cfl, &dt.v, stream));
        iterate(st, q, &dt.v, nullptr, stream);
        ++it;
        check_rc(gust_steady_add(crit, st.res, 6, st.mon, nmon));
        check_rc(gust_steady_relative_residual(crit, st.res, 6, hold, r, &r));
        check_rc(gust_steady_verdict(crit, &v, nullptr, 0));
        if (v.status == GUST_STEADY_STEADY) {
            out->stop = GUST_LEVEL_STEADY;
            break;
        }
        if (v.status == GUST_STEADY_DIVERGED) {
            out->stop = GUST_LEVEL_DIVERGED;
            break;
        }
        check_rc(gust_steady_next_cfl(crit, hold, cfl, r, opt->cfl0, opt->cfl_max, opt->ser_exponent,
                                      opt->ser_max_growth, &cfl));
        if (loose) {
            int32_t stop = 0;
            check_rc(gust_loose_stop(crit, loose, hold, &stop, out->reason, int64_t(sizeof out->reason)));
            if (stop) {
                out->stop = GUST_LEVEL_LOOSE;
                break;
            }
        }
    }
    gust_driver_result& res = out->driver;
    res.iterations = it;
    res.status = v.status;
    res.cfl = cfl;
    res.relative_residual = r;
    for (int n = 0; n < 6; ++n) { res.res[n] = st.res[n]; }
    for (int n = 0; n < 3; ++n) { res.mon[n] = st.mon[n]; }
    res.nmon = nmon;
    out->hold = hold;
    check_rc(gust_steady_references(crit, out->res_ref, 6));
}

int gust_simplec_set_pressure_steps(gust_simplec* s, int32_t steps)
{
    return guarded([&] {
        if (!s) { GUST_FAIL("gust: SimplecStage: null stage"); }
        SimplecStage& st = *as(s);
        check_rc(gust_poisson_set_fixed_steps(st.solver, steps));
        st.drop_graph();
    });
}

int gust_simplec_capture(gust_simplec* s, gust_state* q, const gust_array4* dtau, void* stream)
{
    return guarded([&] {
        if (!s || !q || !dtau) { GUST_FAIL("gust: SimplecStage::capture: null argument"); }
        SimplecStage& st = *as(s);
        check_field("SimplecStage::capture", "dtau", dtau, st.g, GUST_LOC_CELL, 0);
        capture(st, q, dtau, hipStream_t(stream));
    });
}

int gust_simplec_replay(gust_simplec* s, gust_simplec_info* info, void* stream)
{
    return guarded([&] {
        if (!s) { GUST_FAIL("gust: SimplecStage::replay: null stage"); }
        replay(*as(s), info, hipStream_t(stream));
    });
}

int gust_steady_solve_graphed(gust_simplec* s, gust_state* q, const gust_driver_options* opt, int32_t pressure_steps,
                              gust_driver_result* result, void* stream)
{
    return guarded([&] {
        const char* op = "Driver::run";
        if (!s || !q || !opt || !result) { GUST_FAIL("gust: %s: null argument", op); }
        if (pressure_steps < 0 || pressure_steps > 200) {
            GUST_FAIL("gust: %s: pressure_steps must be 0 (from the first iteration) or 1..200", op);
        }
        SimplecStage& st = *as(s);
        hipStream_t hs = hipStream_t(stream);
        // The null stream cannot be captured: the loop then runs on a stream
        // of its own (default priority), after the null stream's work, and
        // the null stream waits for it at the end.
        struct OwnStream {
            hipStream_t s = nullptr;
            ~OwnStream()
            {
                if (s) {
                    (void)hipStreamSynchronize(s);
                    (void)gust_release_stream(s);
                    (void)hipStreamDestroy(s);
                }
            }
        } own;
        if (hs == nullptr) {
            GUST_HIP(hipStreamSynchronize(nullptr));
            GUST_HIP(hipStreamCreateWithFlags(&own.s, hipStreamNonBlocking));
            hs = own.s;
            stream = own.s;
        }
        if (opt->max_iterations < 1) { GUST_FAIL("gust: %s: steady.max_iterations must be >= 1", op); }
        if (!(opt->cfl0 > 0 && std::isfinite(opt->cfl0)) || !(opt->cfl_max >= opt->cfl0 && std::isfinite(opt->cfl_max))) {
            GUST_FAIL("gust: %s: steady.cfl0 must be > 0 and steady.cfl_max >= steady.cfl0", op);
        }
        check_rc(gust_simplec_set_pressure_steps(s, 0));
        struct Restore {
            gust_simplec* s;
            gust_poisson* p;
            ~Restore()
            {
                gust_simplec_set_pressure_steps(s, 0);
                gust_poisson_fixed_steps_misses(p, 0, nullptr);
            }
        } restore{s, st.solver};
        int64_t misses = 0;
        check_rc(gust_poisson_fixed_steps_misses(st.solver, 1, &misses));
        int steps = pressure_steps;
        const char* names[6] = {"u", "v", "w", "p", "theta", "tke"};
        const char* mnames[3] = {"max_speed", "mean_tke", "mean_ustar"};
        const double mscales[3] = {st.opt.scales.velocity, st.opt.scales.tke, 0.1};
        const int nmon = st.wall ? 3 : 2;
        gust_steady* crit = nullptr;
        check_rc(gust_steady_create(&opt->criterion, 6, names, nmon, mnames, mscales, &crit));
        struct Holder {
            gust_steady* c;
            ~Holder() { gust_steady_destroy(c); }
        } hold_c{crit};
        const int hold = opt->criterion.residual_ref_iterations;
        DevField dt;
        dt.alloc(loc_box(st.g, GUST_LOC_CELL), 1, hs);
        gust_array4 nu = st.phys.use_closure ? st.cf[0].v : *st.phys.nu;
        double cfl = opt->cfl0, r = 0;
        gust_steady_result v{};
        int it = 0;
        check_rc(gust_simplec_init(s, q, stream));
        while (it < opt->max_iterations) {
            check_rc(gust_local_dtau(&st.g, q->f[0], q->f[1], q->f[2], &st.z, &st.J, &nu, cfl, &dt.v, stream));
            if (steps == 0) {
                // GMRES to the tolerance, its count sets the fixed steps
                gust_simplec_info inf{};
                iterate(st, q, &dt.v, &inf, stream);
                steps = std::min(200, std::max(1, inf.pressure.poisson.iterations + 1));
                check_rc(gust_simplec_set_pressure_steps(s, steps));
            } else if (!st.exec) {
                // with the steps fixed: the warm-up (buffers, tables, the
                // solver's work space at this step count), then the capture
                check_rc(gust_simplec_set_pressure_steps(s, steps));
                iterate(st, q, &dt.v, nullptr, stream);
                capture(st, q, &dt.v, hs);
            } else {
                replay(st, nullptr, hs);
                int64_t now = 0;
                check_rc(gust_poisson_fixed_steps_misses(st.solver, 1, &now));
                if (now != misses) {
                    misses = now;
                    steps = std::min(200, steps + 2);
                    check_rc(gust_simplec_set_pressure_steps(s, steps));   // drops the graph
                }
            }
            ++it;
            check_rc(gust_steady_add(crit, st.res, 6, st.mon, nmon));
            check_rc(gust_steady_relative_residual(crit, st.res, 6, hold, r, &r));
            check_rc(gust_steady_verdict(crit, &v, nullptr, 0));
            if (v.status != GUST_STEADY_NOT_STEADY) { break; }
            check_rc(gust_steady_next_cfl(crit, hold, cfl, r, opt->cfl0, opt->cfl_max, opt->ser_exponent,
                                          opt->ser_max_growth, &cfl));
        }
        result->iterations = it;
        result->status = v.status;
        result->cfl = cfl;
        result->relative_residual = r;
        for (int n = 0; n < 6; ++n) { result->res[n] = st.res[n]; }
        for (int n = 0; n < 3; ++n) { result->mon[n] = st.mon[n]; }
        result->nmon = nmon;
    });
}

int gust_steady_solve(gust_simplec* s, gust_state* q, const gust_driver_options* opt, gust_driver_result* result,
                      void* stream)
{
    return guarded([&] {
        if (!result) { GUST_FAIL("gust: Driver::run: null argument"); }
        gust_level_result out{};
        solve_loop("Driver::run", s, q, opt, nullptr, &out, stream);
        *result = out.driver;
    });
}

int gust_steady_solve_level(gust_simplec* s, gust_state* q, const gust_driver_options* opt,
                            const gust_level_control* lc, gust_level_result* result, void* stream)
{
    return guarded([&] {
        if (!lc) { GUST_FAIL("gust: grid sequencing: null level control"); }
        solve_loop("Driver::run", s, q, opt, lc, result, stream);
    });
}

int gust_prolong_state(gust_simplec* fine_stage, gust_poisson* solver, const gust_geometry* coarse_geom,
                       int32_t ratio, const gust_state* coarse, gust_state* fine, double tke_floor,
                       gust_prolong_result* result, void* stream)
{
    return guarded([&] {
        const char* op = "prolong_state";
        if (!fine_stage || !solver || !coarse_geom || !coarse || !fine || !result) {
            GUST_FAIL("gust: %s: null argument", op);
        }
        if (!std::isfinite(tke_floor) || !(tke_floor >= 0)) {
            GUST_FAIL("gust: prolong_state: tke_floor must be finite and >= 0 (turb.tke_floor)");
        }
        SimplecStage& st = *as(fine_stage);
        hipStream_t hs = hipStream_t(stream);
        std::memset(result, 0, sizeof *result);
        check_rc(gust_prolong_fields(&st.g, coarse_geom, ratio, coarse, fine, stream));
        {
            const Box3 b = box3(loc_box(st.g, GUST_LOC_CELL));
            auto* cnt = static_cast<unsigned long long*>(Workspace::get(hs, 204, sizeof(unsigned long long)));
            GUST_HIP(hipMemsetAsync(cnt, 0, sizeof *cnt, hs));
            clip_floor_kernel<<<blocks_for(b.npts()), kThreads, 0, hs>>>(b, dview<double>(fine->f[5]), tke_floor, cnt);
            GUST_HIP(hipGetLastError());
            unsigned long long h = 0;
            GUST_HIP(hipMemcpyAsync(&h, cnt, sizeof h, hipMemcpyDeviceToHost, hs));
            GUST_HIP(hipStreamSynchronize(hs));
            result->tke_clipped = int64_t(h);
        }
        check_rc(gust_fill_ghosts(&st.g, fine, &st.mesh, &st.phys, st.wall ? &st.ss.v : nullptr, stream));
        // The projection: the correction with d = 1 off the Dirichlet rows;
        // the stage's d fields are scratch until its next predictor.
        for (int c = 0; c < 3; ++c) {
            const Box3 b = box3(loc_box(st.g, 1 << c));
            one_minus_kernel<<<blocks_for(b.npts()), kThreads, 0, hs>>>(b, dview<double>(&st.d[c].v),
                                                                          dview<const double>(&st.mask[c].v));
            GUST_HIP(hipGetLastError());
        }
        const gust_array4& pv = *fine->f[3];
        IBox pb;
        for (int d = 0; d < 3; ++d) {
            pb.lo[d] = pv.begin[d];
            pb.hi[d] = pv.end[d] - 1;
        }
        DevField keep;
        keep.alloc(pb, 1, hs);
        const Box3 b = box3(pb);
        copy_box_kernel<<<blocks_for(b.npts()), kThreads, 0, hs>>>(b, dview<double>(&keep.v), dview<const double>(&pv));
        GUST_HIP(hipGetLastError());
        check_rc(gust_pressure_correction(solver, &st.g, fine->f[0], fine->f[1], fine->f[2], fine->f[3], &st.z, &st.ax,
                                          &st.ay, &st.az, &st.J, &st.d[0].v, &st.d[1].v, &st.d[2].v, st.pbc, 1.0,
                                          nullptr, 0, &result->projection, stream));
        copy_box_kernel<<<blocks_for(b.npts()), kThreads, 0, hs>>>(b, dview<double>(fine->f[3]),
                                                                    dview<const double>(&keep.v));
        GUST_HIP(hipGetLastError());
        GUST_HIP(hipStreamSynchronize(hs));   // keep is freed on return
    });
}

}   // extern "C"
