// transfer.hip -- the transfers between levels of the horizontal grid
// hierarchy (Source/Transfer.H/.cpp): level lc is coarser than level lf by
// r = 2^(lc - lf) in x and y, z is never coarsened, so every transfer works
// per zeta-level k.
//
//   prolongation: every field at its own staggered point from the four
//     coarse values around it (bilinear per level; the line through the two
//     nearest coarse centres extended in the half cell next to a
//     non-periodic face; the periodic images across a seam; constant along
//     a non-periodic cell-centred direction one coarse cell wide);
//   restriction: cell fields the detJ-weighted mean of the r x r fine cells,
//     u and v the face flux over the coarse area, w from the restricted
//     vertical flux Omega.
//
// One launch prolongs all six fields (a job table, grid-stride over their
// points); one launch restricts the three cell fields. The expressions keep
// the reference's order of operations, so FP64 results are the CPU build's.
#include "cgrid.hpp"

#include <cmath>

namespace gust_amd {

namespace {

using cgrid::blocks_for;
using cgrid::Box3;
using cgrid::box3;
using cgrid::check_geom;
using cgrid::CV;
using cgrid::kThreads;
using cgrid::V;

// Where a fine point sits among the coarse points along one direction
// (Source/Transfer.cpp:16-82): m = 2 (i - flo) + 1 - r over den = 2r for a
// cell-centred direction, m = i - flo over den = r (floor of (m - 1) / den)
// for a nodal one; the coarse index is clamped to [imin, imax] (the linear
// extrapolation beyond), t = lin (m - den c) / den. A periodic cell-centred
// direction reads the images at c = -1 and c = N (wrap).
struct Axis {
    int flo, clo, r, imin, imax, step, n, wrap;
    double lin, inv;
};

Axis make_axis(const gust_geometry& gf, const gust_geometry& gc, int d, int r, bool nodal)
{
    Axis a{};
    a.flo = gf.lo[d];
    a.clo = gc.lo[d];
    a.r = r;
    a.n = gc.hi[d] - gc.lo[d] + 1;
    a.step = 1;
    a.lin = 1.0;
    a.wrap = 0;
    if (nodal) {
        a.imin = 0;
        a.imax = a.n - 1;
        a.inv = 1.0 / double(r);
    } else {
        a.inv = 1.0 / double(2 * r);
        if (gc.periodic[d]) {
            a.imin = -1;
            a.imax = a.n - 1;
            a.wrap = 1;
        } else if (a.n >= 2) {
            a.imin = 0;
            a.imax = a.n - 2;
        } else {
            a.imin = 0;
            a.imax = 0;
            a.step = 0;
            a.lin = 0.0;
        }
    }
    return a;
}

__device__ __forceinline__ int floor_div(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

template <bool Nodal>
__device__ __forceinline__ void locate(const Axis& a, int i, int& ic, int& ic1, double& t)
{
    const int den = Nodal ? a.r : 2 * a.r;
    const int m = Nodal ? (i - a.flo) : (2 * (i - a.flo) + 1 - a.r);
    int c = floor_div(Nodal ? m - 1 : m, den);
    c = min(max(c, a.imin), a.imax);
    t = a.lin * double(m - den * c) * a.inv;
    int c1 = c + a.step;
    if (a.wrap) {   // the coarse periodic images
        if (c < 0) { c += a.n; }
        if (c1 >= a.n) { c1 -= a.n; }
    }
    ic = c + a.clo;
    ic1 = c1 + a.clo;
}

struct ProlongJob {
    V f;
    CV q;
    Box3 b;
    int type;   // bit 0 nodal in x, bit 1 nodal in y
};

struct ProlongSet {
    ProlongJob job[6];
    int64_t off[7];
    Axis cx, cy, nx, ny;
};

template <bool NX, bool NY>
__device__ __forceinline__ void prolong_point(const ProlongSet& s, const ProlongJob& jb, int64_t t)
{
    int i, j, k;
    jb.b.at(t, i, j, k);
    int I, I1, J, J1;
    double tx, ty;
    locate<NX>(NX ? s.nx : s.cx, i, I, I1, tx);
    locate<NY>(NY ? s.ny : s.cy, j, J, J1, ty);
    const double one = 1.0;
    jb.f(i, j, k) = (one - ty) * ((one - tx) * jb.q(I, J, k) + tx * jb.q(I1, J, k)) +
                    ty * ((one - tx) * jb.q(I, J1, k) + tx * jb.q(I1, J1, k));
}

__global__ __launch_bounds__(kThreads) void prolong_kernel(ProlongSet s)
{
    const int64_t total = s.off[6];
    for (int64_t g = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; g < total;
         g += int64_t(gridDim.x) * blockDim.x) {
        int n = 0;
        while (g >= s.off[n + 1]) { ++n; }
        const ProlongJob& jb = s.job[n];
        const int64_t t = g - s.off[n];
        switch (jb.type) {
            case 1: prolong_point<true, false>(s, jb, t); break;
            case 2: prolong_point<false, true>(s, jb, t); break;
            default: prolong_point<false, false>(s, jb, t); break;
        }
    }
}

// c = sum J q / sum J over the r x r fine cells of each coarse cell
// (Source/Transfer.cpp:169-190), for theta, k and p in one launch.
struct CellJob {
    V c;
    CV q;
};

struct RestrictCells {
    CellJob job[3];
    CV J;
    Box3 b;   // the coarse cells
    int r, flo0, flo1, clo0, clo1;
};

__global__ __launch_bounds__(kThreads) void restrict_cells_kernel(RestrictCells s)
{
    const int64_t n = s.b.npts();
    for (int64_t g = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; g < 3 * n;
         g += int64_t(gridDim.x) * blockDim.x) {
        const int f = int(g / n);
        int I, Jc, k;
        s.b.at(g - int64_t(f) * n, I, Jc, k);
        const int i0 = s.flo0 + s.r * (I - s.clo0), j0 = s.flo1 + s.r * (Jc - s.clo1);
        double sjq = 0, sj = 0;
        for (int b = 0; b < s.r; ++b) {
            for (int a = 0; a < s.r; ++a) {
                const double w = s.J(i0 + a, j0 + b, k);
                sjq += w * s.job[f].q(i0 + a, j0 + b, k);
                sj += w;
            }
        }
        s.job[f].c(I, Jc, k) = sjq / sj;
    }
}

// u_c = ((1/r) sum a_f u_f) / a_c over the r fine x-faces of a coarse x-face
// (D = 0), v alike on y-faces (Source/Transfer.cpp:192-214, 276-287).
template <int D>
__global__ __launch_bounds__(kThreads) void restrict_faces_kernel(Box3 b, V c, CV q, CV af, CV ac, int r, int flo0,
                                                                  int flo1, int clo0, int clo1)
{
    const double rinv = 1.0 / double(r);
    GUST_GRID_LOOP(b)
    {
        int I, J, k;
        b.at(t, I, J, k);
        double s = 0;
        for (int bb = 0; bb < r; ++bb) {
            const int i = flo0 + r * (I - clo0) + (D == 0 ? 0 : bb);
            const int j = flo1 + r * (J - clo1) + (D == 0 ? bb : 0);
            s += af(i, j, k) * q(i, j, k);
        }
        c(I, J, k) = (s * rinv) / ac(I, J, k);
    }
}

// Omega_c = ((1/r^2) sum a_z,f Omega_f) / a_z,c over the r x r fine z-faces
// (Source/Transfer.cpp:216-236, 289-303).
__global__ __launch_bounds__(kThreads) void restrict_zfaces_kernel(Box3 b, V c, CV om, CV af, CV ac, int r, int flo0,
                                                                   int flo1, int clo0, int clo1)
{
    const double r2inv = 1.0 / double(r * r);
    GUST_GRID_LOOP(b)
    {
        int I, J, k;
        b.at(t, I, J, k);
        double s = 0;
        for (int bb = 0; bb < r; ++bb) {
            for (int a = 0; a < r; ++a) {
                const int i = flo0 + r * (I - clo0) + a, j = flo1 + r * (J - clo1) + bb;
                s += af(i, j, k) * om(i, j, k);
            }
        }
        c(I, J, k) = (s * r2inv) / ac(I, J, k);
    }
}

// A coarse face field with one ghost layer for w_from_omega: the valid
// faces, their periodic images in x and y, and the rows below the ground
// and above the top from the nearest valid row (zero gradient,
// Source/Transfer.cpp:238-260, 305-317). g: the grown box.
__global__ __launch_bounds__(kThreads) void ghosted_copy_kernel(Box3 g, V dst, CV src, int lo0, int hi0, int lo1,
                                                                int hi1, int lo2, int hi2, int per0, int per1,
                                                                int n0, int n1)
{
    GUST_GRID_LOOP(g)
    {
        int i, j, k;
        g.at(t, i, j, k);
        int si = i, sj = j;
        if (per0) {
            if (si < lo0) { si += n0; }
            if (si > hi0) { si -= n0; }
        }
        if (per1) {
            if (sj < lo1) { sj += n1; }
            if (sj > hi1) { sj -= n1; }
        }
        if (si < lo0 || si > hi0 || sj < lo1 || sj > hi1) {
            dst(i, j, k) = 0.0;   // never read by w_from_omega
            continue;
        }
        const int sk = k < lo2 ? lo2 : (k > hi2 ? hi2 : k);
        dst(i, j, k) = src(si, sj, sk);
    }
}

void check_pair(const char* op, const gust_geometry* fg, const gust_geometry* cg, int r)
{
    check_geom(op, fg);
    check_geom(op, cg);
    if (r < 2 || (r & (r - 1)) != 0 || r > (1 << (GUST_MAX_HIERARCHY_LEVELS - 1))) {
        GUST_FAIL("gust: %s: the ratio %d is not a power of two in [2, 2^%d]", op, r, GUST_MAX_HIERARCHY_LEVELS - 1);
    }
    for (int d = 0; d < 2; ++d) {
        const int nf = fg->hi[d] - fg->lo[d] + 1, nc = cg->hi[d] - cg->lo[d] + 1;
        if (nf != r * nc || fg->lo[d] != r * cg->lo[d] || fg->periodic[d] != cg->periodic[d]) {
            GUST_FAIL("gust: %s: the coarse level is not the fine one coarsened by (%d, %d, 1)", op, r, r);
        }
        if (std::fabs(cg->dx[d] - r * fg->dx[d]) > 1e-12 * cg->dx[d]) {
            GUST_FAIL("gust: %s: the coarse cell size is not %d times the fine one", op, r);
        }
    }
    if (fg->lo[2] != cg->lo[2] || fg->hi[2] != cg->hi[2] || fg->dx[2] != cg->dx[2]) {
        GUST_FAIL("gust: %s: the two levels must have the same z-levels", op);
    }
}

// Scratch of restrict_state: the fine and coarse Omega, u and v with ghosts.
constexpr int kSlotBase = 200;

void check_rc(int rc)
{
    if (rc != 0) { throw Fail{rc}; }
}

const int kLocs[6] = {GUST_LOC_XFACE, GUST_LOC_YFACE, GUST_LOC_ZFACE, GUST_LOC_CELL, GUST_LOC_CELL, GUST_LOC_CELL};
const char* kNames[6] = {"u", "v", "w", "p", "theta", "tke"};

void check_state(const char* op, const char* what, const gust_state* q, const gust_geometry& g)
{
    if (!q) { GUST_FAIL("gust: %s: null %s state", op, what); }
    char name[64];
    for (int n = 0; n < 6; ++n) {
        std::snprintf(name, sizeof name, "%s.%s", what, kNames[n]);
        check_field(op, name, q->f[n], g, kLocs[n], 0);
    }
}

}   // namespace

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_prolong_fields(const gust_geometry* fine_geom, const gust_geometry* coarse_geom, int32_t ratio,
                        const gust_state* coarse, gust_state* fine, void* stream)
{
    return guarded([&] {
        const char* op = "prolong_fields";
        check_pair(op, fine_geom, coarse_geom, ratio);
        check_state(op, "coarse", coarse, *coarse_geom);
        check_state(op, "fine", fine, *fine_geom);
        ProlongSet s{};
        s.cx = make_axis(*fine_geom, *coarse_geom, 0, ratio, false);
        s.cy = make_axis(*fine_geom, *coarse_geom, 1, ratio, false);
        s.nx = make_axis(*fine_geom, *coarse_geom, 0, ratio, true);
        s.ny = make_axis(*fine_geom, *coarse_geom, 1, ratio, true);
        s.off[0] = 0;
        for (int n = 0; n < 6; ++n) {
            const int loc = kLocs[n];
            ProlongJob& jb = s.job[n];
            jb.f = dview<double>(fine->f[n]);
            jb.q = dview<const double>(coarse->f[n]);
            jb.b = box3(loc_box(*fine_geom, loc));
            jb.type = (loc == GUST_LOC_XFACE) ? 1 : (loc == GUST_LOC_YFACE) ? 2 : 0;
            s.off[n + 1] = s.off[n] + jb.b.npts();
        }
        prolong_kernel<<<blocks_for(s.off[6]), kThreads, 0, hipStream_t(stream)>>>(s);
        GUST_HIP(hipGetLastError());
    });
}

int gust_restrict_state(const gust_geometry* fine_geom, const gust_mesh* fine_mesh, const gust_state* fine,
                        const gust_geometry* coarse_geom, const gust_mesh* coarse_mesh, int32_t ratio,
                        gust_state* coarse, gust_array4* omega, void* stream)
{
    return guarded([&] {
        const char* op = "restrict_state";
        check_pair(op, fine_geom, coarse_geom, ratio);
        check_state(op, "fine", fine, *fine_geom);
        check_state(op, "coarse", coarse, *coarse_geom);
        if (!fine_mesh || !coarse_mesh) { GUST_FAIL("gust: %s: null mesh", op); }
        const gust_geometry& fg = *fine_geom;
        const gust_geometry& cg = *coarse_geom;
        check_field(op, "fine.u", fine->f[0], fg, GUST_LOC_XFACE, 1);
        check_field(op, "fine.v", fine->f[1], fg, GUST_LOC_YFACE, 1);
        check_nodal_heights(op, fine_mesh->z_nd, fg);
        check_nodal_heights(op, coarse_mesh->z_nd, cg);
        check_field(op, "fine detJ", fine_mesh->detJ, fg, GUST_LOC_CELL, 0);
        check_field(op, "fine ax", fine_mesh->ax, fg, GUST_LOC_XFACE, 0);
        check_field(op, "fine ay", fine_mesh->ay, fg, GUST_LOC_YFACE, 0);
        check_field(op, "fine az", fine_mesh->az, fg, GUST_LOC_ZFACE, 0);
        check_field(op, "coarse ax", coarse_mesh->ax, cg, GUST_LOC_XFACE, 0);
        check_field(op, "coarse ay", coarse_mesh->ay, cg, GUST_LOC_YFACE, 0);
        check_field(op, "coarse az", coarse_mesh->az, cg, GUST_LOC_ZFACE, 0);
        if (omega) { check_field(op, "omega", omega, cg, GUST_LOC_ZFACE, 0); }
        hipStream_t st = hipStream_t(stream);
        const int r = ratio;
        const int f0 = fg.lo[0], f1 = fg.lo[1], c0 = cg.lo[0], c1 = cg.lo[1];
        {
            RestrictCells s{};
            for (int n = 0; n < 3; ++n) {
                s.job[n].c = dview<double>(coarse->f[3 + n]);
                s.job[n].q = dview<const double>(fine->f[3 + n]);
            }
            s.J = dview<const double>(fine_mesh->detJ);
            s.b = box3(loc_box(cg, GUST_LOC_CELL));
            s.r = r;
            s.flo0 = f0;
            s.flo1 = f1;
            s.clo0 = c0;
            s.clo1 = c1;
            restrict_cells_kernel<<<blocks_for(3 * s.b.npts()), kThreads, 0, st>>>(s);
            GUST_HIP(hipGetLastError());
        }
        {
            const Box3 bx = box3(loc_box(cg, GUST_LOC_XFACE));
            restrict_faces_kernel<0><<<blocks_for(bx.npts()), kThreads, 0, st>>>(
                bx, dview<double>(coarse->f[0]), dview<const double>(fine->f[0]), dview<const double>(fine_mesh->ax),
                dview<const double>(coarse_mesh->ax), r, f0, f1, c0, c1);
            GUST_HIP(hipGetLastError());
            const Box3 by = box3(loc_box(cg, GUST_LOC_YFACE));
            restrict_faces_kernel<1><<<blocks_for(by.npts()), kThreads, 0, st>>>(
                by, dview<double>(coarse->f[1]), dview<const double>(fine->f[1]), dview<const double>(fine_mesh->ay),
                dview<const double>(coarse_mesh->ay), r, f0, f1, c0, c1);
            GUST_HIP(hipGetLastError());
        }
        // w: the fine Omega, restricted; then w from it on the coarse mesh.
        const IBox zf = loc_box(fg, GUST_LOC_ZFACE), zc = loc_box(cg, GUST_LOC_ZFACE);
        const IBox gu = grow(loc_box(cg, GUST_LOC_XFACE), 1, 1, 1), gv = grow(loc_box(cg, GUST_LOC_YFACE), 1, 1, 1);
        const size_t nom_f = size_t(zf.npts()), nom_c = size_t(zc.npts());
        const size_t nu1 = size_t(gu.npts()), nv1 = size_t(gv.npts());
        auto view = [](double* p, const IBox& b) {
            gust_array4 a{};
            a.p = p;
            a.jstride = b.len(0);
            a.kstride = int64_t(b.len(0)) * b.len(1);
            a.nstride = b.npts();
            for (int d = 0; d < 3; ++d) {
                a.begin[d] = b.lo[d];
                a.end[d] = b.hi[d] + 1;
            }
            a.ncomp = 1;
            return a;
        };
        auto slot = [&](int k, size_t n) { return static_cast<double*>(Workspace::get(st, kSlotBase + k, n * 8)); };
        gust_array4 om_f = view(slot(0, nom_f), zf);
        gust_array4 om_local = view(slot(1, nom_c), zc);
        gust_array4 u1 = view(slot(2, nu1), gu);
        gust_array4 v1 = view(slot(3, nv1), gv);
        gust_array4& om_c = omega ? *omega : om_local;
        check_rc(gust_omega_from_w(&fg, fine->f[0], fine->f[1], fine->f[2], fine_mesh->z_nd, &om_f, 0, stream));
        const Box3 bz = box3(zc);
        restrict_zfaces_kernel<<<blocks_for(bz.npts()), kThreads, 0, st>>>(
            bz, dview<double>(&om_c), dview<const double>(&om_f), dview<const double>(fine_mesh->az),
            dview<const double>(coarse_mesh->az), r, f0, f1, c0, c1);
        GUST_HIP(hipGetLastError());
        const int n0 = cg.hi[0] - cg.lo[0] + 1, n1 = cg.hi[1] - cg.lo[1] + 1;
        const IBox vu = loc_box(cg, GUST_LOC_XFACE), vv = loc_box(cg, GUST_LOC_YFACE);
        // A periodic nodal direction's image of face i < lo is face i + n
        // (face lo + n is the seam's other copy), as FillBoundary fills it.
        const Box3 bgu = box3(gu), bgv = box3(gv);
        ghosted_copy_kernel<<<blocks_for(bgu.npts()), kThreads, 0, st>>>(
            bgu, dview<double>(&u1), dview<const double>(coarse->f[0]), vu.lo[0], vu.hi[0], vu.lo[1], vu.hi[1],
            vu.lo[2], vu.hi[2], cg.periodic[0], cg.periodic[1], n0, n1);
        GUST_HIP(hipGetLastError());
        ghosted_copy_kernel<<<blocks_for(bgv.npts()), kThreads, 0, st>>>(
            bgv, dview<double>(&v1), dview<const double>(coarse->f[1]), vv.lo[0], vv.hi[0], vv.lo[1], vv.hi[1],
            vv.lo[2], vv.hi[2], cg.periodic[0], cg.periodic[1], n0, n1);
        GUST_HIP(hipGetLastError());
        check_rc(gust_w_from_omega(&cg, &u1, &v1, &om_c, coarse_mesh->z_nd, coarse->f[2], stream));
    });
}

}   // extern "C"
