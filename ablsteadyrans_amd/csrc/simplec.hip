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
#include "deferred.hpp"
#include "simplec.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace gust_amd {

namespace {

using namespace cgrid;

__device__ __forceinline__ double sc_max(double a, double b) { return (a < b) ? b : a; }   // std::max
__device__ __forceinline__ double sc_min(double a, double b) { return (b < a) ? b : a; }   // std::min

// ---------------------------------------------------------------------------
// Box kernels.

// beta = 1/d and keep of the three face directions in one launch
// (blockIdx.y: the direction; beta_keep_kernel's arithmetic).
struct FaceSet3 {
    Box3 b[3];
    CV d[3];
    V beta[3], keep[3];
};

__global__ __launch_bounds__(kThreads) void beta_keep3_kernel(FaceSet3 s)
{
    const int q = blockIdx.y;
    const Box3& b = s.b[q];
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double dv = s.d[q](i, j, k);
        const bool on = dv > 0;
        s.beta[q](i, j, k) = on ? 1.0 / dv : 1.0;
        s.keep[q](i, j, k) = on ? 1.0 : 0.0;
    }
}

// The correction of the three face fields in one launch: F *= keep, then
// y += F (u, v and Omega; the arithmetic of the two separate passes).
struct CorrSet3 {
    Box3 b[3];
    V F[3], y[3];
    CV keep[3];
};

__global__ __launch_bounds__(kThreads) void correct3_kernel(CorrSet3 s)
{
    const int q = blockIdx.y;
    const Box3& b = s.b[q];
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double f = s.F[q](i, j, k) * s.keep[q](i, j, k);
        s.F[q](i, j, k) = f;
        s.y[q](i, j, k) = s.y[q](i, j, k) + f;
    }
}

// Several reductions in one launch (blockIdx.y: the job), each into its
// slot: max |a|, max |a m|, min a (as ordered bits of |a|; a > 0 expected),
// or a flag (1) where a is negative, NaN or infinite.
enum { kMaxAbs = 0, kMaxAbsMasked = 1, kMinPos = 2, kNonNeg = 3 };
constexpr int kJobs = 8;
struct RJob {
    Box3 b;
    CV a, m;
    int op;
    unsigned long long* out;
};
struct RJobs {
    RJob j[kJobs];
};

__global__ __launch_bounds__(kThreads) void reduce_jobs_kernel(RJobs s)
{
    const RJob& J = s.j[blockIdx.y];
    const Box3& b = J.b;
    const int op = J.op;
    unsigned long long best = op == kMinPos ? ~0ull : 0ull;
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double x = J.a(i, j, k);
        if (op == kNonNeg) {
            if (!(x >= 0) || !(x < INFINITY)) { best = 1; }
        } else {
            const unsigned long long bits = dev::abs_bits(op == kMaxAbsMasked ? x * J.m(i, j, k) : x);
            if (op == kMinPos) { best = bits < best ? bits : best; } else { best = bits > best ? bits : best; }
        }
    }
    if (op == kMinPos) {
        best = dev::block_min_u64<kThreads>(best);
        if (threadIdx.x == 0) { atomicMin(J.out, best); }
    } else {
        best = dev::block_max_u64<kThreads>(best);
        if (threadIdx.x == 0 && best) { atomicMax(J.out, best); }
    }
}

// The same for up to three fields in one launch (blockIdx.y: the field).
struct AddMaxSet {
    Box3 b[3];
    V y[3];
    CV x[3];
    unsigned long long* out[3];
};

__global__ __launch_bounds__(kThreads) void add_max3_kernel(AddMaxSet s)
{
    const int q = blockIdx.y;
    const Box3& b = s.b[q];
    unsigned long long best = 0;
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double v = s.x[q](i, j, k);
        s.y[q](i, j, k) = s.y[q](i, j, k) + v;
        const unsigned long long bits = dev::abs_bits(v);
        best = bits > best ? bits : best;
    }
    best = dev::block_max_u64<kThreads>(best);
    if (threadIdx.x == 0 && best) { atomicMax(s.out[q], best); }
}

// The mixed-precision sweeps' operands: A's six neighbour components, a_P
// and rhs of up to three systems rounded to FP32 compact arrays
// (blockIdx.y: the system).
struct ToF32Set {
    Box3 b[3];
    CV A[3], ap[3], r[3];
    int64_t an[3];
    float* A32[3];
    float* ap32[3];
    float* r32[3];
};

__global__ __launch_bounds__(kThreads) void to_f32_kernel(ToF32Set s)
{
    const int q = blockIdx.y;
    const Box3& b = s.b[q];
    const int64_t n = b.npts();
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double* a = &s.A[q](i, j, k);
#pragma unroll
        for (int c = 1; c < 7; ++c) { s.A32[q][t + c * n] = float(a[c * s.an[q]]); }
        s.ap32[q][t] = float(s.ap[q](i, j, k));
        s.r32[q][t] = float(s.r[q](i, j, k));
    }
}

// y_q += x_q with x_q FP32 (the mixed-precision change, widened exactly)
// and max |x_q| into out[q], for up to three fields (blockIdx.y).
struct AddMaxSetF {
    Box3 b[3];
    V y[3];
    DView<const float> x[3];
    unsigned long long* out[3];
};

__global__ __launch_bounds__(kThreads) void add_max3_f32_kernel(AddMaxSetF s)
{
    const int q = blockIdx.y;
    const Box3& b = s.b[q];
    unsigned long long best = 0;
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double v = double(s.x[q](i, j, k));
        s.y[q](i, j, k) = s.y[q](i, j, k) + v;
        const unsigned long long bits = dev::abs_bits(v);
        best = bits > best ? bits : best;
    }
    best = dev::block_max_u64<kThreads>(best);
    if (threadIdx.x == 0 && best) { atomicMax(s.out[q], best); }
}

// y += x and max |x| into *out (atomic max of ordered bits).
__global__ __launch_bounds__(kThreads) void add_max_kernel(Box3 b, V y, CV x, unsigned long long* out)
{
    unsigned long long best = 0;
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double v = x(i, j, k);
        y(i, j, k) = y(i, j, k) + v;
        const unsigned long long bits = dev::abs_bits(v);
        best = bits > best ? bits : best;
    }
    best = dev::block_max_u64<kThreads>(best);
    if (threadIdx.x == 0 && best) { atomicMax(out, best); }
}

// y = 0 on b, then y = 0 + x inside [lo, hi] (a field and its ghosts
// cleared, the valid points set to x as an add onto zero would).
__global__ __launch_bounds__(kThreads) void zero_add_kernel(Box3 b, V y, CV x, int lo0, int hi0, int lo1, int hi1,
                                                             int lo2, int hi2)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const bool in = i >= lo0 && i <= hi0 && j >= lo1 && j <= hi1 && k >= lo2 && k <= hi2;
        y(i, j, k) = in ? 0.0 + x(i, j, k) : 0.0;
    }
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

// a_P = keep/dtau + A_P, d = keep (a_P - sum a_nb) (Source/Predictor.cpp:186-201).
// With rhs = -R on the same points (the predictor's right-hand side).
__global__ __launch_bounds__(kThreads) void ap_d_kernel(Box3 b, CV A, int64_t an, CV dt, CV mask, int has_mask, V ap,
                                                         V d, CV R, V rhs)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        rhs(i, j, k) = R(i, j, k) * -1.0;
        const double* a = &A(i, j, k);
        const double keep = has_mask ? 1.0 - mask(i, j, k) : 1.0;
        const double p = keep / dt(i, j, k) + a[0];
        ap(i, j, k) = p;
        const double nb = a[1 * an] + a[2 * an] + a[3 * an] + a[4 * an] + a[5 * an] + a[6 * an];
        d(i, j, k) = keep * (p - nb);
    }
}

// The w rows' VNN cap, plain or slope-aware (Source/Predictor.cpp:202-252).
__global__ __launch_bounds__(kThreads) void w_cap_kernel(Box3 b, CV A, int64_t an, CV dt, CV zp, CV mask, int has_mask,
                                                          V ap, V d, double inv_cap, double s, double dxi0,
                                                          double dxi1, int slope_aware)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double* a = &A(i, j, k);
        const double keep = has_mask ? 1.0 - mask(i, j, k) : 1.0;
        double cap;
        if (slope_aware) {
            double term = inv_cap * (a[5 * an] + a[6 * an]);
            const double hx = 0.5 * (zp(i + 1, j + 1, k) + zp(i + 1, j, k) - zp(i, j + 1, k) - zp(i, j, k)) * dxi0;
            const double hy = 0.5 * (zp(i + 1, j + 1, k) + zp(i, j + 1, k) - zp(i + 1, j, k) - zp(i, j, k)) * dxi1;
            const double slope2 = hx * hx + hy * hy;
            if (slope2 > 0) {
                const double root = sqrt(1.0 + slope2) + 1.0;
                const double rc = slope2 / (root * root);
                const double lim = sc_max(0.0, 1.0 / (s * rc) - 1.0) / dt(i, j, k);
                term = sc_min(term, lim);
            }
            cap = keep * term;
        } else {
            cap = keep * inv_cap * (a[5 * an] + a[6 * an]);
        }
        ap(i, j, k) += cap;
        d(i, j, k) += cap;
    }
}

// rhs = -R (MultiFab::Copy then mult(-1)).
// Per-level sums of a face field over the faces counted once: block
// (chunk, level) adds its contiguous share of the level (i fastest, then j)
// and reduces it with a fixed tree; the level's partials are then added in
// chunk order, so the sums are the same on every call.
constexpr int kLevelChunks = 16;

__global__ __launch_bounds__(kThreads) void level_partial_kernel(CV f, int ilo, int ni, int jlo, int nj, int klo,
                                                                  double* part)
{
    const int k = blockIdx.y;
    const int64_t n = int64_t(ni) * nj;
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = int64_t(blockIdx.x) * per;
    const int64_t b1 = b0 + per < n ? b0 + per : n;
    double s = 0.0;
    for (int64_t t = b0 + threadIdx.x; t < b1; t += kThreads) {
        const int i = ilo + int(t % ni), j = jlo + int(t / ni);
        s += f(i, j, klo + k);
    }
    s = dev::block_sum_f64<kThreads>(s);
    if (threadIdx.x == 0) { part[int64_t(k) * gridDim.x + blockIdx.x] = s; }
}

// The level means: the partials in chunk order over the level's face count.
__global__ __launch_bounds__(kThreads) void level_mean_kernel(const double* part, int nchunks, int nz, double npts,
                                                               double* mean)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nz) { return; }
    double s = 0.0;
    for (int c = 0; c < nchunks; ++c) { s += part[int64_t(k) * nchunks + c]; }
    mean[k] = s / npts;
}

// Point-implicit rotation (Source/Predictor.cpp:260-342).
template <bool kIsU>
__global__ __launch_bounds__(kThreads) void coriolis_kernel(Box3 b, CV A, int64_t an, CV dt, CV ro, CV mask,
                                                             int has_mask, V ap, V d, V r, const double* mean, int klo,
                                                             double fs, double relax, int periodic, int lo_edge,
                                                             int hi_edge)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double keep = has_mask ? 1.0 - mask(i, j, k) : 1.0;
        if (keep == 0) { continue; }
        const int n = kIsU ? i : j;
        if (!periodic && (n == lo_edge || n == hi_edge)) { continue; }
        const double* a = &A(i, j, k);
        const double nb = a[1 * an] + a[2 * an] + a[3 * an] + a[4 * an] + a[5 * an] + a[6 * an];
        const double net = sc_max(a[0] - nb, 0.0);
        const double tau_c = 1.0 / sc_max(fabs(fs) / relax, net);
        const double tau_own = 1.0 / (1.0 / dt(i, j, k) + net);
        const double stiff = fs * fs * sc_max(tau_c, tau_own);
        ap(i, j, k) += stiff;
        d(i, j, k) += stiff;
        double s = 0.0;   // interp<YF, XF> (u rows) or interp<XF, YF> (v rows) of the other residual
        if (kIsU) {
            s += ro(i - 1, j, k);
            s += ro(i, j, k);
            s += ro(i - 1, j + 1, k);
            s += ro(i, j + 1, k);
        } else {
            s += ro(i, j - 1, k);
            s += ro(i + 1, j - 1, k);
            s += ro(i, j, k);
            s += ro(i + 1, j, k);
        }
        const double avg = s * 0.25;
        const double m = mean[k - klo];
        r(i, j, k) += (kIsU ? -fs : fs) * (tau_c * m + tau_own * (avg - m));
    }
}

// Point-implicit buoyancy on the w rows (Source/Predictor.cpp:343-395).
__global__ __launch_bounds__(kThreads) void buoyancy_kernel(Box3 b, CV dt, CV th, CV rt, CV At, CV z, CV mask,
                                                             int has_mask, V ap, V d, V r, double gt, int klo, int khi)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double keep = has_mask ? 1.0 - mask(i, j, k) : 1.0;
        if (keep == 0) { continue; }
        // The ground and top faces are boundary rows (Dirichlet under every
        // boundary condition of the reference): no term, and no read of
        // theta's rows outside the domain.
        if (k == klo || k == khi) { continue; }
        auto zc = [&](int kk) {
            return 0.125 * (z(i, j, kk) + z(i + 1, j, kk) + z(i, j + 1, kk) + z(i + 1, j + 1, kk) + z(i, j, kk + 1) +
                            z(i + 1, j, kk + 1) + z(i, j + 1, kk + 1) + z(i + 1, j + 1, kk + 1));
        };
        const double gamma = (th(i, j, k) - th(i, j, k - 1)) / (zc(k) - zc(k - 1));
        const double ap_theta = 0.5 * (At(i, j, k - 1) + At(i, j, k));
        const double tau = 1.0 / (1.0 / dt(i, j, k) + ap_theta);
        const double stiff = gt * sc_max(gamma, 0.0) * tau;
        ap(i, j, k) += stiff;
        d(i, j, k) += stiff;
        r(i, j, k) -= gt * tau * 0.5 * (rt(i, j, k - 1) + rt(i, j, k));
    }
}

// ap = 1/dtau + A_P (Source/Simplec.cpp:247-257).
// a_P = 1/dtau + A_P and rhs = -R (Source/Simplec.cpp:240-247).
__global__ __launch_bounds__(kThreads) void scalar_ap_kernel(Box3 b, CV A, CV dt, V ap, CV R, V rhs)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        ap(i, j, k) = 1.0 / dt(i, j, k) + A(i, j, k);
        rhs(i, j, k) = R(i, j, k) * -1.0;
    }
}

// The k floor: raise values below it, count them (Source/Simplec.cpp:262-277).
__global__ __launch_bounds__(kThreads) void clip_kernel(Box3 b, V f, double floor, unsigned long long* count)
{
    unsigned long long c = 0;
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        if (f(i, j, k) < floor) {
            f(i, j, k) = floor;
            ++c;
        }
    }
    c = dev::block_sum_u64<kThreads>(c);
    if (threadIdx.x == 0 && c) { atomicAdd(count, c); }
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

// A compact FP32 field of ncomp components over box b from the workspace.
gust_array4 scratch_f32(hipStream_t st, int slot, const IBox& b, int ncomp)
{
    const int64_t n = b.npts();
    gust_array4 v{};
    v.p = Workspace::get(st, slot, size_t(n) * size_t(ncomp) * sizeof(float));
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

// The FP32 twins of n <= 3 FP64 z-line systems (A, a_P, rhs rounded in one
// launch; delta an FP32 field with one ghost), in workspace slots 80-91.
void systems_f32(int n, const gust_line_system* in, gust_line_system* out, hipStream_t st)
{
    ToF32Set c{};
    int64_t most = 1;
    for (int q = 0; q < n; ++q) {
        const gust_line_system& a = in[q];
        const IBox vb = loc_box(a.geom, a.loc);
        gust_line_system& b = out[q];
        b.loc = a.loc;
        b.geom = a.geom;
        b.A = scratch_f32(st, 80 + 4 * q, vb, GUST_N_COEFS);
        b.ap = scratch_f32(st, 81 + 4 * q, vb, 1);
        b.rhs = scratch_f32(st, 82 + 4 * q, vb, 1);
        b.delta = scratch_f32(st, 83 + 4 * q, grow(vb, 1, 1, 1), 1);
        c.b[q] = box3(vb);
        c.A[q] = dview<const double>(&a.A);
        c.an[q] = a.A.nstride;
        c.ap[q] = dview<const double>(&a.ap);
        c.r[q] = dview<const double>(&a.rhs);
        c.A32[q] = static_cast<float*>(b.A.p);
        c.ap32[q] = static_cast<float*>(b.ap.p);
        c.r32[q] = static_cast<float*>(b.rhs.p);
        most = std::max<int64_t>(most, c.b[q].npts());
    }
    to_f32_kernel<<<dim3(blocks_for(most), n), kThreads, 0, st>>>(c);
    GUST_HIP(hipGetLastError());
}

// Reductions into a deferred batch: each registers a slot and the callback
// that receives its value when the batch resolves. Requests queue up and
// run() reduces them all in one launch (at most kJobs; a full queue runs by
// itself): the caller runs the queue before anything overwrites a field
// that is in it.
class Reducer {
public:
    explicit Reducer(Deferred& dq) : dq_(dq), st_(dq.stream()) {}
    Reducer(const Reducer&) = delete;
    Reducer& operator=(const Reducer&) = delete;

    void max_abs(const IBox& box, const gust_array4* a, std::function<void(double)> f,
                 const gust_array4* mask = nullptr)
    {
        unsigned long long* slot = dq_.zero_slot([f](unsigned long long v) { f(bits_to_double(v)); });
        push(box, a, mask, mask ? kMaxAbsMasked : kMaxAbs, slot);
    }
    void min_pos(const IBox& box, const gust_array4* a, std::function<void(double)> f)
    {
        unsigned long long* slot = dq_.ones_slot([f](unsigned long long v) { f(bits_to_double(v)); });
        push(box, a, nullptr, kMinPos, slot);
    }
    // Fails with msg when a is negative, NaN or infinite anywhere.
    void check_nonneg(const IBox& box, const gust_array4* a, std::string msg)
    {
        unsigned long long* slot = dq_.zero_slot([msg](unsigned long long v) {
            if (v) { GUST_FAIL("%s", msg.c_str()); }
        });
        push(box, a, nullptr, kNonNeg, slot);
    }
    // Fails with msg unless a is finite and > 0 everywhere.
    void check_positive(const IBox& box, const gust_array4* a, const std::string& msg)
    {
        check_nonneg(box, a, msg);
        min_pos(box, a, [msg](double m) {
            if (!(m > 0)) { GUST_FAIL("%s", msg.c_str()); }
        });
    }
    // y += x with max |x| reported (one pass).
    // y_q += x_q with max |x_q| to f_q (f_q may be empty: no report) for
    // n <= 3 fields in one launch.
    void add_max_abs_n(int n, const IBox* box, gust_array4* const* y, const gust_array4* const* x,
                       std::vector<std::function<void(double)>> f)
    {
        AddMaxSet a{};
        int64_t most = 1;
        for (int q = 0; q < n; ++q) {
            a.b[q] = box3(box[q]);
            a.y[q] = dview<double>(y[q]);
            a.x[q] = dview<const double>(x[q]);
            std::function<void(double)> fq = f[size_t(q)];
            a.out[q] = dq_.zero_slot([fq](unsigned long long v) {
                if (fq) { fq(bits_to_double(v)); }
            });
            most = std::max<int64_t>(most, a.b[q].npts());
        }
        add_max3_kernel<<<dim3(cgrid::reduce_blocks(most), n), kThreads, 0, st_>>>(a);
        GUST_HIP(hipGetLastError());
    }
    // add_max_abs_n with FP32 changes x_q.
    void add_max_abs_n_f32(int n, const IBox* box, gust_array4* const* y, const gust_array4* const* x,
                           std::vector<std::function<void(double)>> f)
    {
        AddMaxSetF a{};
        int64_t most = 1;
        for (int q = 0; q < n; ++q) {
            a.b[q] = box3(box[q]);
            a.y[q] = dview<double>(y[q]);
            a.x[q] = dview<const float>(x[q]);
            std::function<void(double)> fq = f[size_t(q)];
            a.out[q] = dq_.zero_slot([fq](unsigned long long v) {
                if (fq) { fq(bits_to_double(v)); }
            });
            most = std::max<int64_t>(most, a.b[q].npts());
        }
        add_max3_f32_kernel<<<dim3(cgrid::reduce_blocks(most), n), kThreads, 0, st_>>>(a);
        GUST_HIP(hipGetLastError());
    }
    void add_max_abs(const IBox& box, const gust_array4* y, const gust_array4* x, std::function<void(double)> f)
    {
        const Box3 b = box3(box);
        unsigned long long* slot = dq_.zero_slot([f](unsigned long long v) { f(bits_to_double(v)); });
        add_max_kernel<<<cgrid::reduce_blocks(b.npts()), kThreads, 0, st_>>>(b, dview<double>(y),
                                                                             dview<const double>(x), slot);
        GUST_HIP(hipGetLastError());
    }
    void run()
    {
        if (n_ == 0) { return; }
        int64_t most = 1;
        for (int q = 0; q < n_; ++q) { most = std::max<int64_t>(most, jobs_.j[q].b.npts()); }
        reduce_jobs_kernel<<<dim3(cgrid::reduce_blocks(most), n_), kThreads, 0, st_>>>(jobs_);
        GUST_HIP(hipGetLastError());
        n_ = 0;
    }

private:
    void push(const IBox& box, const gust_array4* a, const gust_array4* m, int op, unsigned long long* slot)
    {
        if (n_ == kJobs) { run(); }
        RJob& J = jobs_.j[n_++];
        J.b = box3(box);
        J.a = dview<const double>(a);
        J.m = dview<const double>(m ? m : a);
        J.op = op;
        J.out = slot;
    }

    Deferred& dq_;
    hipStream_t st_;
    RJobs jobs_{};
    int n_ = 0;
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

// theta's and k's scalar sweeps of one SIMPLEC iteration (Source/Simplec.cpp:
// 222-281 and 494-503: theta's, then k's with its floor clip) with the two
// z-line Jacobi solves in one launch per sweep. Neither system depends on the
// other field (k's rows come from the fluxes, the closure and the wall
// state), so the bits are those of the two scalar_sweep calls in a row.
void scalar_sweeps_theta_tke(const gust_geometry& g, gust_array4* theta, const gust_array4* R_theta,
                             const gust_array4* A_theta, gust_array4* tke, const gust_array4* R_tke,
                             const gust_array4* A_tke, const gust_array4* dtau, int sweeps, double floor,
                             const gust_line_options& sweep, gust_scalar_sweep_result* res_theta,
                             gust_scalar_sweep_result* res_tke, hipStream_t st, int sweep_precision)
{
    check_rc(gust_check_line_options(&sweep));
    // both kinds solve theta and k in shared launches (their systems come
    // from R, A and dtau), in FP64 or (sweep_precision 1) FP32
    const bool lusgs = sweep.kind == GUST_SWEEP_LUSGS;
    const char* op = "scalar_sweep";
    check_field(op, "field", theta, g, GUST_LOC_CELL, 1);
    check_field(op, "field", tke, g, GUST_LOC_CELL, 1);
    check_field(op, "R", R_theta, g, GUST_LOC_CELL, 0);
    check_field(op, "R", R_tke, g, GUST_LOC_CELL, 0);
    check_field(op, "dtau", dtau, g, GUST_LOC_CELL, 0);
    check_field(op, "A", A_theta, g, GUST_LOC_CELL, 0);
    check_field(op, "A", A_tke, g, GUST_LOC_CELL, 0);
    if (A_theta->ncomp != GUST_N_COEFS || A_tke->ncomp != GUST_N_COEFS) {
        GUST_FAIL("gust: scalar_sweep: A needs 7 components");
    }
    if (std::isnan(floor) || !(floor >= 0)) { GUST_FAIL("gust: scalar_sweep: the floor must be >= 0"); }
    DeferredUse dq(st);
    Reducer red(*dq);
    const IBox cells = loc_box(g, GUST_LOC_CELL);
    red.check_positive(cells, dtau, "gust: scalar_sweep: dtau must be finite and > 0 everywhere");
    red.run();
    const Box3 b = box3(cells);
    gust_array4* field[2] = {theta, tke};
    const gust_array4* R[2] = {R_theta, R_tke};
    const gust_array4* A[2] = {A_theta, A_tke};
    gust_scalar_sweep_result* res[2] = {res_theta, res_tke};
    const int slot[2][3] = {{51, 52, 53}, {63, 64, 65}};
    gust_line_system sys[2];
    for (int q = 0; q < 2; ++q) {
        Scratch ap = scratch(st, slot[q][0], cells, 0), rhs = scratch(st, slot[q][1], cells, 0),
                delta = scratch(st, slot[q][2], cells, 1);
        scalar_ap_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, dview<const double>(A[q]),
                                                                    dview<const double>(dtau), dview<double>(&ap.v),
                                                                    dview<const double>(R[q]), dview<double>(&rhs.v));
        GUST_HIP(hipGetLastError());
        if (res[q]) { *res[q] = gust_scalar_sweep_result{}; }
        sys[q] = gust_line_system{GUST_LOC_CELL, g, *A[q], ap.v, rhs.v, delta.v};
    }
    std::vector<Deferred::Callback> fs;
    for (int q = 0; q < 2; ++q) {
        gust_scalar_sweep_result* r = res[q];
        fs.emplace_back([r](unsigned long long v) {
            if (r) { r->last_sweep = bits_to_double(v); }
        });
    }
    gust_line_system s32[2];
    if (lusgs && sweep_precision == 1) {
        systems_f32(2, sys, s32, st);
        zline_lusgs_into_f32(2, s32, &sweep, sweeps, dq->zero_run(std::move(fs)), st);
    } else if (lusgs) {
        zline_lusgs_into(2, sys, &sweep, sweeps, dq->zero_run(std::move(fs)), st);
    } else if (sweep_precision == 1) {
        systems_f32(2, sys, s32, st);
        zline_jacobi_into_f32(2, s32, sweeps, dq->zero_run(std::move(fs)), st);
    } else {
        zline_jacobi_into(2, sys, sweeps, dq->zero_run(std::move(fs)), st);
    }
    {
        const IBox boxes[2] = {cells, cells};
        const gust_line_system* used = sweep_precision == 1 ? s32 : sys;
        const gust_array4* xs[2] = {&used[0].delta, &used[1].delta};
        std::vector<std::function<void(double)>> fs2(2);
        for (int q = 0; q < 2; ++q) {
            gust_scalar_sweep_result* r = res[q];
            if (r) { fs2[size_t(q)] = [r](double x) { r->max_dq = x; }; }
        }
        if (sweep_precision == 1) { red.add_max_abs_n_f32(2, boxes, field, xs, std::move(fs2)); }
        else { red.add_max_abs_n(2, boxes, field, xs, std::move(fs2)); }
    }
    unsigned long long* cnt = dq->zero_slot([res_tke](unsigned long long v) {
        if (res_tke) { res_tke->clipped = static_cast<int64_t>(v); }
    });
    clip_kernel<<<cgrid::reduce_blocks(b.npts()), kThreads, 0, st>>>(b, dview<double>(tke), floor, cnt);
    GUST_HIP(hipGetLastError());
    dq.finish();
}

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
        DeferredUse dq(st);
        Reducer red(*dq);
        // the values the checks and the report need, read when the batch resolves
        struct Vals {
            double speed[3] = {0, 0, 0};
            double jmin = 0;
            gust_pressure_correction_result res{};
        };
        auto vals = std::make_shared<Vals>();
        // face_coefficients (Simplec.cpp:87-112)
        FaceSet3 bk{};
        int64_t most = 1;
        for (int d = 0; d < 3; ++d) {
            beta[d] = scratch(st, 34 + d, fb[d], 0);
            keep[d] = scratch(st, 37 + d, fb[d], 0);
            F[d] = scratch(st, 40 + d, fb[d], 0);
            red.check_nonneg(fb[d], dd[d], std::string("gust: pressure_correction: the SIMPLEC denominator ") +
                                               dname[d] + " must be finite and >= 0 everywhere");
            bk.b[d] = box3(fb[d]);
            bk.d[d] = dview<const double>(dd[d]);
            bk.beta[d] = dview<double>(&beta[d].v);
            bk.keep[d] = dview<double>(&keep[d].v);
            most = std::max<int64_t>(most, bk.b[d].npts());
        }
        beta_keep3_kernel<<<dim3(blocks_for(most), 3), kThreads, 0, st>>>(bk);
        GUST_HIP(hipGetLastError());
        check_rc(gust_velocity_divergence(geom, u, v, w, z_nd, ax, ay, az, detJ, &rhs.v, stream));
        red.max_abs(cells, &rhs.v, [vals](double x) { vals->res.divergence_before = x; });
        if (source) {
            red.run();
            box_op(st, cells, &rhs.v, source, 0.0, 2);
        }
        red.max_abs(fb[0], u, [vals](double x) { vals->speed[0] = x; });
        red.max_abs(fb[1], v, [vals](double x) { vals->speed[1] = x; });
        red.max_abs(fb[2], w, [vals](double x) { vals->speed[2] = x; });
        red.min_pos(cells, detJ, [vals](double x) { vals->jmin = x; });
        red.run();
        gust_face_coeffs coeffs{&beta[0].v, &beta[1].v, &beta[2].v, nullptr, 1.0};
        box_op(st, grow(cells, 1, 1, 1), &phi.v, nullptr, 0.0, 3);
        {
            const int rc = gust_poisson_solve(solver, &rhs.v, &phi.v, &coeffs, bc, &F[0].v, &F[1].v, &F[2].v,
                                              &vals->res.poisson, stream);
            if (rc != 0) {
                // an earlier check's error (a bad denominator) explains a failed solve
                std::string err = gust_last_error();
                dq->resolve();
                set_error("%s", err.c_str());
                throw Fail{rc};
            }
        }
        const double dx0 = g.dx[0], dx1 = g.dx[1], dz = g.dx[2];
        dq->then([vals, project_net_source, dx0, dx1, dz] {
            const double speed = std::max({vals->speed[0], vals->speed[1], vals->speed[2]});
            const double dz_min = vals->jmin * dz;
            const double inv_h = 1.0 / dx0 + 1.0 / dx1 + 1.0 / dz_min;
            const gust_poisson_result& pr = vals->res.poisson;
            if (!project_net_source && pr.singular && std::abs(pr.removed_mean) > 1.0e-8 * speed * inv_h) {
                GUST_FAIL("gust: pressure_correction: the predicted velocity has a net source the boundaries cannot "
                          "carry: J-weighted mean divergence %s 1/s against the scale %s 1/s (a boundary condition "
                          "or the predictor is inconsistent)",
                          std::to_string(pr.removed_mean).c_str(), std::to_string(speed * inv_h).c_str());
            }
        });
        // p += alpha_p phi; u, v, Omega += the masked fluxes; w follows.
        box_op(st, cells, p, &phi.v, alpha_p, 0);
        red.max_abs(cells, &phi.v, [vals, alpha_p](double x) { vals->res.max_dp = alpha_p * x; });
        red.max_abs(fb[0], &F[0].v, [vals](double x) { vals->res.max_du = x; }, &keep[0].v);
        red.max_abs(fb[1], &F[1].v, [vals](double x) { vals->res.max_dv = x; }, &keep[1].v);
        red.max_abs(fb[2], &F[2].v, [vals](double x) { vals->res.max_dw = x; }, &keep[2].v);
        red.run();
        // Omega from the predicted velocities, then F *= keep and u, v, Omega
        // += F in one launch (Omega does not read F)
        check_rc(gust_omega_from_w(geom, u, v, w, z_nd, &omega.v, 0, stream));
        {
            CorrSet3 cs{};
            gust_array4* y3[3] = {u, v, &omega.v};
            int64_t most_c = 1;
            for (int d = 0; d < 3; ++d) {
                cs.b[d] = box3(fb[d]);
                cs.F[d] = dview<double>(&F[d].v);
                cs.y[d] = dview<double>(y3[d]);
                cs.keep[d] = dview<const double>(&keep[d].v);
                most_c = std::max<int64_t>(most_c, cs.b[d].npts());
            }
            correct3_kernel<<<dim3(blocks_for(most_c), 3), kThreads, 0, st>>>(cs);
            GUST_HIP(hipGetLastError());
        }
        override_sync(st, g, GUST_LOC_XFACE, u);
        override_sync(st, g, GUST_LOC_YFACE, v);
        override_sync(st, g, GUST_LOC_ZFACE, &omega.v);
        fill_boundary(st, g, GUST_LOC_XFACE, u);
        fill_boundary(st, g, GUST_LOC_YFACE, v);
        check_rc(gust_w_from_omega(geom, u, v, &omega.v, z_nd, w, stream));
        override_sync(st, g, GUST_LOC_ZFACE, w);
        fill_boundary(st, g, GUST_LOC_ZFACE, w);
        check_rc(gust_velocity_divergence(geom, u, v, w, z_nd, ax, ay, az, detJ, &div.v, stream));
        red.max_abs(cells, &div.v, [vals, result](double x) {
            vals->res.divergence_after = x;
            if (result) { *result = vals->res; }
        });
        red.run();
        dq.finish();
    });
}

int gust_momentum_predictor(const gust_geometry* geom, gust_faces* vel, const gust_array4* theta,
                            const gust_array4* z_nd, const gust_faces* R, const gust_array4* R_theta,
                            const gust_faces* A, const gust_array4* A_theta, const gust_faces* dtau,
                            const gust_faces* mask, const gust_predictor_options* opt, gust_faces* ap, gust_faces* d,
                            gust_momentum_predictor_result* result, void* stream)
{
    return guarded([&] {
        const char* op = "momentum_predictor";
        hipStream_t st = (hipStream_t)stream;
        check_geom(op, geom);
        if (!vel || !R || !A || !dtau || !opt || !ap || !d) { GUST_FAIL("gust: momentum_predictor: null argument"); }
        const double relax = opt->rotation_relax, cap = opt->w_vnn_cap, margin = opt->w_vnn_slope_margin;
        if (std::isnan(relax) || !(relax > 0 && relax <= 1)) {
            GUST_FAIL("gust: momentum_predictor: rotation_relax must be > 0 and <= 1 (simplec.rotation_relax)");
        }
        if (std::isnan(cap) || !(cap == 0 || (cap > 0 && std::isfinite(cap) && std::isfinite(1 / cap)))) {
            GUST_FAIL("gust: momentum_predictor: w_vnn_cap must be 0 (off), or finite and > 0 with a finite "
                      "reciprocal (simplec.w_vnn_cap)");
        }
        if (std::isnan(margin) || !(margin == 0 || (margin >= 1 && std::isfinite(margin)))) {
            GUST_FAIL("gust: momentum_predictor: w_vnn_slope_margin must be 0 (off), or finite and >= 1 "
                      "(simplec.w_vnn_cap_slope_margin)");
        }
        if (opt->sweeps < 1) { GUST_FAIL("gust: momentum_predictor: sweeps must be >= 1 (got %d)", opt->sweeps); }
        const gust_geometry& g = *geom;
        const char* dname[3] = {"dtau_x", "dtau_y", "dtau_z"};
        DeferredUse dq(st);
        Reducer red(*dq);
        for (int c = 0; c < 3; ++c) {
            check_field(op, dname[c], dtau->c[c], g, 1 << c, 0);
            const IBox fb = loc_box(g, 1 << c);
            // finite and > 0: the non-negative flag and a zero minimum both fail
            red.check_positive(fb, dtau->c[c], "gust: momentum_predictor: dtau_faces must be finite and > 0 everywhere");
        }
        red.run();
        const bool rotation = opt->coriolis_f != 0;
        const bool buoyancy = opt->gravity > 0;
        if ((rotation || buoyancy) && !mask) {
            GUST_FAIL("gust: momentum_predictor: the point-implicit buoyancy and rotation need the boundary "
                      "conditions (the Dirichlet rows they skip); give in.bc, or gravity 0 and Coriolis off");
        }
        check_rc(gust_check_line_options(&opt->sweep));
        if (opt->sweep_precision != 0 && opt->sweep_precision != 1) {
            GUST_FAIL("gust: momentum_predictor: sweep_precision must be 0 (FP64) or 1 (FP32 sweeps)");
        }
        double* dmax[3] = {nullptr, nullptr, nullptr};
        double* last[3] = {nullptr, nullptr, nullptr};
        if (result) {
            *result = gust_momentum_predictor_result{};
            dmax[0] = &result->max_du;
            dmax[1] = &result->max_dv;
            dmax[2] = &result->max_dw;
            last[0] = &result->last_sweep_du;
            last[1] = &result->last_sweep_dv;
            last[2] = &result->last_sweep_dw;
        }
        const double dxi0 = 1.0 / g.dx[0], dxi1 = 1.0 / g.dx[1];
        // The three line solves share launches (z-line Jacobi: one per
        // sweep; LU-SGS: one for all sweeps): their systems (a_P, d, rhs)
        // come from R, A and dtau only, never from the updated velocities,
        // so solving them together gives the bits of the reference's
        // one-after-the-other loop.
        const bool batched = true;
        const bool lusgs = opt->sweep.kind == GUST_SWEEP_LUSGS;
        const int rslot[3] = {44, 55, 57}, dslot[3] = {45, 56, 58};
        gust_line_system sys[3];
        for (int c = 0; c < 3; ++c) {
            const int loc = 1 << c;
            const IBox fb = loc_box(g, loc);
            const Box3 b = box3(fb);
            check_field(op, "q", vel->c[c], g, loc, 0);
            check_field(op, "R", R->c[c], g, loc, 0);
            check_field(op, "A", A->c[c], g, loc, 0);
            if (A->c[c]->ncomp != GUST_N_COEFS) { GUST_FAIL("gust: momentum_predictor: A needs 7 components"); }
            check_field(op, "ap", ap->c[c], g, loc, 0);
            check_field(op, "d", d->c[c], g, loc, 0);
            const gust_array4* m = mask ? mask->c[c] : nullptr;
            if (m) { check_field(op, "mask", m, g, loc, 0); }
            const CV mk = dview<const double>(m ? m : A->c[c]);
            const int64_t an = A->c[c]->nstride;
            Scratch rhs = scratch(st, batched ? rslot[c] : 44, fb, 0), delta = scratch(st, batched ? dslot[c] : 45, fb, 1);
            ap_d_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, dview<const double>(A->c[c]), an,
                                                                   dview<const double>(dtau->c[c]), mk, m ? 1 : 0,
                                                                   dview<double>(ap->c[c]), dview<double>(d->c[c]),
                                                                   dview<const double>(R->c[c]), dview<double>(&rhs.v));
            if (c == 2 && cap > 0) {
                check_nodal_heights(op, z_nd, g);
                w_cap_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(
                    b, dview<const double>(A->c[c]), an, dview<const double>(dtau->c[c]), dview<const double>(z_nd),
                    mk, m ? 1 : 0, dview<double>(ap->c[c]), dview<double>(d->c[c]), 1.0 / cap, margin, dxi0, dxi1,
                    margin > 0 ? 1 : 0);
            }
            GUST_HIP(hipGetLastError());
            if (c < 2 && rotation) {
                // the other residual with its periodic images (FillBoundary)
                const int co = 1 - c, lo_other = 1 << co;
                const IBox ob = loc_box(g, lo_other);
                Scratch og = scratch(st, 46, ob, 1);
                {
                    // og = 0 on the grown box, then 0 + R on the valid box
                    // (one pass; the two passes' bits)
                    const Box3 gb = box3(grow(ob, 1, 1, 1));
                    zero_add_kernel<<<blocks_for(gb.npts()), kThreads, 0, st>>>(
                        gb, dview<double>(&og.v), dview<const double>(R->c[co]), ob.lo[0], ob.hi[0], ob.lo[1],
                        ob.hi[1], ob.lo[2], ob.hi[2]);
                    GUST_HIP(hipGetLastError());
                }
                fill_boundary(st, g, lo_other, &og.v);
                // per-level mean over every face once
                const bool per_o = g.periodic[co] != 0;
                IBox sb = ob;
                if (per_o) { sb.hi[co] -= 1; }
                const int nz = g.hi[2] - g.lo[2] + 1;
                auto* part = static_cast<double*>(Workspace::get(st, 47, size_t(nz) * kLevelChunks * sizeof(double)));
                level_partial_kernel<<<dim3(kLevelChunks, nz), kThreads, 0, st>>>(
                    dview<const double>(&og.v), sb.lo[0], sb.hi[0] - sb.lo[0] + 1, sb.lo[1], sb.hi[1] - sb.lo[1] + 1,
                    g.lo[2], part);
                const double npts_level = double((g.hi[co] - g.lo[co] + 1) + (per_o ? 0 : 1)) *
                                          double(g.hi[c] - g.lo[c] + 1);
                auto* dmean = static_cast<double*>(Workspace::get(st, 48, size_t(nz) * sizeof(double)));
                level_mean_kernel<<<(nz + kThreads - 1) / kThreads, kThreads, 0, st>>>(part, kLevelChunks, nz,
                                                                                      npts_level, dmean);
                GUST_HIP(hipGetLastError());
                const int lo_edge = g.lo[c], hi_edge = g.hi[c] + 1;
                const int periodic = g.periodic[c] ? 1 : 0;
                if (c == 0) {
                    coriolis_kernel<true><<<blocks_for(b.npts()), kThreads, 0, st>>>(
                        b, dview<const double>(A->c[c]), an, dview<const double>(dtau->c[c]),
                        dview<const double>(&og.v), mk, m ? 1 : 0, dview<double>(ap->c[c]), dview<double>(d->c[c]),
                        dview<double>(&rhs.v), dmean, g.lo[2], opt->coriolis_f, relax, periodic, lo_edge, hi_edge);
                } else {
                    coriolis_kernel<false><<<blocks_for(b.npts()), kThreads, 0, st>>>(
                        b, dview<const double>(A->c[c]), an, dview<const double>(dtau->c[c]),
                        dview<const double>(&og.v), mk, m ? 1 : 0, dview<double>(ap->c[c]), dview<double>(d->c[c]),
                        dview<double>(&rhs.v), dmean, g.lo[2], opt->coriolis_f, relax, periodic, lo_edge, hi_edge);
                }
                GUST_HIP(hipGetLastError());
            }
            if (c == 2 && buoyancy) {
                check_field(op, "R_theta", R_theta, g, GUST_LOC_CELL, 0);
                check_field(op, "theta", theta, g, GUST_LOC_CELL, 1);
                if (!A_theta) { GUST_FAIL("gust: momentum_predictor: the buoyancy needs theta's implicit rows"); }
                check_field(op, "A_theta", A_theta, g, GUST_LOC_CELL, 0);
                check_nodal_heights(op, z_nd, g);
                const double gt = opt->gravity / opt->theta0;
                buoyancy_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(
                    b, dview<const double>(dtau->c[c]), dview<const double>(theta), dview<const double>(R_theta),
                    dview<const double>(A_theta), dview<const double>(z_nd), mk, m ? 1 : 0, dview<double>(ap->c[c]),
                    dview<double>(d->c[c]), dview<double>(&rhs.v), gt, g.lo[2], g.hi[2] + 1);
                GUST_HIP(hipGetLastError());
            }
            if (batched) {
                sys[c] = gust_line_system{loc, g, *A->c[c], *ap->c[c], rhs.v, delta.v};
                continue;
            }
            check_rc(gust_line_sweeps(&opt->sweep, loc, A->c[c], ap->c[c], &rhs.v, opt->sweeps, geom, &delta.v,
                                      last[c], stream));
            if (dmax[c]) {
                double* out = dmax[c];
                red.add_max_abs(fb, vel->c[c], &delta.v, [out](double x) { *out = x; });
            } else {
                box_add(st, fb, vel->c[c], &delta.v);
            }
        }
        if (batched) {
            std::vector<Deferred::Callback> fs;
            for (int c = 0; c < 3; ++c) {
                double* out = last[c];
                fs.emplace_back([out](unsigned long long v) {
                    if (out) { *out = bits_to_double(v); }
                });
            }
            gust_line_system s32[3];
            const bool f32 = opt->sweep_precision == 1;
            if (lusgs && f32) {
                systems_f32(3, sys, s32, st);
                zline_lusgs_into_f32(3, s32, &opt->sweep, opt->sweeps, dq->zero_run(std::move(fs)), st);
            } else if (lusgs) {
                zline_lusgs_into(3, sys, &opt->sweep, opt->sweeps, dq->zero_run(std::move(fs)), st);
            } else if (f32) {
                systems_f32(3, sys, s32, st);
                zline_jacobi_into_f32(3, s32, opt->sweeps, dq->zero_run(std::move(fs)), st);
            } else {
                zline_jacobi_into(3, sys, opt->sweeps, dq->zero_run(std::move(fs)), st);
            }
            IBox boxes[3];
            const gust_array4* xs[3];
            std::vector<std::function<void(double)>> fs3(3);
            for (int c = 0; c < 3; ++c) {
                boxes[c] = loc_box(g, 1 << c);
                xs[c] = f32 ? &s32[c].delta : &sys[c].delta;
                if (dmax[c]) {
                    double* out = dmax[c];
                    fs3[size_t(c)] = [out](double x) { *out = x; };
                }
            }
            if (f32) { red.add_max_abs_n_f32(3, boxes, vel->c, xs, std::move(fs3)); }
            else { red.add_max_abs_n(3, boxes, vel->c, xs, std::move(fs3)); }
        }
        dq.finish();
    });
}

int gust_scalar_sweep(const gust_geometry* geom, int32_t eq, gust_array4* field, const gust_array4* R_eq,
                      const gust_array4* A, const gust_array4* dtau, int32_t sweeps, double floor, int32_t clip,
                      const gust_line_options* sweep, gust_scalar_sweep_result* result, void* stream)
{
    return guarded([&] {
        const char* op = "scalar_sweep";
        hipStream_t st = (hipStream_t)stream;
        check_geom(op, geom);
        if (eq != GUST_EQ_THETA && eq != GUST_EQ_TKE) { GUST_FAIL("gust: scalar_sweep: eq must be theta or tke"); }
        const gust_geometry& g = *geom;
        check_field(op, "field", field, g, GUST_LOC_CELL, 1);
        check_field(op, "R", R_eq, g, GUST_LOC_CELL, 0);
        check_field(op, "dtau", dtau, g, GUST_LOC_CELL, 0);
        check_field(op, "A", A, g, GUST_LOC_CELL, 0);
        if (A->ncomp != GUST_N_COEFS) { GUST_FAIL("gust: scalar_sweep: A needs 7 components"); }
        if (std::isnan(floor) || !(floor >= 0)) { GUST_FAIL("gust: scalar_sweep: the floor must be >= 0"); }
        if (!sweep) { GUST_FAIL("gust: scalar_sweep: null line options"); }
        DeferredUse dq(st);
        Reducer red(*dq);
        const IBox cells = loc_box(g, GUST_LOC_CELL);
        red.check_positive(cells, dtau, "gust: scalar_sweep: dtau must be finite and > 0 everywhere");
        red.run();
        const Box3 b = box3(cells);
        Scratch ap = scratch(st, 51, cells, 0), rhs = scratch(st, 52, cells, 0), delta = scratch(st, 53, cells, 1);
        scalar_ap_kernel<<<blocks_for(b.npts()), kThreads, 0, st>>>(b, dview<const double>(A), dview<const double>(dtau),
                                                                    dview<double>(&ap.v), dview<const double>(R_eq),
                                                                    dview<double>(&rhs.v));
        GUST_HIP(hipGetLastError());
        if (result) { *result = gust_scalar_sweep_result{}; }
        check_rc(gust_line_sweeps(sweep, GUST_LOC_CELL, A, &ap.v, &rhs.v, sweeps, geom, &delta.v,
                                  result ? &result->last_sweep : nullptr, stream));
        if (result) {
            red.add_max_abs(cells, field, &delta.v, [result](double x) { result->max_dq = x; });
        } else {
            box_add(st, cells, field, &delta.v);
        }
        if (eq == GUST_EQ_TKE && clip) {
            unsigned long long* cnt = dq->zero_slot([result](unsigned long long v) {
                if (result) { result->clipped = static_cast<int64_t>(v); }
            });
            clip_kernel<<<cgrid::reduce_blocks(b.npts()), kThreads, 0, st>>>(b, dview<double>(field), floor, cnt);
            GUST_HIP(hipGetLastError());
        }
        dq.finish();
    });
}

}   // extern "C"
