// newton.hip -- the Newton-Krylov finish of the steady solve on a tile.
//
// Reference behaviour (GUST):
//   Source/Residual.cpp:324-396        jvp: a forward difference of the full residual
//   Source/Newton.H, Newton.cpp        (D + J) dq = -R(q) by flexible GMRES in scaled
//                                      variables, the operator's rows, the usable update
//   Source/NewtonPreconditioner.*      one SIMPLEC iteration on the linearised system
//                                      (the diagonal pressure denominators)
//   Source/Continuation.*              pseudo-transient continuation: SER CFL, Armijo
//                                      line search, retries, the switch to infinite dtau
//   Source/NewtonStage.cpp:91-207      the stage: init, one continuation step per
//                                      iteration, the run's norms, the hand-backs
//   Source/Driver.cpp:335-400          the switch rules between the stages
//
// Design (MI355X): a Newton vector is the six fields' valid points packed one
// after the other in one device array (u, v, w at their faces, p, theta, k at
// the cells, i fastest), so every vector operation of FGMRES is one launch
// over one array. Inner products carry an owner mask (the hi copy of a
// periodic seam face is not counted) and reduce in a fixed tree, so a tile
// gives the same bits on every call. The Arnoldi j + 1 dots of classical
// Gram-Schmidt (with one reorthogonalisation) share one launch and one host
// read; the Hessenberg matrix and its Givens rotations live on the host. J v
// is the library's residual on q + eps v (no Jacobian is formed), the
// preconditioner the library's predictor, pressure correction and scalar
// sweeps on the increment.
#include "cgrid.hpp"
#include "deferred.hpp"
#include "heights.hpp"
#include "simplec.hpp"
#include "stage.hpp"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
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

constexpr int kNF = 6;
constexpr int kDotBlocks = 256;

// ---------------------------------------------------------------------------
// Device kernels.

// The six fields' valid boxes and their offsets in a packed vector.
struct Packing {
    Box3 box[kNF];
    int64_t off[kNF + 1];
    int64_t n() const { return off[kNF]; }
};

struct PackArgs {
    CV src[kNF];
    Box3 box[kNF];
    int64_t off[kNF];
    double scale[kNF];
};

// out[off_f + t] = scale_f * src_f at the t-th valid point of field f.
__global__ __launch_bounds__(kThreads) void pack_kernel(PackArgs a, double* __restrict__ out)
{
    const int f = blockIdx.y;
    const Box3 b = a.box[f];
    const CV s = a.src[f];
    const double sc = a.scale[f];
    double* o = out + a.off[f];
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        o[t] = sc * s(i, j, k);
    }
}

struct UnpackArgs {
    V dst[kNF];
    Box3 box[kNF];
    int64_t off[kNF];
    double scale[kNF];
};

// dst_f = scale_f * in[off_f + t] on the valid points.
__global__ __launch_bounds__(kThreads) void unpack_kernel(UnpackArgs a, const double* __restrict__ in)
{
    const int f = blockIdx.y;
    const Box3 b = a.box[f];
    const V d = a.dst[f];
    const double sc = a.scale[f];
    const double* x = in + a.off[f];
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        d(i, j, k) = sc * x[t];
    }
}

// Six whole views (ghosts included) at once: y = a x + b z over the box.
struct ViewSet {
    V y[kNF];
    CV x[kNF];
    CV z[kNF];
    Box3 box[kNF];
};

__global__ __launch_bounds__(kThreads) void view_lincomb_kernel(ViewSet s, double a, double b)
{
    const int f = blockIdx.y;
    const Box3 bx = s.box[f];
    GUST_GRID_LOOP(bx)
    {
        int i, j, k;
        bx.at(t, i, j, k);
        s.y[f](i, j, k) = a * s.x[f](i, j, k) + b * s.z[f](i, j, k);
    }
}

__global__ __launch_bounds__(kThreads) void view_copy_kernel(ViewSet s)
{
    const int f = blockIdx.y;
    const Box3 bx = s.box[f];
    GUST_GRID_LOOP(bx)
    {
        int i, j, k;
        bx.at(t, i, j, k);
        s.y[f](i, j, k) = s.x[f](i, j, k);
    }
}

__global__ __launch_bounds__(kThreads) void view_zero_kernel(ViewSet s)
{
    const int f = blockIdx.y;
    const Box3 bx = s.box[f];
    GUST_GRID_LOOP(bx)
    {
        int i, j, k;
        bx.at(t, i, j, k);
        s.y[f](i, j, k) = 0.0;
    }
}

// y += a x on the valid points of six fields (MultiFab::Saxpy, nghost 0).
__global__ __launch_bounds__(kThreads) void view_saxpy_kernel(ViewSet s, double a)
{
    const int f = blockIdx.y;
    const Box3 bx = s.box[f];
    GUST_GRID_LOOP(bx)
    {
        int i, j, k;
        bx.at(t, i, j, k);
        s.y[f](i, j, k) += a * s.x[f](i, j, k);
    }
}

// Jv = (Jv - R0) * inv_eps on the valid points (jvp's Subtract then mult).
__global__ __launch_bounds__(kThreads) void jv_finish_kernel(ViewSet s, double inv_eps)
{
    const int f = blockIdx.y;
    const Box3 bx = s.box[f];
    GUST_GRID_LOOP(bx)
    {
        int i, j, k;
        bx.at(t, i, j, k);
        const double d = s.y[f](i, j, k) - s.x[f](i, j, k);
        s.y[f](i, j, k) = d * inv_eps;
    }
}

__global__ __launch_bounds__(kThreads) void fill_kernel(int64_t n, double v, double* __restrict__ y)
{
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        y[t] = v;
    }
}

// y = a x + b y
__global__ __launch_bounds__(kThreads) void axpby_kernel(int64_t n, double a, const double* __restrict__ x, double b,
                                                         double* __restrict__ y)
{
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        y[t] = a * x[t] + b * y[t];
    }
}

// y = a x (y is not read)
__global__ __launch_bounds__(kThreads) void scale_copy_kernel(int64_t n, double a, const double* __restrict__ x,
                                                              double* __restrict__ y)
{
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        y[t] = a * x[t];
    }
}

// w += sign * sum_j c[j] V_j (c on the device)
__global__ __launch_bounds__(kThreads) void multi_axpy_kernel(int64_t n, int nv, const double* __restrict__ Vb,
                                                              int64_t ldv, const double* __restrict__ c, double sign,
                                                              double* __restrict__ w)
{
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        double s = 0.0;
        for (int j = 0; j < nv; ++j) { s += c[j] * Vb[j * ldv + t]; }
        w[t] += sign * s;
    }
}

// partial[j][block] = sum over the block's points of own * V_j * w, a fixed
// order per thread and a fixed tree per block.
__global__ __launch_bounds__(kThreads) void mdot_kernel(int64_t n, const double* __restrict__ Vb, int64_t ldv,
                                                        const double* __restrict__ w,
                                                        const unsigned char* __restrict__ own,
                                                        double* __restrict__ partial)
{
    __shared__ double sh[kThreads];
    const int j = blockIdx.y;
    const double* v = Vb + int64_t(j) * ldv;
    double s = 0.0;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        if (own[t]) { s += v[t] * w[t]; }
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) { sh[threadIdx.x] += sh[threadIdx.x + h]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { partial[j * kDotBlocks + blockIdx.x] = sh[0]; }
}

__global__ __launch_bounds__(kThreads) void mdot_final_kernel(const double* __restrict__ partial,
                                                              double* __restrict__ out)
{
    __shared__ double sh[kThreads];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < kDotBlocks; b += kThreads) { s += partial[j * kDotBlocks + b]; }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) { sh[threadIdx.x] += sh[threadIdx.x + h]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[j] = sh[0]; }
}

// partial[block] = (max |x|, min x, any NaN) over the block's points of
// x[0, n): a fixed tree per block; the final pass in block order.
__global__ __launch_bounds__(kThreads) void range_kernel(int64_t n, const double* __restrict__ x,
                                                         double* __restrict__ partial)
{
    __shared__ double smax[kThreads], smin[kThreads], snan[kThreads];
    double mx = 0.0, mn = DBL_MAX, nan = 0.0;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        const double v = x[t];
        if (v != v) {
            nan = 1.0;
            continue;
        }
        mx = fmax(mx, fabs(v));
        mn = fmin(mn, v);
    }
    smax[threadIdx.x] = mx;
    smin[threadIdx.x] = mn;
    snan[threadIdx.x] = nan;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + h]);
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + h]);
            snan[threadIdx.x] = fmax(snan[threadIdx.x], snan[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x] = smax[0];
        partial[3 * blockIdx.x + 1] = smin[0];
        partial[3 * blockIdx.x + 2] = snan[0];
    }
}

__global__ void range_final_kernel(const double* __restrict__ partial, int nblk, double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) { return; }
    double mx = 0.0, mn = DBL_MAX, nan = 0.0;
    for (int b = 0; b < nblk; ++b) {
        mx = fmax(mx, partial[3 * b]);
        mn = fmin(mn, partial[3 * b + 1]);
        nan = fmax(nan, partial[3 * b + 2]);
    }
    out[0] = mx;
    out[1] = mn;
    out[2] = nan;
}

struct FieldOf {
    int64_t off[kNF + 1];
    double tinv[kNF];
    __device__ __forceinline__ int of(int64_t t) const
    {
        int f = 0;
        while (f < kNF - 1 && t >= off[f + 1]) { ++f; }
        return f;
    }
};

// NewtonOperator::apply's combination (Source/Newton.cpp:56-91): y = (inv dq
// + Jv) / t + dir x, with or without each of inv and dir.
__global__ __launch_bounds__(kThreads) void combine_kernel(int64_t n, FieldOf fo, const double* __restrict__ inv,
                                                           const double* __restrict__ dir,
                                                           const double* __restrict__ dq,
                                                           const double* __restrict__ jv,
                                                           const double* __restrict__ x, double* __restrict__ y)
{
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        const int f = fo.of(t);
        const double ti = fo.tinv[f];
        const bool has_inv = inv != nullptr && f != 3;
        const bool has_dir = dir != nullptr && f < 3;
        if (has_inv && has_dir) {
            y[t] = (inv[t] * dq[t] + jv[t]) * ti + dir[t] * x[t];
        } else if (has_inv) {
            y[t] = (inv[t] * dq[t] + jv[t]) * ti;
        } else if (has_dir) {
            y[t] = jv[t] * ti + dir[t] * x[t];
        } else {
            y[t] = jv[t] * ti;
        }
    }
}

// The Dirichlet rows of u, v, w take r: y = m r + (1 - m) y
// (NewtonPreconditioner.cpp:37-49).
__global__ __launch_bounds__(kThreads) void dirichlet_take_kernel(int64_t n, const double* __restrict__ m,
                                                                  const double* __restrict__ r,
                                                                  double* __restrict__ y)
{
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        y[t] = m[t] * r[t] + (1.0 - m[t]) * y[t];
    }
}

// inv = (1 - dir) / dtau on a face field's valid points (Newton.cpp:93-116).
__global__ __launch_bounds__(kThreads) void inv_dtau_kernel(Box3 b, CV dt, CV m, int has_m, double* __restrict__ out)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        out[t] = has_m ? (1.0 - m(i, j, k)) / dt(i, j, k) : 1.0 / dt(i, j, k);
    }
}

// The packed periodic copies: the hi seam face takes the lo seam face's value
// (OverrideSync with the low owner) for a face field periodic along its own
// direction, in a packed vector.
__global__ __launch_bounds__(kThreads) void seam_override_kernel(double* __restrict__ x, int64_t off, int n0, int n1,
                                                                 int n2, int axis)
{
    const int64_t cnt = axis == 0 ? int64_t(n1) * n2 : int64_t(n0) * n2;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < cnt; t += int64_t(gridDim.x) * blockDim.x) {
        int64_t lo, hi;
        if (axis == 0) {
            const int64_t j = t % n1, k = t / n1;
            lo = (k * n1 + j) * n0;
            hi = lo + (n0 - 1);
        } else {
            const int64_t i = t % n0, k = t / n0;
            lo = k * int64_t(n1) * n0 + i;
            hi = lo + int64_t(n1 - 1) * n0;
        }
        x[off + hi] = x[off + lo];
    }
}

// The k floor on the valid cells (the trial points, Continuation.cpp:373).
__global__ __launch_bounds__(kThreads) void floor_kernel(Box3 b, V k, double floor)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, kk;
        b.at(t, i, j, kk);
        if (k(i, j, kk) < floor) { k(i, j, kk) = floor; }
    }
}

// d = a_P on the faces with d > 0 (NewtonPreconditioner.cpp:51-64).
__global__ __launch_bounds__(kThreads) void diag_denominator_kernel(Box3 b, V d, CV ap)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        if (d(i, j, k) > 0) { d(i, j, k) = ap(i, j, k); }
    }
}

// R += a x on the valid cells (theta's coupling term).
__global__ __launch_bounds__(kThreads) void box_saxpy_kernel(Box3 b, V y, double a, CV x, CV x0)
{
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double d = x(i, j, k) - x0(i, j, k);
        y(i, j, k) += a * d;
    }
}

// ---------------------------------------------------------------------------
// Host side.

struct DBuf {
    double* p = nullptr;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf()
    {
        if (p) { (void)hipFree(p); }
    }
    void alloc(int64_t n)
    {
        GUST_HIP(hipMalloc(&p, size_t(std::max<int64_t>(n, 1)) * sizeof(double)));
    }
};

// A state of the stage's layout (two ghost layers) owned here.
struct DevState {
    DevField f[kNF];
    gust_state s{};
    void alloc(const gust_geometry& g, hipStream_t st)
    {
        for (int n = 0; n < kNF; ++n) {
            f[n].alloc(grow(loc_box(g, kStageLocs[n]), 2, 2, 2), 1, st);
            s.f[n] = &f[n].v;
        }
    }
};

// Residual-shaped fields (valid points only).
struct DevRes {
    DevField f[kNF];
    gust_residuals r{};
    void alloc(const gust_geometry& g, hipStream_t st)
    {
        for (int n = 0; n < kNF; ++n) {
            f[n].alloc(loc_box(g, kStageLocs[n]), 1, st);
            r.R[n] = &f[n].v;
        }
    }
    gust_residuals mut()
    {
        return r;
    }
};

Box3 view_box3(const gust_array4* v)
{
    return Box3{v->begin[0], v->begin[1], v->begin[2], v->end[0] - v->begin[0], v->end[1] - v->begin[1],
                v->end[2] - v->begin[2]};
}

int grid_for(int64_t n)
{
    return int(std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, 8192)));
}

double field_scale(const gust_jvp_scales& s, int f)
{
    return f < 3 ? s.velocity : f == 3 ? s.pressure : f == 4 ? s.theta : s.tke;
}

// The continuation's SER law (PseudoTime.cpp:53-61 through gust_ser_cfl).
double ser(double cfl, double r, double cfl0, double cfl_max, double p, double growth)
{
    double out = 0;
    check_rc(gust_ser_cfl(cfl, r, cfl0, cfl_max, p, growth, &out));
    return out;
}

}   // namespace

struct NewtonSolver {
    SimplecStage* stage = nullptr;
    gust_newton_options opt{};
    gust_geometry g{};
    Packing pk{};
    int64_t n = 0;
    hipStream_t st0 = nullptr;
    // constant per tile
    std::unique_ptr<unsigned char, void (*)(unsigned char*)> own{nullptr, [](unsigned char* p) { (void)hipFree(p); }};
    DBuf dir;            // 1 on the Dirichlet rows of u, v, w
    DevField theta_ref;  // theta_ref on the cells (jvp's offset)
    // FGMRES
    DBuf Vb, Zb, w, x, b, tmp, scratch, partial, dots;
    int m = 0;
    // the operator of a step
    DBuf inv, dqp, jvp_p;
    bool has_inv = false;
    DevState dq, qp;     // apply's increment (ghosts filled), jvp's perturbed state
    DevRes R0, Jv;
    int jvp_calls = 0;
    double eps_min = 0, eps_max = 0;
    // the continuation
    DevState qb, qt, dqs;
    DevField dtau;
    gust_continuation_state cst{};
    // the preconditioner (one SIMPLEC iteration on the linearised system)
    DevState q0, d, qn;
    DevRes Rp;
    DevField pcf[7], pss, F0[3], F1[3], dtf[3], A[3], At, Ak, ap[3], dd[3], conv0, conv1, pdtau;
    DevField At1, Ak1;   // stage denominators: theta's and k's rows at the corrected transport
    gust_closure_fields pcfv{};
    double speed = 1;
    bool pc_ready = false;
    // the stage's closure fields at the current state (the continuation's nu)
    double* host = nullptr;   // pinned, for the dots
    int host_cap = 0;

    ~NewtonSolver()
    {
        if (host) { (void)hipHostFree(host); }
    }

    hipStream_t stream(void* s) const { return static_cast<hipStream_t>(s); }

    // ---- vector operations ----
    void dot_many(const double* Vb_, int64_t ldv, int nv, const double* w_, double* out, hipStream_t st)
    {
        mdot_kernel<<<dim3(kDotBlocks, nv), kThreads, 0, st>>>(n, Vb_, ldv, w_, own.get(), partial.p);
        mdot_final_kernel<<<nv, kThreads, 0, st>>>(partial.p, dots.p);
        GUST_HIP(hipGetLastError());
        GUST_HIP(hipMemcpyAsync(host, dots.p, size_t(nv) * sizeof(double), hipMemcpyDeviceToHost, st));
        GUST_HIP(hipStreamSynchronize(st));
        std::memcpy(out, host, size_t(nv) * sizeof(double));
    }
    double dot(const double* a, const double* b_, hipStream_t st)
    {
        double r = 0;
        dot_many(a, 0, 1, b_, &r, st);
        return r;
    }
    double norm2(const double* a, hipStream_t st) { return std::sqrt(dot(a, a, st)); }
    // (max |x|, min x, NaN seen) over x[0, len)
    void range(const double* xv, int64_t len, double out[3], hipStream_t st)
    {
        range_kernel<<<kDotBlocks, kThreads, 0, st>>>(len, xv, partial.p);
        range_final_kernel<<<1, 64, 0, st>>>(partial.p, kDotBlocks, dots.p);
        GUST_HIP(hipGetLastError());
        GUST_HIP(hipMemcpyAsync(host, dots.p, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
        GUST_HIP(hipStreamSynchronize(st));
        std::memcpy(out, host, 3 * sizeof(double));
    }

    PackArgs pack_args(const gust_array4* const* src, const double* scale) const
    {
        PackArgs a{};
        for (int f = 0; f < kNF; ++f) {
            a.src[f] = dview<const double>(src[f]);
            a.box[f] = pk.box[f];
            a.off[f] = pk.off[f];
            a.scale[f] = scale[f];
        }
        return a;
    }
    void pack(const gust_array4* const* src, const double* scale, double* out, hipStream_t st) const
    {
        int64_t most = 1;
        for (int f = 0; f < kNF; ++f) { most = std::max<int64_t>(most, pk.box[f].npts()); }
        pack_kernel<<<dim3(grid_for(most), kNF), kThreads, 0, st>>>(pack_args(src, scale), out);
        GUST_HIP(hipGetLastError());
    }
    void unpack(gust_array4* const* dst, const double* scale, const double* in, hipStream_t st) const
    {
        UnpackArgs a{};
        int64_t most = 1;
        for (int f = 0; f < kNF; ++f) {
            a.dst[f] = dview<double>(dst[f]);
            a.box[f] = pk.box[f];
            a.off[f] = pk.off[f];
            a.scale[f] = scale[f];
            most = std::max<int64_t>(most, pk.box[f].npts());
        }
        unpack_kernel<<<dim3(grid_for(most), kNF), kThreads, 0, st>>>(a, in);
        GUST_HIP(hipGetLastError());
    }
    // whole views (ghosts included) of the stage's layout
    ViewSet views(const gust_state& y, const gust_state* x, const gust_state* z) const
    {
        ViewSet s{};
        for (int f = 0; f < kNF; ++f) {
            s.y[f] = dview<double>(y.f[f]);
            if (x) { s.x[f] = dview<const double>(x->f[f]); }
            if (z) { s.z[f] = dview<const double>(z->f[f]); }
            s.box[f] = view_box3(y.f[f]);
        }
        return s;
    }
    int64_t most_view(const gust_state& y) const
    {
        int64_t most = 1;
        for (int f = 0; f < kNF; ++f) { most = std::max<int64_t>(most, view_box3(y.f[f]).npts()); }
        return most;
    }
    // y = x on every point of y's views (x covers them)
    void copy_state(const gust_state& y, const gust_state& x, hipStream_t st) const
    {
        view_copy_kernel<<<dim3(grid_for(most_view(y)), kNF), kThreads, 0, st>>>(views(y, &x, nullptr));
        GUST_HIP(hipGetLastError());
    }
    void zero_state(const gust_state& y, hipStream_t st) const
    {
        view_zero_kernel<<<dim3(grid_for(most_view(y)), kNF), kThreads, 0, st>>>(views(y, nullptr, nullptr));
        GUST_HIP(hipGetLastError());
    }
    // y += a x on the valid points of the six fields (fields in mask only)
    void saxpy_valid(const gust_state& y, double a, const gust_state& x, int first, int last, hipStream_t st) const
    {
        ViewSet s = views(y, &x, nullptr);
        int64_t most = 1;
        for (int f = 0; f < kNF; ++f) {
            s.box[f] = (f >= first && f <= last) ? pk.box[f] : Box3{0, 0, 0, 0, 0, 0};
            most = std::max<int64_t>(most, s.box[f].npts());
        }
        view_saxpy_kernel<<<dim3(grid_for(most), kNF), kThreads, 0, st>>>(s, a);
        GUST_HIP(hipGetLastError());
    }

    // ---- the stage's pieces ----
    void fill(gust_state* q, gust_array4* ss, hipStream_t st)
    {
        check_rc(gust_fill_ghosts(&g, q, &stage->mesh, &stage->phys, ss, st));
    }
    void fill_h(gust_state* q, hipStream_t st)
    {
        check_rc(gust_fill_ghosts_homogeneous(&g, q, &stage->mesh, &stage->phys, st));
    }
    void residual(const gust_state* q, DevRes& R, gust_closure_fields* cf, gust_array4* ss, hipStream_t st)
    {
        gust_residuals r = R.mut();
        check_rc(gust_residual(&g, q, &stage->mesh, &stage->phys, &r, cf, ss, st));
    }

    // jvp (Source/Residual.cpp:324-396): Jv from q (filled), v (filled), R0 = R(q).
    double jvp(const gust_state& q, const gust_state& v, const gust_residuals& r0, const gust_residuals& jv,
               hipStream_t st)
    {
        const gust_jvp_scales& js = stage->opt.scales;
        double sums[3] = {0, 0, 0};
        for (int f = 0; f < kNF; ++f) {
            double s[3] = {0, 0, 0};
            check_rc(gust_scaled_sums(&g, kStageLocs[f], q.f[f], v.f[f], f == 4 ? &theta_ref.v : nullptr,
                                      field_scale(js, f), s, st));
            for (int c = 0; c < 3; ++c) { sums[c] += s[c]; }
        }
        ViewSet js_set{};
        int64_t most = 1;
        for (int f = 0; f < kNF; ++f) {
            js_set.y[f] = dview<double>(jv.R[f]);
            js_set.x[f] = dview<const double>(r0.R[f]);
            js_set.box[f] = pk.box[f];
            most = std::max<int64_t>(most, pk.box[f].npts());
        }
        if (sums[1] == 0) {
            ViewSet z = js_set;
            view_zero_kernel<<<dim3(grid_for(most), kNF), kThreads, 0, st>>>(z);
            GUST_HIP(hipGetLastError());
            return 0;
        }
        const double mean_q = sums[0] / sums[2];
        const double rms_v = std::sqrt(sums[1] / sums[2]);
        const double eps = opt.fd_eps_scale * (1 + mean_q) / rms_v;
        // q + eps v on every point (ghosts included), refilled
        view_lincomb_kernel<<<dim3(grid_for(most_view(qp.s)), kNF), kThreads, 0, st>>>(views(qp.s, &q, &v), 1.0,
                                                                                       eps);
        GUST_HIP(hipGetLastError());
        fill(&qp.s, stage->wall ? &stage->ss.v : nullptr, st);
        gust_residuals jr = jv;
        check_rc(gust_residual(&g, &qp.s, &stage->mesh, &stage->phys, &jr, nullptr,
                               stage->wall ? &stage->ss.v : nullptr, st));
        jv_finish_kernel<<<dim3(grid_for(most), kNF), kThreads, 0, st>>>(js_set, 1 / eps);
        GUST_HIP(hipGetLastError());
        return eps;
    }

    // the scales S (unknowns) and T (rows)
    void scales(double s[kNF], double t[kNF]) const
    {
        const gust_jvp_scales& js = stage->opt.scales;
        for (int f = 0; f < kNF; ++f) {
            s[f] = field_scale(js, f);
            t[f] = f == 3 ? 1.0 : field_scale(js, f);
        }
    }

    // The operator of a step at q (filled) with R0 = R(q): 1/dtau on the rows.
    void set_operator(const gust_array4* dt, hipStream_t st)
    {
        has_inv = dt != nullptr;
        jvp_calls = 0;
        eps_min = eps_max = 0;
        if (!dt) { return; }
        check_rc(gust_face_dtau(&g, dt, &stage->dtf[0].v, &stage->dtf[1].v, &stage->dtf[2].v, st));
        for (int f = 0; f < kNF; ++f) {
            if (f == 3) {
                fill_kernel<<<grid_for(pk.box[f].npts()), kThreads, 0, st>>>(pk.box[f].npts(), 0.0, inv.p + pk.off[f]);
                continue;
            }
            const gust_array4* src = f < 3 ? &stage->dtf[f].v : dt;
            const bool has_m = f < 3 && !(f < 2 && g.periodic[f]);
            inv_dtau_kernel<<<grid_for(pk.box[f].npts()), kThreads, 0, st>>>(
                pk.box[f], dview<const double>(src), has_m ? dview<const double>(&stage->mask[f].v) : CV{},
                has_m ? 1 : 0, inv.p + pk.off[f]);
        }
        GUST_HIP(hipGetLastError());
    }

    // Ax = T^-1 (D + J) S x (Source/Newton.cpp:282-306).
    void apply(const gust_state& q, const double* xv, double* Ax, hipStream_t st)
    {
        double s[kNF], t[kNF];
        scales(s, t);
        zero_state(dq.s, st);
        unpack(dq.s.f, s, xv, st);
        fill_h(&dq.s, st);
        const double eps = jvp(q, dq.s, R0.r, Jv.r, st);
        if (eps > 0) {
            eps_min = eps_min > 0 ? std::min(eps_min, eps) : eps;
            eps_max = std::max(eps_max, eps);
        }
        ++jvp_calls;
        const double ones[kNF] = {1, 1, 1, 1, 1, 1};
        pack(dq.s.f, ones, dqp.p, st);
        const gust_array4* jr[kNF];
        for (int f = 0; f < kNF; ++f) { jr[f] = Jv.r.R[f]; }
        pack(jr, ones, jvp_p.p, st);
        FieldOf fo{};
        for (int f = 0; f <= kNF; ++f) { fo.off[f] = pk.off[f]; }
        for (int f = 0; f < kNF; ++f) { fo.tinv[f] = 1.0 / t[f]; }
        combine_kernel<<<grid_for(n), kThreads, 0, st>>>(n, fo, has_inv ? inv.p : nullptr, dir.p, dqp.p, jvp_p.p,
                                                         xv, Ax);
        GUST_HIP(hipGetLastError());
    }

    // z = S^-1 M T r (Source/Newton.cpp:308-324).
    void precond(const double* r, double* z, hipStream_t st)
    {
        if (opt.preconditioner == 0) {
            GUST_HIP(hipMemcpyAsync(z, r, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, st));
            return;
        }
        double s[kNF], t[kNF];
        scales(s, t);
        // the hook works in physical units: tmp = T r, then z = S^-1 M tmp
        for (int f = 0; f < kNF; ++f) {
            const int64_t len = pk.off[f + 1] - pk.off[f];
            scale_copy_kernel<<<grid_for(len), kThreads, 0, st>>>(len, t[f], r + pk.off[f], tmp.p + pk.off[f]);
        }
        pc_apply(tmp.p, z, st);
        for (int f = 0; f < kNF; ++f) {
            const int64_t len = pk.off[f + 1] - pk.off[f];
            scale_copy_kernel<<<grid_for(len), kThreads, 0, st>>>(len, 1.0 / s[f], z + pk.off[f], z + pk.off[f]);
        }
        GUST_HIP(hipGetLastError());
    }

    // ---- the SIMPLEC preconditioner (Source/NewtonPreconditioner.cpp) ----
    void pc_setup(const gust_state& q, const gust_array4* dt, hipStream_t st)
    {
        SimplecStage& S = *stage;
        copy_state(q0.s, q, st);
        gust_array4* ssv = S.wall ? &pss.v : nullptr;
        fill(&q0.s, ssv, st);
        const bool closure = S.phys.use_closure != 0;
        if (closure) { residual(&q0.s, Rp, &pcfv, ssv, st); }
        if (dt) {
            ViewSet c{};
            c.y[0] = dview<double>(&pdtau.v);
            c.x[0] = dview<const double>(dt);
            c.box[0] = pk.box[3];
            view_copy_kernel<<<dim3(grid_for(pk.box[3].npts()), 1), kThreads, 0, st>>>(c);
            GUST_HIP(hipGetLastError());
        } else {
            const gust_array4* nu = closure ? pcfv.nu : S.phys.nu;
            check_rc(gust_local_dtau(&g, q0.s.f[0], q0.s.f[1], q0.s.f[2], &S.z, &S.J, nu, opt.start_cfl, &pdtau.v,
                                     st));
        }
        {
            // once per setup: the bound of the sweeps
            PackArgs a{};
            a.src[0] = dview<const double>(&pdtau.v);
            a.box[0] = pk.box[3];
            a.off[0] = 0;
            a.scale[0] = 1.0;
            pack_kernel<<<dim3(grid_for(pk.box[3].npts()), 1), kThreads, 0, st>>>(a, scratch.p);
            GUST_HIP(hipGetLastError());
            double rg[3];
            range(scratch.p, pk.box[3].npts(), rg, st);
            if (rg[2] != 0 || !(rg[1] > 0) || !std::isfinite(rg[0])) {
                GUST_FAIL("gust: SimplecPreconditioner::setup: dtau must be finite and > 0 everywhere");
            }
        }
        check_rc(gust_transport_fluxes(&g, q0.s.f[0], q0.s.f[1], q0.s.f[2], &S.z, &S.ax, &S.ay, &F0[0].v, &F0[1].v,
                                       &F0[2].v, st));
        check_rc(gust_face_dtau(&g, &pdtau.v, &dtf[0].v, &dtf[1].v, &dtf[2].v, st));
        check_rc(gust_convection(GUST_LOC_CELL, &g, q0.s.f[4], &F0[0].v, &F0[1].v, &F0[2].v, &S.J,
                                 S.phys.theta_limiter, S.phys.theta_eps, 0, &conv0.v, st));
        // the velocity scale of apply's directional derivative (at least 1 m/s)
        speed = std::max(1.0, velocity_max(q0.s, st));
        // A1 and the folds at q's transport (the rows of every sweep of apply)
        gust_faces Ff{{&F0[0].v, &F0[1].v, &F0[2].v}};
        const gust_closure_fields* cf = closure ? &pcfv : nullptr;
        const int eqs[3] = {GUST_EQ_U, GUST_EQ_V, GUST_EQ_W};
        for (int c = 0; c < 3; ++c) {
            check_rc(gust_implicit_rows(eqs[c], &g, &Ff, &S.mesh, &S.phys, cf, ssv, &A[c].v, st));
        }
        check_rc(gust_implicit_rows(GUST_EQ_THETA, &g, &Ff, &S.mesh, &S.phys, cf, ssv, &At.v, st));
        check_rc(gust_implicit_rows(GUST_EQ_TKE, &g, &Ff, &S.mesh, &S.phys, cf, ssv, &Ak.v, st));
        pc_ready = true;
    }

    // max |u|, |v|, |w| over their valid points (AMReX's norm0)
    double velocity_max(const gust_state& q, hipStream_t st)
    {
        const double ones[kNF] = {1, 1, 1, 1, 1, 1};
        pack(q.f, ones, scratch.p, st);
        double rg[3];
        range(scratch.p, pk.off[3], rg, st);
        return rg[0];
    }

    void pc_apply(const double* r, double* out, hipStream_t st)
    {
        if (!pc_ready) { GUST_FAIL("gust: SimplecPreconditioner::apply: setup has not been called"); }
        SimplecStage& S = *stage;
        const bool closure = S.phys.use_closure != 0;
        gust_array4* ssv = S.wall ? &pss.v : nullptr;
        const gust_closure_fields* cf = closure ? &pcfv : nullptr;
        // -r in place of the residual
        const double neg[kNF] = {-1, -1, -1, -1, -1, -1};
        gust_array4* rp[kNF];
        for (int f = 0; f < kNF; ++f) { rp[f] = &Rp.f[f].v; }
        unpack(rp, neg, r, st);
        // 1. the momentum predictor on the increment from zero (theta: q's)
        zero_state(d.s, st);
        {
            ViewSet c{};
            c.y[0] = dview<double>(d.s.f[4]);
            c.x[0] = dview<const double>(q0.s.f[4]);
            c.box[0] = view_box3(d.s.f[4]);
            view_copy_kernel<<<dim3(grid_for(c.box[0].npts()), 1), kThreads, 0, st>>>(c);
            GUST_HIP(hipGetLastError());
        }
        gust_faces vel{{d.s.f[0], d.s.f[1], d.s.f[2]}};
        gust_faces Rf{{&Rp.f[0].v, &Rp.f[1].v, &Rp.f[2].v}};
        gust_faces Af{{&A[0].v, &A[1].v, &A[2].v}};
        gust_faces Df{{&dtf[0].v, &dtf[1].v, &dtf[2].v}};
        gust_faces Mf{{&S.mask[0].v, &S.mask[1].v, &S.mask[2].v}};
        gust_faces Pf{{&ap[0].v, &ap[1].v, &ap[2].v}};
        gust_faces Qf{{&dd[0].v, &dd[1].v, &dd[2].v}};
        gust_predictor_options po{};
        po.sweeps = S.opt.sweeps;
        po.rotation_relax = S.opt.rotation_relax;
        po.w_vnn_cap = S.opt.w_vnn_cap;
        po.w_vnn_slope_margin = S.opt.w_vnn_cap_slope_margin;
        {
            const gust_forcing& fc = S.phys.forcing;
            const double pi = 3.14159265358979323846;
            po.coriolis_f = fc.coriolis ? 4 * pi * std::sin(fc.latitude * pi / 180) / fc.rotation_period : 0.0;
        }
        po.gravity = S.phys.forcing.gravity;
        po.theta0 = S.phys.forcing.theta0;
        po.sweep = S.opt.sweep;
        po.sweep_precision = 0;
        const bool buoy = S.phys.forcing.gravity > 0;
        gust_momentum_predictor_result pres{};
        check_rc(gust_momentum_predictor(&g, &vel, d.s.f[4], &S.z, &Rf, &Rp.f[4].v, &Af, buoy ? &At.v : nullptr, &Df,
                                         &Mf, &po, &Pf, &Qf, &pres, st));
        // 2. the pressure correction of du* with r_p as the source
        {
            ViewSet c{};
            c.y[0] = dview<double>(d.s.f[4]);
            c.box[0] = view_box3(d.s.f[4]);
            view_zero_kernel<<<dim3(grid_for(c.box[0].npts()), 1), kThreads, 0, st>>>(c);
            GUST_HIP(hipGetLastError());
        }
        fill_h(&d.s, st);
        // the denominators: diagonal, d = a_P where the stage's d > 0; stage,
        // the predictor's d = a_P - sum a_nb as it left them
        const bool linear = opt.pc_denominator == 1;
        if (linear) {
            for (int c = 0; c < 3; ++c) {
                diag_denominator_kernel<<<grid_for(pk.box[c].npts()), kThreads, 0, st>>>(
                    pk.box[c], dview<double>(&dd[c].v), dview<const double>(&ap[c].v));
            }
            GUST_HIP(hipGetLastError());
        }
        // the source of the correction: r_p (a net source is projected out)
        scale_copy_field_from_packed(r, 3, 1.0, st);
        gust_pressure_correction_result cres{};
        check_rc(gust_pressure_correction(S.solver, &g, d.s.f[0], d.s.f[1], d.s.f[2], d.s.f[3], &S.z, &S.ax, &S.ay,
                                          &S.az, &S.J, &dd[0].v, &dd[1].v, &dd[2].v, S.pbc, S.opt.alpha_p,
                                          &Rp.f[3].v, 1, &cres, st));
        // 3. theta's coupling (Rp's theta still holds -r_theta). diagonal:
        // the directional derivative of theta's convection along du with the
        // step sigma du (sigma = sqrt(eps) U / max|du|), divided by sigma;
        // stage: the difference at u + du itself (sigma = 1)
        double sigma = 1.0;
        if (linear) {
            const double dmax = velocity_max(d.s, st);
            sigma = dmax > 0 ? std::sqrt(std::numeric_limits<double>::epsilon()) * speed / dmax : 1.0;
        }
        copy_state(qn.s, q0.s, st);
        saxpy_valid(qn.s, sigma, d.s, 0, 2, st);
        fill(&qn.s, nullptr, st);
        check_rc(gust_transport_fluxes(&g, qn.s.f[0], qn.s.f[1], qn.s.f[2], &S.z, &S.ax, &S.ay, &F1[0].v, &F1[1].v,
                                       &F1[2].v, st));
        check_rc(gust_convection(GUST_LOC_CELL, &g, qn.s.f[4], &F1[0].v, &F1[1].v, &F1[2].v, &S.J,
                                 S.phys.theta_limiter, S.phys.theta_eps, 0, &conv1.v, st));
        box_saxpy_kernel<<<grid_for(pk.box[4].npts()), kThreads, 0, st>>>(
            pk.box[4], dview<double>(&Rp.f[4].v), 1.0 / sigma, dview<const double>(&conv1.v),
            dview<const double>(&conv0.v));
        GUST_HIP(hipGetLastError());
        // 4. the theta and k steps on the increments, no floor. diagonal:
        // with q's transport (the rows of setup); stage: with the corrected
        // transport F1 (rows rebuilt from it)
        gust_array4* rows_t = &At.v;
        gust_array4* rows_k = &Ak.v;
        if (!linear) {
            gust_faces F1f{{&F1[0].v, &F1[1].v, &F1[2].v}};
            check_rc(gust_implicit_rows(GUST_EQ_THETA, &g, &F1f, &S.mesh, &S.phys, cf, ssv, &At1.v, st));
            check_rc(gust_implicit_rows(GUST_EQ_TKE, &g, &F1f, &S.mesh, &S.phys, cf, ssv, &Ak1.v, st));
            rows_t = &At1.v;
            rows_k = &Ak1.v;
        }
        gust_scalar_sweep_result rt{}, rk{};
        check_rc(gust_scalar_sweep(&g, GUST_EQ_THETA, d.s.f[4], &Rp.f[4].v, rows_t, &pdtau.v, S.opt.sweeps, 0.0, 0,
                                   &S.opt.sweep, &rt, st));
        check_rc(gust_scalar_sweep(&g, GUST_EQ_TKE, d.s.f[5], &Rp.f[5].v, rows_k, &pdtau.v, S.opt.sweeps, 0.0, 0,
                                   &S.opt.sweep, &rk, st));
        // 5. dq: the increment's valid points; the Dirichlet rows of u, v, w
        // take r; the periodic seam copies take their owner's
        const double ones[kNF] = {1, 1, 1, 1, 1, 1};
        pack(d.s.f, ones, out, st);
        dirichlet_take_kernel<<<grid_for(pk.off[3]), kThreads, 0, st>>>(pk.off[3], dir.p, r, out);
        GUST_HIP(hipGetLastError());
        seam_override(out, st);
    }

    // Rp_f = a * (packed r's field f) on the valid points
    void scale_copy_field_from_packed(const double* r, int f, double a, hipStream_t st)
    {
        UnpackArgs u{};
        u.dst[0] = dview<double>(&Rp.f[f].v);
        u.box[0] = pk.box[f];
        u.off[0] = pk.off[f];
        u.scale[0] = a;
        unpack_kernel<<<dim3(grid_for(pk.box[f].npts()), 1), kThreads, 0, st>>>(u, r);
        GUST_HIP(hipGetLastError());
    }

    void seam_override(double* xv, hipStream_t st)
    {
        for (int f = 0; f < 2; ++f) {
            if (!g.periodic[f]) { continue; }
            const Box3& b = pk.box[f];
            const int64_t cnt = f == 0 ? int64_t(b.n1) * b.n2 : int64_t(b.n0) * b.n2;
            seam_override_kernel<<<grid_for(cnt), kThreads, 0, st>>>(xv, pk.off[f], b.n0, b.n1, b.n2, f);
        }
        GUST_HIP(hipGetLastError());
    }

    // ---- FGMRES (flexible, right preconditioned, restarted) ----
    // Solves A x = b from x = 0; returns the iteration count, status and the
    // final residual estimate.
    void fgmres(const gust_state& q, double bnorm, gust_newton_step_info& info, hipStream_t st)
    {
        const int64_t ld = n;
        fill_kernel<<<grid_for(n), kThreads, 0, st>>>(n, 0.0, x.p);
        GUST_HIP(hipGetLastError());
        info.gmres_iterations = 0;
        info.gmres_status = 0;
        info.krylov_residual = 0;
        if (!(bnorm > 0) || !std::isfinite(bnorm)) { return; }
        const double target = opt.krylov_rtol * bnorm;
        std::vector<double> H(size_t(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m), h(m + 1), h2(m + 1);
        double beta = bnorm;
        // r0 = b
        GUST_HIP(hipMemcpyAsync(w.p, b.p, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, st));
        int iters = 0;
        double resid = beta;
        bool done = false;
        while (!done) {
            scale_copy_kernel<<<grid_for(n), kThreads, 0, st>>>(n, 1.0 / beta, w.p, Vb.p);
            GUST_HIP(hipGetLastError());
            std::fill(gv.begin(), gv.end(), 0.0);
            gv[0] = beta;
            int j = 0;
            for (; j < m; ++j) {
                double* vj = Vb.p + int64_t(j) * ld;
                double* zj = Zb.p + int64_t(j) * ld;
                precond(vj, zj, st);
                apply(q, zj, w.p, st);
                // classical Gram-Schmidt with one reorthogonalisation
                dot_many(Vb.p, ld, j + 1, w.p, h.data(), st);
                GUST_HIP(hipMemcpyAsync(dots.p, h.data(), size_t(j + 1) * sizeof(double), hipMemcpyHostToDevice, st));
                multi_axpy_kernel<<<grid_for(n), kThreads, 0, st>>>(n, j + 1, Vb.p, ld, dots.p, -1.0, w.p);
                dot_many(Vb.p, ld, j + 1, w.p, h2.data(), st);
                GUST_HIP(hipMemcpyAsync(dots.p, h2.data(), size_t(j + 1) * sizeof(double), hipMemcpyHostToDevice, st));
                multi_axpy_kernel<<<grid_for(n), kThreads, 0, st>>>(n, j + 1, Vb.p, ld, dots.p, -1.0, w.p);
                GUST_HIP(hipGetLastError());
                for (int i = 0; i <= j; ++i) { h[i] += h2[i]; }
                const double hn = norm2(w.p, st);
                for (int i = 0; i <= j; ++i) { H[size_t(i) * m + j] = h[i]; }
                H[size_t(j + 1) * m + j] = hn;
                // the earlier rotations on the new column, then a new one
                for (int i = 0; i < j; ++i) {
                    const double a = H[size_t(i) * m + j], c = H[size_t(i + 1) * m + j];
                    H[size_t(i) * m + j] = cs[i] * a + sn[i] * c;
                    H[size_t(i + 1) * m + j] = -sn[i] * a + cs[i] * c;
                }
                const double a = H[size_t(j) * m + j], c = H[size_t(j + 1) * m + j];
                const double rr = std::hypot(a, c);
                cs[j] = rr > 0 ? a / rr : 1.0;
                sn[j] = rr > 0 ? c / rr : 0.0;
                H[size_t(j) * m + j] = rr;
                H[size_t(j + 1) * m + j] = 0.0;
                gv[j + 1] = -sn[j] * gv[j];
                gv[j] = cs[j] * gv[j];
                resid = std::abs(gv[j + 1]);
                ++iters;
                const bool breakdown = !(hn > 0) || !std::isfinite(hn);
                if (!breakdown && j + 1 < m + 1) {
                    scale_copy_kernel<<<grid_for(n), kThreads, 0, st>>>(n, 1.0 / hn, w.p,
                                                                        Vb.p + int64_t(j + 1) * ld);
                    GUST_HIP(hipGetLastError());
                }
                if (resid <= target || iters >= opt.gmres_max_iters || breakdown || !std::isfinite(resid)) {
                    ++j;
                    done = true;
                    break;
                }
            }
            const int kk = std::min(j, m);
            // y = H^-1 g (upper triangular), x += Z y
            for (int i = kk - 1; i >= 0; --i) {
                double s = gv[i];
                for (int l = i + 1; l < kk; ++l) { s -= H[size_t(i) * m + l] * y[l]; }
                const double dgl = H[size_t(i) * m + i];
                y[i] = dgl != 0 ? s / dgl : 0.0;
            }
            GUST_HIP(hipMemcpyAsync(dots.p, y.data(), size_t(kk) * sizeof(double), hipMemcpyHostToDevice, st));
            multi_axpy_kernel<<<grid_for(n), kThreads, 0, st>>>(n, kk, Zb.p, ld, dots.p, 1.0, x.p);
            GUST_HIP(hipGetLastError());
            if (done) { break; }
            // restart: r = b - A x
            apply(q, x.p, w.p, st);
            axpby_kernel<<<grid_for(n), kThreads, 0, st>>>(n, 1.0, b.p, -1.0, w.p);
            GUST_HIP(hipGetLastError());
            beta = norm2(w.p, st);
            resid = beta;
            if (!(beta > target) || !std::isfinite(beta)) { break; }
        }
        info.gmres_iterations = iters;
        info.gmres_status = resid <= target ? 0 : 1;
        info.krylov_residual = resid;
    }

    // fill(q), R0, the solve; the update S x goes to use(x) when usable
    // (Source/Newton.cpp:455-499).
    template <class Use>
    gust_newton_step_info solve(gust_state* q, const gust_array4* dt, Use use, hipStream_t st)
    {
        SimplecStage& S = *stage;
        fill(q, S.wall ? &S.ss.v : nullptr, st);
        residual(q, R0, nullptr, S.wall ? &S.ss.v : nullptr, st);
        set_operator(dt, st);
        if (opt.preconditioner == 1) { pc_setup(*q, dt, st); }
        double s[kNF], t[kNF];
        scales(s, t);
        double mt[kNF];
        for (int f = 0; f < kNF; ++f) { mt[f] = -1.0 / t[f]; }
        const gust_array4* rr[kNF];
        for (int f = 0; f < kNF; ++f) { rr[f] = R0.r.R[f]; }
        pack(rr, mt, b.p, st);
        gust_newton_step_info info{};
        info.residual_norm = norm2(b.p, st);
        fgmres(*q, info.residual_norm, info, st);
        info.krylov_relative = info.residual_norm > 0 ? info.krylov_residual / info.residual_norm : 0.0;
        if (!(info.residual_norm > 0)) { info.krylov_residual = 0; }
        info.eps_min = eps_min;
        info.eps_max = eps_max;
        info.jvp_calls = jvp_calls;
        info.applied = std::isfinite(info.residual_norm) && std::isfinite(info.krylov_relative) &&
                       info.krylov_relative <= 1;
        if (info.applied) { use(s); }
        GUST_HIP(hipStreamSynchronize(st));
        return info;
    }

    double residual_norm(gust_state* q, hipStream_t st)
    {
        SimplecStage& S = *stage;
        fill(q, S.wall ? &S.ss.v : nullptr, st);
        residual(q, R0, nullptr, S.wall ? &S.ss.v : nullptr, st);
        double s[kNF], t[kNF];
        scales(s, t);
        double mt[kNF];
        for (int f = 0; f < kNF; ++f) { mt[f] = -1.0 / t[f]; }
        const gust_array4* rr[kNF];
        for (int f = 0; f < kNF; ++f) { rr[f] = R0.r.R[f]; }
        pack(rr, mt, b.p, st);
        return norm2(b.p, st);
    }

    // to += S x on the valid points (NewtonOperator::add_update)
    void add_update(const gust_state& to, const double s[kNF], hipStream_t st)
    {
        zero_state(qt.s, st);
        unpack(qt.s.f, s, x.p, st);
        saxpy_valid(to, 1.0, qt.s, 0, kNF - 1, st);
    }

    // ---- the continuation (Source/Continuation.cpp:339-427) ----
    int continuation_step(gust_state* q, gust_continuation_record* last, hipStream_t st)
    {
        SimplecStage& S = *stage;
        for (int attempt = 0; attempt <= opt.max_retries; ++attempt) {
            gust_continuation_record rec{};
            rec.step = cst.steps;
            rec.attempt = attempt;
            rec.cfl = cst.cfl;
            rec.infinite = cst.infinite;
            copy_state(qb.s, *q, st);
            fill(&qb.s, S.wall ? &S.ss.v : nullptr, st);
            const gust_array4* dt = nullptr;
            if (!cst.infinite) {
                const gust_array4* nu = S.phys.use_closure ? &S.cf[0].v : S.phys.nu;
                check_rc(gust_local_dtau(&g, qb.s.f[0], qb.s.f[1], qb.s.f[2], &S.z, &S.J, nu, cst.cfl, &dtau.v, st));
                dt = &dtau.v;
            }
            zero_state(dqs.s, st);
            const gust_newton_step_info info =
                solve(&qb.s, dt, [&](const double* s) { add_update(dqs.s, s, st); }, st);
            rec.residual = info.residual_norm;
            rec.residual_new = info.residual_norm;
            rec.gmres_iterations = info.gmres_iterations;
            rec.gmres_status = info.gmres_status;
            rec.krylov_relative = info.krylov_relative;
            const bool finite_r = std::isfinite(info.residual_norm);
            if (cst.r0 == 0 && finite_r) { cst.r0 = info.residual_norm; }
            if (finite_r && info.residual_norm == 0) {
                rec.outcome = 0;
                if (last) { *last = rec; }
                accept(0.0);
                return 0;
            }
            if (info.applied) {
                const double r = info.residual_norm;
                auto phi = [&](double lambda) {
                    copy_state(qt.s, qb.s, st);
                    saxpy_valid(qt.s, lambda, dqs.s, 0, kNF - 1, st);
                    const Box3 cb = pk.box[5];
                    floor_kernel<<<grid_for(cb.npts()), kThreads, 0, st>>>(cb, dview<double>(qt.s.f[5]),
                                                                           S.opt.tke_floor);
                    GUST_HIP(hipGetLastError());
                    return residual_norm(&qt.s, st);
                };
                // Armijo backtracking (Continuation.cpp:179-208)
                const double growth = cst.infinite ? 1.0 : opt.ptc_growth;
                double lambda = 1, r_full = 0, res = 0;
                int backtracks = 0;
                bool accepted = false, by_growth = false;
                while (true) {
                    res = phi(lambda);
                    if (lambda == 1) { r_full = res; }
                    if (std::isfinite(res) && res <= (1 - opt.armijo_c * lambda) * r) {
                        accepted = true;
                        break;
                    }
                    const double next = opt.backtrack * lambda;
                    if (!(next >= opt.lambda_min)) { break; }
                    lambda = next;
                    ++backtracks;
                }
                if (!accepted && growth > 1 && std::isfinite(r_full) && r_full <= growth * r) {
                    lambda = 1;
                    res = phi(lambda);
                    accepted = true;
                    by_growth = true;
                }
                rec.lambda = lambda;
                rec.backtracks = backtracks;
                rec.residual_new = res;
                rec.by_growth = by_growth ? 1 : 0;
                if (accepted) {
                    copy_state(*q, qt.s, st);
                    GUST_HIP(hipStreamSynchronize(st));
                    rec.outcome = 0;
                    if (last) { *last = rec; }
                    accept(res);
                    return 0;
                }
                rec.outcome = 2;
            } else {
                rec.outcome = 1;
            }
            if (last) { *last = rec; }
            // a failed attempt: CFL cut (from infinite dtau: cfl_cut cfl_max)
            cst.cfl = std::max(opt.cfl_cut * (cst.infinite ? opt.cfl_max : cst.cfl),
                               std::numeric_limits<double>::min());
            cst.infinite = 0;
        }
        return 1;
    }

    // Continuation.cpp:210-219 with the Newton ramp (cfl0 = cfl_max = the cap)
    void accept(double r_new)
    {
        cst.steps += 1;
        cst.r_last = r_new;
        const double r = cst.r0 > 0 ? r_new / cst.r0 : 0.0;
        const double from = cst.infinite ? opt.cfl_max : cst.cfl;
        cst.cfl = ser(from, r, opt.cfl_max, opt.cfl_max, opt.ser_exponent, opt.ser_max_growth);
        cst.infinite = (opt.dtau_inf_at_cap == 1 && cst.cfl >= opt.cfl_max) ? 1 : 0;
    }
};

namespace {

// The Newton stage (Source/NewtonStage.cpp:91-207): one continuation step per
// iteration and the run's norms at the new state; the hand-backs.
struct NewtonStageRun {
    NewtonSolver& N;
    SimplecStage& S;
    const gust_driver_options& dopt;
    const gust_switch_options& sw;
    bool request = false, handed_back = false, converged = false;
    int failures = 0, failed_visits = 0, idle = 0;
    std::string reason;
    static constexpr int kMaxFailedVisits = 3;

    // the run's residual at q with the closure's live fields, its norms
    void run_norms(gust_state* q, hipStream_t st)
    {
        gust_array4* ssv = S.wall ? &S.ss.v : nullptr;
        check_rc(gust_fill_ghosts(&S.g, q, &S.mesh, &S.phys, ssv, st));
        gust_residuals R = S.residuals();
        check_rc(gust_residual(&S.g, q, &S.mesh, &S.phys, &R, S.phys.use_closure ? &S.cfv : nullptr, ssv, st));
        check_rc(gust_stage_norms(&S.g, &R, q->f[0], q->f[1], q->f[2], q->f[5], ssv, &S.opt.scales, S.opt.height,
                                  S.res, S.mon, &S.nmon, st));
    }

    void init(gust_state* q, hipStream_t st)
    {
        request = false;
        reason.clear();
        failures = 0;
        check_rc(gust_fill_ghosts(&S.g, q, &S.mesh, &S.phys, S.wall ? &S.ss.v : nullptr, st));
        // a fresh continuation at the SIMPLEC cap, ramping to the Newton cap
        N.cst = gust_continuation_state{};
        N.cst.cfl = N.opt.start_cfl;
        idle = 0;
        run_norms(q, st);   // the closure's fields at q: the continuation's diffusivity
    }

    void iterate(gust_state* q, hipStream_t st)
    {
        request = false;
        const gust_continuation_state before = N.cst;
        const double drop = std::pow(10.0, -dopt.criterion.residual_drop_orders);
        if (before.steps > 0 && before.r0 > 0 && before.r_last <= drop * before.r0) {
            if (++idle > dopt.criterion.monitor_window) {
                request = true;
                handed_back = true;
                reason = "Newton converged but the steady criterion did not hold for steady.monitor_window "
                         "iterations";
                run_norms(q, st);
                return;
            }
            converged = true;
            run_norms(q, st);
            return;
        }
        converged = false;
        const int status = N.continuation_step(q, nullptr, st);
        failures = status == 1 ? failures + 1 : 0;
        run_norms(q, st);
        if (failures >= sw.newton_fallback_failures) {
            request = true;
            ++failed_visits;
            reason = std::to_string(failures) + " consecutive failed Newton steps (steady.newton_fallback_failures " +
                     std::to_string(sw.newton_fallback_failures) + ")";
            if (failed_visits >= kMaxFailedVisits) {
                handed_back = true;
                reason += "; " + std::to_string(failed_visits) + " failed visits: no more switches to Newton";
            }
        }
        // the stage solves the run's own equations (no limiter swap, no
        // frozen switches): no final SIMPLEC pass is needed
    }

    void leave()
    {
        request = false;
        failures = 0;
    }
};

// q's six fields must cover their valid boxes grown by two (the stage's
// layout: every copy of a state here reads and writes that box).
void check_state2(const char* op, const gust_geometry& g, const gust_state* q, const char* what)
{
    if (!q) { GUST_FAIL("gust: %s: %s is null", op, what); }
    for (int f = 0; f < kNF; ++f) {
        if (!covers(q->f[f], grow(loc_box(g, kStageLocs[f]), 2, 2, 2), 1)) {
            GUST_FAIL("gust: %s: %s needs two ghost layers on every field", op, what);
        }
    }
}

NewtonSolver* as_n(gust_newton* n) { return reinterpret_cast<NewtonSolver*>(n); }
const NewtonSolver* as_n(const gust_newton* n) { return reinterpret_cast<const NewtonSolver*>(n); }
SimplecStage* as_s(gust_simplec* s) { return reinterpret_cast<SimplecStage*>(s); }

std::vector<std::string> newton_option_errors(const gust_newton_options& o)
{
    std::vector<std::string> err;
    auto num = [](double x) {
        char b[40];
        std::snprintf(b, sizeof(b), "%.17g", x);
        return std::string(b);
    };
    if (o.gmres_restart < 5 || o.gmres_restart > 200) {
        err.push_back("steady.newton_gmres_restart = " + std::to_string(o.gmres_restart) +
                      ": must be an integer >= 5 and <= 200");
    }
    if (o.gmres_max_iters < 1) {
        err.push_back("steady.newton_gmres_max_iters = " + std::to_string(o.gmres_max_iters) +
                      ": must be an integer >= 1");
    }
    const double e_min = std::pow(std::numeric_limits<double>::epsilon(), 0.75);
    if (std::isnan(o.fd_eps_scale) || !(o.fd_eps_scale >= e_min) || !(o.fd_eps_scale <= 1.0e-2)) {
        err.push_back("steady.newton_fd_eps_scale = " + num(o.fd_eps_scale) + ": must be >= " + num(e_min) +
                      " (machine epsilon^(3/4)) and <= 1e-2");
    }
    if (std::isnan(o.krylov_rtol) || !(o.krylov_rtol > 0) || !(o.krylov_rtol < 1)) {
        err.push_back("steady.newton_krylov_rtol = " + num(o.krylov_rtol) + ": must be > 0 and < 1");
    }
    if (o.preconditioner != 0 && o.preconditioner != 1) {
        err.push_back("steady.newton_preconditioner must be identity or simplec");
    }
    if (o.pc_denominator != 0 && o.pc_denominator != 1) {
        err.push_back("the preconditioner's pressure denominator must be stage (0) or diagonal (1)");
    }
    if (std::isnan(o.cfl_max) || !(o.cfl_max > 0) || !std::isfinite(o.cfl_max)) {
        err.push_back("steady.newton_cfl_max = " + num(o.cfl_max) + ": must be finite and > 0");
    }
    if (std::isnan(o.start_cfl) || !(o.start_cfl > 0) || !std::isfinite(o.start_cfl) || o.start_cfl > o.cfl_max) {
        err.push_back("steady.cfl_max = " + num(o.start_cfl) + ": must be finite, > 0 and <= steady.newton_cfl_max");
    }
    if (std::isnan(o.armijo_c) || !(o.armijo_c > 0) || !(o.armijo_c < 0.5)) {
        err.push_back("steady.newton_armijo_c = " + num(o.armijo_c) + ": must be > 0 and < 0.5");
    }
    if (std::isnan(o.backtrack) || !(o.backtrack > 0) || !(o.backtrack < 1)) {
        err.push_back("steady.newton_backtrack = " + num(o.backtrack) + ": must be > 0 and < 1");
    }
    if (std::isnan(o.lambda_min) || !(o.lambda_min > 0) || !(o.lambda_min < 1)) {
        err.push_back("steady.newton_lambda_min = " + num(o.lambda_min) + ": must be > 0 and < 1");
    }
    if (std::isnan(o.cfl_cut) || !(o.cfl_cut > 0) || !(o.cfl_cut < 1)) {
        err.push_back("steady.newton_cfl_cut = " + num(o.cfl_cut) + ": must be > 0 and < 1");
    }
    if (o.max_retries < 1) { err.push_back("steady.newton_max_retries must be >= 1"); }
    if (o.dtau_inf_at_cap != 0 && o.dtau_inf_at_cap != 1) {
        err.push_back("steady.newton_dtau_inf_at_cap must be 0 or 1");
    }
    if (std::isnan(o.ptc_growth) || !(o.ptc_growth >= 1) || !std::isfinite(o.ptc_growth)) {
        err.push_back("steady.newton_ptc_growth = " + num(o.ptc_growth) + ": must be finite and >= 1");
    }
    if (std::isnan(o.ser_exponent) || !(o.ser_exponent > 0) || std::isnan(o.ser_max_growth) ||
        !(o.ser_max_growth >= 1)) {
        err.push_back("steady.ser_exponent must be > 0 and steady.ser_max_growth >= 1");
    }
    return err;
}

void fail_on(const char* op, const std::vector<std::string>& err)
{
    if (err.empty()) { return; }
    std::string all = std::string("gust: ") + op + ": invalid inputs:";
    for (const auto& e : err) { all += "\n  " + e; }
    GUST_FAIL("%s", all.c_str());
}

}   // namespace

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_newton_default_options(gust_newton_options* o)
{
    return guarded([&] {
        if (!o) { GUST_FAIL("gust: newton_default_options: null options"); }
        *o = gust_newton_options{};
        o->gmres_restart = 30;
        o->gmres_max_iters = 2000;
        o->fd_eps_scale = std::sqrt(std::numeric_limits<double>::epsilon());
        o->krylov_rtol = 1.0e-2;
        o->preconditioner = 1;
        o->cfl_max = 128;
        o->start_cfl = 10;
        o->ser_exponent = 1;
        o->ser_max_growth = 2;
        o->armijo_c = 1.0e-4;
        o->backtrack = 0.5;
        o->lambda_min = 1.0e-2;
        o->cfl_cut = 0.25;
        o->max_retries = 4;
        o->dtau_inf_at_cap = 1;
        o->ptc_growth = 1;
        o->pc_denominator = 1;
    });
}

int gust_check_newton_options(const gust_newton_options* o, char* buf, int64_t cap, int32_t* count)
{
    return guarded([&] {
        if (!o) { GUST_FAIL("gust: check_newton_options: null options"); }
        const std::vector<std::string> err = newton_option_errors(*o);
        if (count) { *count = int32_t(err.size()); }
        if (buf && cap > 0) {
            std::string all;
            for (const auto& e : err) { all += e + "\n"; }
            std::snprintf(buf, size_t(cap), "%s", all.c_str());
        }
    });
}

int gust_newton_create(gust_simplec* stage, const gust_newton_options* opt, gust_newton** out, void* stream)
{
    return guarded([&] {
        const char* op = "NewtonKrylov";
        if (!stage || !opt || !out) { GUST_FAIL("gust: %s: null argument", op); }
        *out = nullptr;
        fail_on(op, newton_option_errors(*opt));
        SimplecStage& S = *as_s(stage);
        hipStream_t st = static_cast<hipStream_t>(stream);
        auto nw = std::make_unique<NewtonSolver>();
        nw->stage = &S;
        nw->opt = *opt;
        nw->g = S.g;
        const gust_geometry& g = S.g;
        int64_t off = 0;
        for (int f = 0; f < kNF; ++f) {
            nw->pk.box[f] = box3(loc_box(g, kStageLocs[f]));
            nw->pk.off[f] = off;
            off += nw->pk.box[f].npts();
        }
        nw->pk.off[kNF] = off;
        nw->n = off;
        const int64_t n = off;
        // the owner mask: the hi copy of a periodic seam face is not counted
        {
            std::vector<unsigned char> own(size_t(n), 1);
            for (int f = 0; f < 2; ++f) {
                if (!g.periodic[f]) { continue; }
                const Box3& b = nw->pk.box[f];
                for (int k = 0; k < b.n2; ++k)
                    for (int j = 0; j < b.n1; ++j)
                        for (int i = 0; i < b.n0; ++i) {
                            const bool hi = f == 0 ? i == b.n0 - 1 : j == b.n1 - 1;
                            if (hi) { own[size_t(nw->pk.off[f] + (int64_t(k) * b.n1 + j) * b.n0 + i)] = 0; }
                        }
            }
            unsigned char* p = nullptr;
            GUST_HIP(hipMalloc(&p, size_t(n)));
            nw->own.reset(p);
            GUST_HIP(hipMemcpyAsync(p, own.data(), size_t(n), hipMemcpyHostToDevice, st));
            GUST_HIP(hipStreamSynchronize(st));
        }
        // the Dirichlet rows of u, v, w: the stage's masks, packed
        nw->dir.alloc(n);
        fill_kernel<<<grid_for(n), kThreads, 0, st>>>(n, 0.0, nw->dir.p);
        GUST_HIP(hipGetLastError());
        {
            const gust_array4* m[kNF] = {&S.mask[0].v, &S.mask[1].v, &S.mask[2].v, &S.mask[0].v, &S.mask[0].v,
                                         &S.mask[0].v};
            const double sc[kNF] = {1, 1, 1, 0, 0, 0};
            PackArgs a = nw->pack_args(m, sc);
            for (int f = 3; f < kNF; ++f) { a.box[f] = Box3{0, 0, 0, 0, 0, 0}; }
            int64_t most = 1;
            for (int f = 0; f < 3; ++f) { most = std::max<int64_t>(most, nw->pk.box[f].npts()); }
            pack_kernel<<<dim3(grid_for(most), kNF), kThreads, 0, st>>>(a, nw->dir.p);
            GUST_HIP(hipGetLastError());
        }
        // theta_ref on the cells (jvp's offset of theta)
        {
            const IBox cells = loc_box(g, GUST_LOC_CELL);
            nw->theta_ref.alloc(cells, 1, st);
            const heights::Sounding snd = heights::sounding(S.phys.n_theta_ref, S.phys.theta_ref_z, S.phys.theta_ref);
            const Box3 tb = box3(cells);
            heights::theta_ref_kernel<<<blocks_for(tb.npts()), kThreads, 0, st>>>(tb, dview<const double>(&S.z), snd,
                                                                                  dview<double>(&nw->theta_ref.v));
            GUST_HIP(hipGetLastError());
        }
        // FGMRES and the operator
        nw->m = opt->gmres_restart;
        nw->Vb.alloc(n * (nw->m + 1));
        nw->Zb.alloc(n * nw->m);
        for (DBuf* b : {&nw->w, &nw->x, &nw->b, &nw->tmp, &nw->scratch, &nw->inv, &nw->dqp, &nw->jvp_p}) {
            b->alloc(n);
        }
        nw->partial.alloc(int64_t(kDotBlocks) * std::max(nw->m + 1, 3));
        nw->dots.alloc(std::max(nw->m + 1, 3));
        GUST_HIP(hipHostMalloc(&nw->host, size_t(std::max(nw->m + 1, 3)) * sizeof(double)));
        for (DevState* s : {&nw->dq, &nw->qp, &nw->qb, &nw->qt, &nw->dqs}) { s->alloc(g, st); }
        nw->R0.alloc(g, st);
        nw->Jv.alloc(g, st);
        nw->dtau.alloc(loc_box(g, GUST_LOC_CELL), 1, st);
        nw->cst.cfl = opt->start_cfl;
        // the preconditioner's fields
        if (opt->preconditioner == 1) {
            for (DevState* s : {&nw->q0, &nw->d, &nw->qn}) { s->alloc(g, st); }
            const IBox cells = loc_box(g, GUST_LOC_CELL);
            for (int c = 0; c < 3; ++c) { nw->pcf[c].alloc(grow(cells, 1, 1, 1), 1, st); }
            for (int c = 3; c < 7; ++c) { nw->pcf[c].alloc(cells, 1, st); }
            nw->pcfv = gust_closure_fields{&nw->pcf[0].v, &nw->pcf[1].v, &nw->pcf[2].v, &nw->pcf[3].v,
                                           &nw->pcf[4].v, &nw->pcf[5].v, &nw->pcf[6].v};
            IBox slab = grow(cells, 1, 1, 0);
            slab.hi[2] = slab.lo[2];
            nw->pss.alloc(slab, GUST_SS_N, st);
            for (int c = 0; c < 3; ++c) {
                const IBox fb = loc_box(g, 1 << c);
                nw->F0[c].alloc(grow(fb, 1, 1, 1), 1, st);
                nw->F1[c].alloc(grow(fb, 1, 1, 1), 1, st);
                nw->dtf[c].alloc(fb, 1, st);
                nw->A[c].alloc(fb, GUST_N_COEFS, st);
                nw->ap[c].alloc(fb, 1, st);
                nw->dd[c].alloc(fb, 1, st);
            }
            nw->At.alloc(cells, GUST_N_COEFS, st);
            nw->Ak.alloc(cells, GUST_N_COEFS, st);
            if (opt->pc_denominator == 0) {
                nw->At1.alloc(cells, GUST_N_COEFS, st);
                nw->Ak1.alloc(cells, GUST_N_COEFS, st);
            }
            nw->conv0.alloc(cells, 1, st);
            nw->conv1.alloc(cells, 1, st);
            nw->pdtau.alloc(cells, 1, st);
        }
        nw->Rp.alloc(g, st);
        GUST_HIP(hipStreamSynchronize(st));
        *out = reinterpret_cast<gust_newton*>(nw.release());
    });
}

void gust_newton_destroy(gust_newton* n)
{
    delete as_n(n);
}

int gust_jvp(gust_simplec* stage, const gust_state* q, const gust_state* v, const gust_residuals* R0,
             gust_residuals* Jv, double* eps, void* stream)
{
    return guarded([&] {
        if (!stage || !q || !v || !R0 || !Jv || !eps) { GUST_FAIL("gust: jvp: null argument"); }
        const gust_geometry& g = as_s(stage)->g;
        check_state2("jvp", g, q, "q");
        check_state2("jvp", g, v, "v");
        for (int f = 0; f < kNF; ++f) {
            if (!covers(R0->R[f], loc_box(g, kStageLocs[f]), 1) || !covers(Jv->R[f], loc_box(g, kStageLocs[f]), 1)) {
                GUST_FAIL("gust: jvp: R0 and Jv must cover every field's valid points");
            }
        }
        hipStream_t st = static_cast<hipStream_t>(stream);
        // a solver of the smallest restart for its work state
        gust_newton_options o{};
        check_rc(gust_newton_default_options(&o));
        o.gmres_restart = 5;
        o.preconditioner = 0;
        o.fd_eps_scale = as_s(stage)->opt.scales.step;
        NewtonSolver* nw = nullptr;
        check_rc(gust_newton_create(stage, &o, reinterpret_cast<gust_newton**>(&nw), stream));
        std::unique_ptr<NewtonSolver> hold(nw);
        *eps = nw->jvp(*q, *v, *R0, *Jv, st);
        GUST_HIP(hipStreamSynchronize(st));
    });
}

int gust_newton_step(gust_newton* n, gust_state* q, const gust_array4* dtau, gust_newton_step_info* info,
                     void* stream)
{
    return guarded([&] {
        if (!n || !q) { GUST_FAIL("gust: NewtonKrylov::step: null argument"); }
        NewtonSolver& N = *as_n(n);
        check_state2("NewtonKrylov::step", N.g, q, "q");
        if (dtau && !covers(dtau, loc_box(N.g, GUST_LOC_CELL), 1)) {
            GUST_FAIL("gust: NewtonKrylov::step: dtau must cover the valid cells");
        }
        hipStream_t st = static_cast<hipStream_t>(stream);
        const gust_newton_step_info r =
            N.solve(q, dtau, [&](const double* s) { N.add_update(*q, s, st); }, st);
        N.fill(q, N.stage->wall ? &N.stage->ss.v : nullptr, st);
        GUST_HIP(hipStreamSynchronize(st));
        if (info) { *info = r; }
    });
}

int gust_newton_residual_norm(gust_newton* n, gust_state* q, double* r, void* stream)
{
    return guarded([&] {
        if (!n || !q || !r) { GUST_FAIL("gust: NewtonKrylov::residual_norm: null argument"); }
        check_state2("NewtonKrylov::residual_norm", as_n(n)->g, q, "q");
        *r = as_n(n)->residual_norm(q, static_cast<hipStream_t>(stream));
    });
}

int gust_newton_pc_setup(gust_newton* n, const gust_state* q, const gust_array4* dtau, void* stream)
{
    return guarded([&] {
        if (!n || !q) { GUST_FAIL("gust: SimplecPreconditioner::setup: null argument"); }
        NewtonSolver& N = *as_n(n);
        if (N.opt.preconditioner != 1) {
            GUST_FAIL("gust: SimplecPreconditioner::setup: the Newton stage has the identity preconditioner");
        }
        check_state2("SimplecPreconditioner::setup", N.g, q, "q");
        if (dtau && !covers(dtau, loc_box(N.g, GUST_LOC_CELL), 1)) {
            GUST_FAIL("gust: SimplecPreconditioner::setup: dtau must cover the valid cells");
        }
        hipStream_t st = static_cast<hipStream_t>(stream);
        N.pc_setup(*q, dtau, st);
        GUST_HIP(hipStreamSynchronize(st));
    });
}

int gust_newton_pc_apply(gust_newton* n, const gust_residuals* r, gust_residuals* dq, void* stream)
{
    return guarded([&] {
        if (!n || !r || !dq) { GUST_FAIL("gust: SimplecPreconditioner::apply: null argument"); }
        NewtonSolver& N = *as_n(n);
        if (N.opt.preconditioner != 1) {
            GUST_FAIL("gust: SimplecPreconditioner::apply: the Newton stage has the identity preconditioner");
        }
        for (int f = 0; f < kNF; ++f) {
            const IBox vb = loc_box(N.g, kStageLocs[f]);
            if (!covers(r->R[f], vb, 1) || !covers(dq->R[f], vb, 1)) {
                GUST_FAIL("gust: SimplecPreconditioner::apply: r and dq must cover every field's valid points");
            }
        }
        hipStream_t st = static_cast<hipStream_t>(stream);
        const double ones[kNF] = {1, 1, 1, 1, 1, 1};
        N.pack(r->R, ones, N.b.p, st);
        N.pc_apply(N.b.p, N.x.p, st);
        gust_array4* dst[kNF];
        for (int f = 0; f < kNF; ++f) { dst[f] = const_cast<gust_array4*>(dq->R[f]); }
        N.unpack(dst, ones, N.x.p, st);
        GUST_HIP(hipStreamSynchronize(st));
    });
}

int gust_continuation_step(gust_newton* n, gust_state* q, int32_t* status, gust_continuation_record* last,
                           void* stream)
{
    return guarded([&] {
        if (!n || !q || !status) { GUST_FAIL("gust: Continuation::step: null argument"); }
        check_state2("Continuation::step", as_n(n)->g, q, "q");
        *status = as_n(n)->continuation_step(q, last, static_cast<hipStream_t>(stream));
    });
}

int gust_continuation_get(const gust_newton* n, gust_continuation_state* st)
{
    return guarded([&] {
        if (!n || !st) { GUST_FAIL("gust: Continuation::state: null argument"); }
        *st = as_n(n)->cst;
    });
}

int gust_continuation_restore(gust_newton* n, const gust_continuation_state* st)
{
    return guarded([&] {
        if (!n || !st) { GUST_FAIL("gust: Continuation::restore: null argument"); }
        auto finite_nonneg = [](double x) { return std::isfinite(x) && x >= 0; };
        if (st->steps < 0 || !finite_nonneg(st->cfl) || !(st->cfl > 0) || !finite_nonneg(st->r0) ||
            !finite_nonneg(st->r_last)) {
            GUST_FAIL("gust: Continuation::restore: the state is inconsistent (steps >= 0, CFL finite and > 0, |R| "
                      "finite and >= 0)");
        }
        NewtonSolver& N = *as_n(n);
        N.cst = *st;
        N.cst.cfl = std::min(N.cst.cfl, N.opt.cfl_max);
    });
}

int gust_switch_default_options(gust_switch_options* sw)
{
    return guarded([&] {
        if (!sw) { GUST_FAIL("gust: switch_default_options: null options"); }
        sw->switch_window = 20;
        sw->switch_residual = 1.0e-4;
        sw->switch_back_growth = 10;
        sw->switch_monitor_tol = 1.0e-2;
        sw->newton_fallback_failures = 2;
    });
}

int gust_steady_solve_newton(gust_simplec* s, gust_newton* n, gust_state* q, const gust_driver_options* opt,
                             const gust_switch_options* sw, gust_newton_run_result* out, void* stream)
{
    return guarded([&] {
        const char* op = "Driver::run";
        if (!s || !n || !q || !opt || !sw || !out) { GUST_FAIL("gust: %s: null argument", op); }
        SimplecStage& S = *as_s(s);
        NewtonSolver& N = *as_n(n);
        if (N.stage != &S) { GUST_FAIL("gust: %s: the Newton stage was made on another SIMPLEC stage", op); }
        check_state2(op, S.g, q, "q");
        if (opt->max_iterations < 1) { GUST_FAIL("gust: %s: steady.max_iterations must be >= 1", op); }
        if (!(opt->cfl0 > 0 && std::isfinite(opt->cfl0)) ||
            !(opt->cfl_max >= opt->cfl0 && std::isfinite(opt->cfl_max))) {
            GUST_FAIL("gust: %s: steady.cfl0 must be > 0 and steady.cfl_max >= steady.cfl0", op);
        }
        if (sw->switch_window < 1) { GUST_FAIL("gust: %s: steady.switch_window must be >= 1", op); }
        if (!(sw->switch_residual > 0 && sw->switch_residual <= 1)) {
            GUST_FAIL("gust: %s: steady.switch_residual must be > 0 and <= 1", op);
        }
        if (!(std::isfinite(sw->switch_back_growth) && sw->switch_back_growth > 1)) {
            GUST_FAIL("gust: %s: steady.switch_back_growth must be finite and > 1", op);
        }
        if (!(sw->switch_monitor_tol >= 0 && sw->switch_monitor_tol < 1)) {
            GUST_FAIL("gust: %s: steady.switch_monitor_tol must be 0 (off), or > 0 and < 1", op);
        }
        if (sw->newton_fallback_failures < 1) { GUST_FAIL("gust: %s: steady.newton_fallback_failures must be >= 1", op); }
        hipStream_t st = static_cast<hipStream_t>(stream);
        const char* names[6] = {"u", "v", "w", "p", "theta", "tke"};
        const char* mnames[3] = {"max_speed", "mean_tke", "mean_ustar"};
        const double mscales[3] = {S.opt.scales.velocity, S.opt.scales.tke, 0.1};
        const int nmon = S.wall ? 3 : 2;
        gust_steady* crit = nullptr;
        check_rc(gust_steady_create(&opt->criterion, 6, names, nmon, mnames, mscales, &crit));
        struct Holder {
            gust_steady* c;
            ~Holder() { gust_steady_destroy(c); }
        } hold_c{crit};
        const int hold = opt->criterion.residual_ref_iterations;
        const int W = sw->switch_window;
        double cfl = opt->cfl0, r = 0;
        DevField dt;
        dt.alloc(loc_box(S.g, GUST_LOC_CELL), 1, st);
        gust_array4 nu = S.phys.use_closure ? S.cf[0].v : *S.phys.nu;
        NewtonStageRun ns{N, S, *opt, *sw};
        std::vector<double> rhist;
        std::vector<std::array<double, 3>> mhist;
        int stage = 0, below = 0, last_switch = -1;
        *out = gust_newton_run_result{};
        out->first_switch = -1;
        gust_steady_result v{};
        check_rc(gust_simplec_init(s, q, stream));
        int it = 0;
        while (it < opt->max_iterations) {
            if (stage == 0) {
                check_rc(gust_local_dtau(&S.g, q->f[0], q->f[1], q->f[2], &S.z, &S.J, &nu, cfl, &dt.v, stream));
                simplec_iterate(S, q, &dt.v, nullptr, stream);
                ++out->simplec_iterations;
            } else {
                ns.iterate(q, st);
                ++out->newton_iterations;
            }
            const int index = it;
            ++it;
            check_rc(gust_steady_add(crit, S.res, 6, S.mon, nmon));
            check_rc(gust_steady_relative_residual(crit, S.res, 6, hold, r, &r));
            rhist.push_back(r);
            mhist.push_back({S.mon[0], S.mon[1], S.mon[2]});
            check_rc(gust_steady_verdict(crit, &v, nullptr, 0));
            if (v.status != GUST_STEADY_NOT_STEADY) { break; }
            check_rc(gust_steady_next_cfl(crit, hold, cfl, r, opt->cfl0, opt->cfl_max, opt->ser_exponent,
                                          opt->ser_max_growth, &cfl));
            // the switch rules (Source/Driver.cpp:335-385)
            below = r <= sw->switch_residual ? below + 1 : 0;
            int to = stage;
            if (stage > 0 && ns.request) {
                to = stage - 1;
            } else if (last_switch >= 0 && index - last_switch < W) {
                continue;
            } else if (below >= W && stage == 0 && !ns.handed_back) {
                bool settled = true;
                if (sw->switch_monitor_tol > 0) {
                    if (int(mhist.size()) < W + 1) {
                        settled = false;
                    } else {
                        const auto& now = mhist.back();
                        const auto& then = mhist[mhist.size() - 1 - size_t(W)];
                        for (int m = 0; m < nmon; ++m) {
                            const double scale = std::max(std::abs(now[size_t(m)]), mscales[m]);
                            if (!(std::abs(now[size_t(m)] - then[size_t(m)]) <= sw->switch_monitor_tol * scale)) {
                                settled = false;
                            }
                        }
                    }
                }
                if (settled) { to = 1; }
            } else if (stage > 0 && int(rhist.size()) > W && index - W > last_switch) {
                const double before = rhist[rhist.size() - 1 - size_t(W)];
                if (r >= sw->switch_back_growth * before) { to = 0; }
            }
            if (to == stage) { continue; }
            ++out->switches;
            if (stage == 1) { ns.leave(); }
            stage = to;
            below = 0;
            last_switch = index;
            if (to == 1) {
                if (out->first_switch < 0) { out->first_switch = index; }
                ns.init(q, st);
            } else {
                check_rc(gust_simplec_init(s, q, stream));
            }
        }
        gust_driver_result& res = out->driver;
        res.iterations = it;
        res.status = v.status;
        res.cfl = cfl;
        res.relative_residual = r;
        for (int k = 0; k < 6; ++k) { res.res[k] = S.res[k]; }
        for (int k = 0; k < 3; ++k) { res.mon[k] = S.mon[k]; }
        res.nmon = nmon;
        out->newton_steps = N.cst.steps;
        out->stage = stage;
        out->handed_back = ns.handed_back ? 1 : 0;
        GUST_HIP(hipStreamSynchronize(st));
    });
}

}   // extern "C"
