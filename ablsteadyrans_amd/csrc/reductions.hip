// reductions.hip -- the reductions of the SIMPLEC control loop on a tile:
// the column-ordered deterministic sum (Source/Reductions.cpp:11-53), the
// NaN-reporting maximum (55-66), the scaled sums of the Newton step
// (Source/Residual.cpp:101-150), the stage norms and monitors
// (Source/Simplec.cpp:374-409) and the confidence share (411-445).
//
// Column-ordered sums keep the reference's order exactly: one thread per
// (i, j) column adds its k values from 0, the host adds the column sums in
// storage order, so the bits equal a CPU loop in that order. The norms use a
// fixed two-level tree instead (block partials, then one block per field):
// the same tile gives the same bits on every call and any stream.
#include "cgrid.hpp"
#include "deferred.hpp"
#include "reduce.hpp"

#include <array>
#include <cmath>
#include <functional>
#include <memory>
#include <cstring>
#include <limits>

namespace gust_amd {

namespace {

using cgrid::blocks_for;
using cgrid::Box3;
using cgrid::box3;
using cgrid::check_geom;
using cgrid::CV;
using cgrid::kThreads;

// Workspace slots of this file.
constexpr int kSlotCols = 60;
constexpr int kSlotPart = 61;
constexpr int kSlotOut = 62;

CV comp_view(const gust_array4* a, int comp)
{
    CV v = dview<const double>(a);
    v.p += int64_t(comp) * a->nstride;
    return v;
}

// One thread per (i, j) column of b: the sum of its values in k order.
__global__ __launch_bounds__(kThreads) void column_sum_kernel(Box3 b, CV a, double* cs)
{
    const int ncol = b.n0 * b.n1;
    for (int t = blockIdx.x * kThreads + threadIdx.x; t < ncol; t += gridDim.x * kThreads) {
        const int i = b.lo0 + t % b.n0, j = b.lo1 + t / b.n0;
        double s = 0;
        for (int k = b.lo2; k < b.lo2 + b.n2; ++k) { s += a(i, j, k); }
        cs[t] = s;
    }
}

// The three scaled sums of one column; the columns of a skipped face (the
// periodic copy) contribute zeros, as the owner mask's points do.
__global__ __launch_bounds__(kThreads) void scaled_sums_kernel(Box3 b, CV q, CV v, CV off, int has_off,
                                                              double sinv, int skip_i, int skip_j, double* cs)
{
    const int ncol = b.n0 * b.n1;
    for (int t = blockIdx.x * kThreads + threadIdx.x; t < ncol; t += gridDim.x * kThreads) {
        const int i = b.lo0 + t % b.n0, j = b.lo1 + t / b.n0;
        double s0 = 0, s1 = 0, s2 = 0;
        if (i != skip_i && j != skip_j) {
            for (int k = b.lo2; k < b.lo2 + b.n2; ++k) {
                const double qq = has_off ? q(i, j, k) - off(i, j, k) : q(i, j, k);
                const double vv = v(i, j, k) * sinv;
                s0 += std::fabs(qq) * sinv;
                s1 += vv * vv;
                s2 += 1.0;
            }
        }
        cs[3 * t] = s0;
        cs[3 * t + 1] = s1;
        cs[3 * t + 2] = s2;
    }
}

// max over double keys ordered as the reals (negative values flipped).
__device__ __forceinline__ unsigned long long order_key(double x)
{
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

double from_key(unsigned long long k)
{
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double x;
    std::memcpy(&x, &b, sizeof x);
    return x;
}

// Flag any NaN; max of the others as ordered keys.
__global__ __launch_bounds__(kThreads) void max_kernel(Box3 b, CV a, unsigned long long* out)
{
    unsigned long long m = 0, nan = 0;
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const double x = a(i, j, k);
        if (x != x) {
            nan = 1;
        } else {
            const unsigned long long key = order_key(x);
            m = key > m ? key : m;
        }
    }
    m = dev::block_max_u64<kThreads>(m);
    nan = dev::block_max_u64<kThreads>(nan);
    if (threadIdx.x == 0) {
        if (m) { atomicMax(out, m); }
        if (nan) { atomicOr(out + 1, 1ull); }
    }
}

// ---- fixed-tree reductions of several fields in one launch ---------------

enum RedOp : int { kSumSq = 0, kSum = 1, kMaxAbs = 2 };

struct RField {
    const double* p;
    int64_t js, ks;
    int b0, b1, b2;
    int lo0, lo1, lo2, n0, n1, n2;
    int op;
};

constexpr int kMaxFields = 12;
constexpr int kPartBlocks = 128;

struct RSet {
    RField f[kMaxFields];
};

__device__ __forceinline__ double combine(double a, double b, int op)
{
    if (op != kMaxAbs) { return a + b; }
    if (a != a) { return a; }
    if (b != b) { return b; }
    return b > a ? b : a;
}

// Block-wide reduction in a fixed order: a shuffle tree per wave, then the
// waves' values in index order.
__device__ double block_reduce(double v, int op)
{
    __shared__ double wv[kThreads / 64];
    for (int off = 32; off >= 1; off >>= 1) { v = combine(v, __shfl_down(v, off, 64), op); }
    if ((threadIdx.x & 63) == 0) { wv[threadIdx.x >> 6] = v; }
    __syncthreads();
    double r = wv[0];
    for (int w = 1; w < kThreads / 64; ++w) { r = combine(r, wv[w], op); }
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kThreads) void multi_partial_kernel(RSet s, double* part)
{
    const RField f = s.f[blockIdx.y];
    const unsigned n = unsigned(f.n0) * unsigned(f.n1) * unsigned(f.n2);
    double acc = 0;
    for (unsigned t = blockIdx.x * kThreads + threadIdx.x; t < n; t += gridDim.x * kThreads) {
        const unsigned r = t / unsigned(f.n0);
        const int i = f.lo0 + int(t - r * unsigned(f.n0));
        const int j = f.lo1 + int(r % unsigned(f.n1));
        const int k = f.lo2 + int(r / unsigned(f.n1));
        const double x = f.p[int64_t(i - f.b0) + int64_t(j - f.b1) * f.js + int64_t(k - f.b2) * f.ks];
        if (f.op == kSumSq) {
            acc += x * x;
        } else if (f.op == kSum) {
            acc += x;
        } else {
            acc = combine(acc, std::fabs(x), kMaxAbs);
        }
    }
    const double r = block_reduce(acc, f.op);
    if (threadIdx.x == 0) { part[blockIdx.y * gridDim.x + blockIdx.x] = r; }
}

__global__ __launch_bounds__(kThreads) void multi_final_kernel(RSet s, const double* part, int nb, double* out)
{
    const int op = s.f[blockIdx.x].op;
    double acc = 0;
    for (int t = threadIdx.x; t < nb; t += kThreads) { acc = combine(acc, part[blockIdx.x * nb + t], op); }
    const double r = block_reduce(acc, op);
    if (threadIdx.x == 0) { out[blockIdx.x] = r; }
}

// ---- the confidence share --------------------------------------------------

struct Thresholds {
    double t[6];
};

__device__ __forceinline__ double amax(double a, double b) { return (a < b) ? b : a; }   // std::max

__global__ __launch_bounds__(kThreads) void confidence_kernel(Box3 b, CV ru, CV rv, CV rw, CV rp, CV rt, CV rk,
                                                             CV z, int klo, Thresholds th,
                                                             unsigned long long* out)
{
    unsigned long long bad = 0, top = 0;
    GUST_GRID_LOOP(b)
    {
        int i, j, k;
        b.at(t, i, j, k);
        const bool ok = amax(std::fabs(ru(i, j, k)), std::fabs(ru(i + 1, j, k))) <= th.t[0] &&
                        amax(std::fabs(rv(i, j, k)), std::fabs(rv(i, j + 1, k))) <= th.t[1] &&
                        amax(std::fabs(rw(i, j, k)), std::fabs(rw(i, j, k + 1))) <= th.t[2] &&
                        std::fabs(rp(i, j, k)) <= th.t[3] && std::fabs(rt(i, j, k)) <= th.t[4] &&
                        std::fabs(rk(i, j, k)) <= th.t[5];
        double val = 0.0;
        if (!ok) {
            ++bad;
            double zc = 0, zg = 0;
            for (int a = 0; a < 2; ++a) {
                for (int c = 0; c < 2; ++c) {
                    zg += z(i + a, j + c, klo);
                    zc += z(i + a, j + c, k) + z(i + a, j + c, k + 1);
                }
            }
            val = 0.125 * zc - 0.25 * zg;
        }
        const unsigned long long key = order_key(val);
        top = key > top ? key : top;
    }
    bad = dev::block_sum_u64<kThreads>(bad);
    top = dev::block_max_u64<kThreads>(top);
    if (threadIdx.x == 0) {
        if (bad) { atomicAdd(out, bad); }
        atomicMax(out + 1, top);
    }
}

// ---- host helpers -----------------------------------------------------------

void need(bool ok, const char* op, const char* what)
{
    if (!ok) { GUST_FAIL("gust: %s: %s", op, what); }
}

void check_view(const char* op, const char* what, const gust_array4* a, const IBox& b, int comp = 0)
{
    if (!covers(a, b, comp + 1)) {
        GUST_FAIL("gust: %s: %s does not cover the tile's valid points", op, what);
    }
}

template <typename T>
T* fetch(hipStream_t st, const void* dev, size_t n, int slot)
{
    T* h = static_cast<T*>(Workspace::get_host_pinned(slot, n * sizeof(T)));
    GUST_HIP(hipGetLastError());
    GUST_HIP(hipMemcpyAsync(h, dev, n * sizeof(T), hipMemcpyDeviceToHost, st));
    GUST_HIP(hipStreamSynchronize(st));
    return h;
}

RField rfield(const gust_array4* a, const IBox& b, int op)
{
    RField f;
    f.p = static_cast<const double*>(a->p);
    f.js = a->jstride;
    f.ks = a->kstride;
    f.b0 = a->begin[0];
    f.b1 = a->begin[1];
    f.b2 = a->begin[2];
    f.lo0 = b.lo[0];
    f.lo1 = b.lo[1];
    f.lo2 = b.lo[2];
    f.n0 = b.len(0);
    f.n1 = b.len(1);
    f.n2 = b.len(2);
    f.op = op;
    return f;
}

// Reduce every field of s (n of them) with the fixed tree into device memory.
double* multi_reduce_dev(hipStream_t st, const RSet& s, int n)
{
    int64_t most = 1;
    for (int q = 0; q < n; ++q) {
        most = std::max<int64_t>(most, int64_t(s.f[q].n0) * s.f[q].n1 * s.f[q].n2);
    }
    const int nb = int(std::min<int64_t>(kPartBlocks, (most + kThreads - 1) / kThreads));
    double* part = static_cast<double*>(Workspace::get(st, kSlotPart, size_t(n) * nb * sizeof(double)));
    double* dout = static_cast<double*>(Workspace::get(st, kSlotOut, kMaxFields * sizeof(double)));
    multi_partial_kernel<<<dim3(nb, n), kThreads, 0, st>>>(s, part);
    multi_final_kernel<<<n, kThreads, 0, st>>>(s, part, nb, dout);
    return dout;
}

// ... and host values out.
void multi_reduce(hipStream_t st, const RSet& s, int n, double* out)
{
    const double* h = fetch<double>(st, multi_reduce_dev(st, s, n), n, kSlotOut);
    for (int q = 0; q < n; ++q) { out[q] = h[q]; }
}

struct SlotPtrs {
    unsigned long long* p[kMaxFields];
};

__global__ void to_slots_kernel(const double* src, SlotPtrs dst, int n)
{
    const int t = threadIdx.x;
    if (t < n) { *dst.p[t] = __double_as_longlong(src[t]); }
}

// ... read with the caller's batch: done(values) runs when it resolves.
void multi_reduce_then(hipStream_t st, const RSet& s, int n, std::function<void(const double*)> done)
{
    const double* dout = multi_reduce_dev(st, s, n);
    DeferredUse dq(st);
    auto vals = std::make_shared<std::array<double, kMaxFields>>();
    SlotPtrs dst{};
    for (int q = 0; q < n; ++q) {
        dst.p[q] = dq->zero_slot([vals, q](unsigned long long b) { (*vals)[size_t(q)] = bits_to_double(b); });
    }
    to_slots_kernel<<<1, 64, 0, st>>>(dout, dst, n);
    GUST_HIP(hipGetLastError());
    dq->then([vals, f = std::move(done)] { f(vals->data()); });
    dq.finish();
}

const char* const kResNames[6] = {"R.u", "R.v", "R.w", "R.p", "R.theta", "R.tke"};
const int kResLoc[6] = {GUST_LOC_XFACE, GUST_LOC_YFACE, GUST_LOC_ZFACE, GUST_LOC_CELL, GUST_LOC_CELL,
                        GUST_LOC_CELL};

void check_scales(const char* op, const gust_jvp_scales* s)
{
    need(s != nullptr, op, "null residual scales");
    for (const double x : {s->velocity, s->theta, s->tke}) {
        if (!std::isfinite(x) || !(x > 0)) {
            GUST_FAIL("gust: %s: the residual scales (jvp_scales) must be finite and > 0", op);
        }
    }
}

void scales_of(const gust_jvp_scales& s, double H, double scale[6])
{
    const double U = s.velocity;
    scale[0] = scale[1] = scale[2] = U * U / H;
    scale[3] = U / H;
    scale[4] = U * s.theta / H;
    scale[5] = U * s.tke / H;
}

void check_residuals(const char* op, const gust_geometry& g, const gust_residuals* R)
{
    need(R != nullptr, op, "null residual fields");
    for (int n = 0; n < 6; ++n) { check_view(op, kResNames[n], R->R[n], loc_box(g, kResLoc[n])); }
}

}   // namespace

double sum_squares(hipStream_t st, const gust_array4* a, const IBox& b)
{
    if (b.npts() * 2 >= int64_t(std::numeric_limits<unsigned>::max())) {
        GUST_FAIL("gust: sum_squares: tile too large");
    }
    RSet set{};
    set.f[0] = rfield(a, b, kSumSq);
    double out = 0;
    multi_reduce(st, set, 1, &out);
    return out;
}

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_deterministic_sum(const gust_geometry* geom, const gust_array4* a, int32_t comp, double* sum, void* stream)
{
    return guarded([&] {
        const char* op = "deterministic_sum";
        check_geom(op, geom);
        need(sum != nullptr, op, "null output");
        need(comp >= 0, op, "negative component");
        const IBox vb = loc_box(*geom, GUST_LOC_CELL);
        check_view(op, "the field", a, vb, comp);
        hipStream_t st = static_cast<hipStream_t>(stream);
        const Box3 b = box3(vb);
        const int ncol = b.n0 * b.n1;
        double* cs = static_cast<double*>(Workspace::get(st, kSlotCols, size_t(ncol) * sizeof(double)));
        column_sum_kernel<<<blocks_for(ncol), kThreads, 0, st>>>(b, comp_view(a, comp), cs);
        const double* h = fetch<double>(st, cs, size_t(ncol), kSlotCols);
        double s = 0;
        for (int c = 0; c < ncol; ++c) { s += h[c]; }
        *sum = s;
    });
}

int gust_global_max(const gust_geometry* geom, int32_t loc, const gust_array4* a, int32_t comp, double* max,
                    void* stream)
{
    return guarded([&] {
        const char* op = "global_max";
        check_geom(op, geom);
        need(valid_loc(loc), op, "invalid location");
        need(max != nullptr, op, "null output");
        need(comp >= 0, op, "negative component");
        const IBox vb = loc_box(*geom, loc);
        check_view(op, "the field", a, vb, comp);
        hipStream_t st = static_cast<hipStream_t>(stream);
        auto* d = static_cast<unsigned long long*>(Workspace::get(st, kSlotOut, 2 * sizeof(unsigned long long)));
        GUST_HIP(hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), st));
        const Box3 b = box3(vb);
        max_kernel<<<cgrid::reduce_blocks(b.npts()), kThreads, 0, st>>>(b, comp_view(a, comp), d);
        const unsigned long long* h = fetch<unsigned long long>(st, d, 2, kSlotOut);
        if (h[1]) {
            *max = std::numeric_limits<double>::quiet_NaN();
        } else {
            const double m = from_key(h[0]);
            *max = (m == 0.0) ? 0.0 : m;
        }
    });
}

int gust_scaled_sums(const gust_geometry* geom, int32_t loc, const gust_array4* q, const gust_array4* v,
                     const gust_array4* offset, double s, double sums[3], void* stream)
{
    return guarded([&] {
        const char* op = "scaled_sums";
        check_geom(op, geom);
        need(valid_loc(loc), op, "invalid location");
        need(sums != nullptr, op, "null output");
        need(std::isfinite(s) && s > 0, op, "the scale must be finite and > 0");
        const IBox vb = loc_box(*geom, loc);
        check_view(op, "q", q, vb);
        check_view(op, "v", v, vb);
        if (offset) { check_view(op, "the offset", offset, vb); }
        // The owner mask of one periodic box: the hi face of a nodal periodic
        // axis is the image of the lo face and is not owned.
        const int skip_i = (geom->periodic[0] && (loc & 1)) ? vb.hi[0] : std::numeric_limits<int>::min();
        const int skip_j = (geom->periodic[1] && (loc & 2)) ? vb.hi[1] : std::numeric_limits<int>::min();
        hipStream_t st = static_cast<hipStream_t>(stream);
        const Box3 b = box3(vb);
        const int ncol = b.n0 * b.n1;
        double* cs = static_cast<double*>(Workspace::get(st, kSlotCols, size_t(3 * ncol) * sizeof(double)));
        const CV qv = dview<const double>(q);
        scaled_sums_kernel<<<blocks_for(ncol), kThreads, 0, st>>>(b, qv, dview<const double>(v),
                                                                  offset ? dview<const double>(offset) : qv,
                                                                  offset ? 1 : 0, 1.0 / s, skip_i, skip_j, cs);
        const double* h = fetch<double>(st, cs, size_t(3 * ncol), kSlotCols);
        double t[3] = {0, 0, 0};
        for (int c = 0; c < ncol; ++c) {
            for (int n = 0; n < 3; ++n) { t[n] += h[3 * c + n]; }
        }
        for (int n = 0; n < 3; ++n) { sums[n] = t[n]; }
    });
}

int gust_residual_scales(const gust_jvp_scales* s, double height, double scale[6])
{
    return guarded([&] {
        check_scales("residual_scales", s);
        need(scale != nullptr, "residual_scales", "null output");
        need(std::isfinite(height) && height > 0, "residual_scales", "the height must be finite and > 0");
        scales_of(*s, height, scale);
    });
}

int gust_residual_scale_height(const double* zlevels, int32_t nlevels, double prob_lo_z, double prob_hi_z,
                               double* height)
{
    return guarded([&] {
        if (!zlevels || nlevels < 1) {
            GUST_FAIL("gust: residual_scale_height: the mesh has no levels (not built)");
        }
        need(height != nullptr, "residual_scale_height", "null output");
        const double depth = zlevels[nlevels - 1] - zlevels[0];
        const double box = prob_hi_z - prob_lo_z;
        const double tol = 1000.0 * std::numeric_limits<double>::epsilon() * box;
        *height = (std::fabs(depth - box) <= tol) ? box : depth;
    });
}

int gust_stage_norms(const gust_geometry* geom, const gust_residuals* R, const gust_array4* u, const gust_array4* v,
                     const gust_array4* w, const gust_array4* tke, const gust_array4* ustar,
                     const gust_jvp_scales* s, double height, double res[6], double mon[3], int32_t* nmon,
                     void* stream)
{
    return guarded([&] {
        const char* op = "stage_norms";
        check_geom(op, geom);
        check_scales(op, s);
        need(std::isfinite(height) && height > 0, op, "the height must be finite and > 0");
        need(res && mon && nmon, op, "null output");
        const gust_geometry& g = *geom;
        check_residuals(op, g, R);
        check_view(op, "u", u, loc_box(g, GUST_LOC_XFACE));
        check_view(op, "v", v, loc_box(g, GUST_LOC_YFACE));
        check_view(op, "w", w, loc_box(g, GUST_LOC_ZFACE));
        check_view(op, "tke", tke, loc_box(g, GUST_LOC_CELL));
        IBox slab = loc_box(g, GUST_LOC_CELL);
        slab.hi[2] = slab.lo[2];
        if (ustar) { check_view(op, "ustar", ustar, slab); }
        const IBox cells = loc_box(g, GUST_LOC_CELL);
        need(cells.npts() * 2 < int64_t(std::numeric_limits<unsigned>::max()), op, "tile too large");

        RSet set{};
        int n = 0;
        for (int q = 0; q < 6; ++q) { set.f[n++] = rfield(R->R[q], loc_box(g, kResLoc[q]), kSumSq); }
        set.f[n++] = rfield(u, loc_box(g, GUST_LOC_XFACE), kMaxAbs);
        set.f[n++] = rfield(v, loc_box(g, GUST_LOC_YFACE), kMaxAbs);
        set.f[n++] = rfield(w, loc_box(g, GUST_LOC_ZFACE), kMaxAbs);
        set.f[n++] = rfield(tke, cells, kSum);
        if (ustar) { set.f[n++] = rfield(ustar, slab, kSum); }
        // The norms arrive with the caller's batch (at once when there is none).
        double scale[6], npts[6];
        scales_of(*s, height, scale);
        for (int q = 0; q < 6; ++q) { npts[q] = double(loc_box(g, kResLoc[q]).npts()); }
        const double ncells = double(cells.npts()), ncols = double(int64_t(cells.len(0)) * cells.len(1));
        const bool wall = ustar != nullptr;
        std::array<double, 6> sc, np;
        std::copy(scale, scale + 6, sc.begin());
        std::copy(npts, npts + 6, np.begin());
        multi_reduce_then(static_cast<hipStream_t>(stream), set, n,
                          [=](const double* out) {
                              for (int q = 0; q < 6; ++q) { res[q] = std::sqrt(out[q]) / std::sqrt(np[q]) / sc[q]; }
                              mon[0] = std::max(std::max(out[6], out[7]), out[8]);
                              mon[1] = out[9] / ncells;
                              mon[2] = wall ? out[10] / ncols : 0.0;
                              *nmon = wall ? 3 : 2;
                          });
    });
}

int gust_confidence(const gust_geometry* geom, const gust_residuals* R, const gust_array4* z_nd,
                    const gust_jvp_scales* s, double height, const double threshold[6], double* share, double* top,
                    void* stream)
{
    return guarded([&] {
        const char* op = "confidence";
        check_geom(op, geom);
        check_scales(op, s);
        need(std::isfinite(height) && height > 0, op, "the height must be finite and > 0");
        need(threshold && share && top, op, "null argument");
        const gust_geometry& g = *geom;
        check_residuals(op, g, R);
        IBox nodes = loc_box(g, GUST_LOC_XFACE | GUST_LOC_YFACE | GUST_LOC_ZFACE);
        check_view(op, "z_phys_nd", z_nd, nodes);
        double scale[6];
        scales_of(*s, height, scale);
        Thresholds th;
        for (int q = 0; q < 6; ++q) { th.t[q] = threshold[q] * scale[q]; }
        hipStream_t st = static_cast<hipStream_t>(stream);
        auto* d = static_cast<unsigned long long*>(Workspace::get(st, kSlotOut, 2 * sizeof(unsigned long long)));
        GUST_HIP(hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), st));
        const IBox cells = loc_box(g, GUST_LOC_CELL);
        const Box3 b = box3(cells);
        confidence_kernel<<<cgrid::reduce_blocks(b.npts()), kThreads, 0, st>>>(
            b, dview<const double>(R->R[0]), dview<const double>(R->R[1]), dview<const double>(R->R[2]),
            dview<const double>(R->R[3]), dview<const double>(R->R[4]), dview<const double>(R->R[5]),
            dview<const double>(z_nd), cells.lo[2], th, d);
        const unsigned long long* h = fetch<unsigned long long>(st, d, 2, kSlotOut);
        *share = 1.0 - double(h[0]) / double(cells.npts());
        *top = from_key(h[1]);
    });
}

}   // extern "C"
