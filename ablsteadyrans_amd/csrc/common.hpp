// common.hpp -- shared host/device plumbing of libgust_amd: error reporting,
// view checks, the per-(device, stream) workspace, and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "gust_amd.h"

#include "errors.hpp"

namespace gust_amd {

#define GUST_HIP(expr)                                                                                    \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) {                                                                           \
            ::gust_amd::set_error("gust: HIP error %s at %s:%d (%s)", hipGetErrorString(e_), __FILE__,     \
                                  __LINE__, #expr);                                                       \
            throw ::gust_amd::Fail{2};                                                                    \
        }                                                                                                 \
    } while (0)

// ---------------------------------------------------------------------------
// Boxes and views.

struct IBox {
    int lo[3], hi[3];   // inclusive
    int len(int d) const { return hi[d] - lo[d] + 1; }
    int64_t npts() const { return int64_t(len(0)) * len(1) * len(2); }
};

// The domain box converted to a field location (bit d of loc = nodal in d).
inline IBox loc_box(const gust_geometry& g, int loc)
{
    IBox b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = g.lo[d];
        b.hi[d] = g.hi[d] + ((loc >> d) & 1);
    }
    return b;
}

inline bool valid_loc(int loc)
{
    return loc == GUST_LOC_CELL || loc == GUST_LOC_XFACE || loc == GUST_LOC_YFACE || loc == GUST_LOC_ZFACE;
}

const char* loc_words(int loc);

// True when the view's box [begin, end) covers b with at least ncomp comps.
inline bool covers(const gust_array4* v, const IBox& b, int ncomp)
{
    if (!v || !v->p || v->ncomp < ncomp) { return false; }
    for (int d = 0; d < 3; ++d) {
        if (v->begin[d] > b.lo[d] || v->end[d] - 1 < b.hi[d]) { return false; }
    }
    return true;
}

inline IBox grow(IBox b, int gx, int gy, int gz)
{
    b.lo[0] -= gx; b.hi[0] += gx;
    b.lo[1] -= gy; b.hi[1] += gy;
    b.lo[2] -= gz; b.hi[2] += gz;
    return b;
}

// Element offset of (i, j, k, n) in a view.
inline int64_t offset_of(const gust_array4& v, int i, int j, int k, int n = 0)
{
    return int64_t(i - v.begin[0]) + int64_t(j - v.begin[1]) * v.jstride + int64_t(k - v.begin[2]) * v.kstride +
           int64_t(n) * v.nstride;
}

// A view as kernels read it: component 0 at (i, j, k).
template <typename T>
struct DView {
    T* p;
    int64_t js, ks;
    int b0, b1, b2;
    __host__ __device__ __forceinline__ T& operator()(int i, int j, int k) const
    {
        return p[int64_t(i - b0) + int64_t(j - b1) * js + int64_t(k - b2) * ks];
    }
};

template <typename T>
DView<T> dview(const gust_array4* v)
{
    return DView<T>{static_cast<T*>(v->p), v->jstride, v->kstride, v->begin[0], v->begin[1], v->begin[2]};
}

// check_operator_field (Source/CVFluxes.cpp:95-110) on a view: the field at
// loc must cover the domain's box at loc grown by ngrow.
void check_field(const char* op, const char* what, const gust_array4* v, const gust_geometry& g, int loc,
                 int ngrow);
// check_nodal_heights (Source/CVFluxes.cpp:112-119).
void check_nodal_heights(const char* op, const gust_array4* z, const gust_geometry& g);

// ---------------------------------------------------------------------------
// Workspace: grow-only device buffers keyed by (device, stream, slot). A
// buffer is reused by later calls on the same stream (stream order keeps
// them apart); growing frees the old buffer after synchronising the stream.

class Workspace {
public:
    static void* get(hipStream_t s, int slot, size_t bytes);
    // Pinned host memory, one buffer per (host thread, slot).
    static void* get_host_pinned(int slot, size_t bytes);
    // Device bytes held for one stream of the current device (workspace
    // buffers plus what the per-stream caches report).
    static size_t bytes_held(hipStream_t s);
    // Synchronises the stream and frees everything held for it on the
    // current device; the next call on that stream allocates afresh.
    static void release(hipStream_t s);
};

// Per-stream device caches (tables, profiles) register a hook that reports
// and drops their entries of one (device, stream); release() calls them.
struct StreamCacheHook {
    size_t (*bytes)(int device, hipStream_t s);
    void (*drop)(int device, hipStream_t s);
};
void register_stream_cache(StreamCacheHook h);

// Device properties used for launch shapes.
struct DeviceInfo {
    int cus = 256;
    size_t lds_per_block = 65536;
};
const DeviceInfo& device_info();

}   // namespace gust_amd

// ---------------------------------------------------------------------------
// Device helpers.

namespace gust_amd::dev {

// Block index remapped so that the blocks the hardware places on one XCD
// (round-robin, block b on XCD b % 8) cover one contiguous range of logical
// tiles: neighbouring column rows then share an L2.
__device__ __forceinline__ int xcd_remap(int b, int nblocks)
{
    constexpr int kXcd = 8;
    const int q = nblocks / kXcd, r = nblocks % kXcd;
    const int x = b % kXcd, idx = b / kXcd;
    return x * q + (x < r ? x : r) + idx;
}

// |x| as ordered bits: max over these is max |x| with NaN winning.
__device__ __forceinline__ unsigned long long abs_bits(double x)
{
    return static_cast<unsigned long long>(__double_as_longlong(x)) & 0x7fffffffffffffffull;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v)
{
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v)
{
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
    for (int off = 32; off >= 1; off >>= 1) { v += __shfl_xor(v, off, 64); }
    return v;
}

// Block-wide max / min / sum of v over a block of kBlock threads (a multiple
// of 64); every thread of the block must call it (it holds two barriers).
// The result is valid in thread 0. One atomic per block instead of one per wave keeps a reduction
// over millions of points from serialising on a single address.
template <int kBlock>
__device__ __forceinline__ unsigned long long block_max_u64(unsigned long long v)
{
    __shared__ unsigned long long s[kBlock / 64];
    v = wave_max_u64(v);
    if ((threadIdx.x & 63) == 0) { s[threadIdx.x >> 6] = v; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) { v = s[w] > v ? s[w] : v; }
    }
    __syncthreads();   // s is reused by the next call
    return v;
}

template <int kBlock>
__device__ __forceinline__ unsigned long long block_min_u64(unsigned long long v)
{
    __shared__ unsigned long long s[kBlock / 64];
    v = wave_min_u64(v);
    if ((threadIdx.x & 63) == 0) { s[threadIdx.x >> 6] = v; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) { v = s[w] < v ? s[w] : v; }
    }
    __syncthreads();   // s is reused by the next call
    return v;
}

template <int kBlock>
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v)
{
    __shared__ unsigned long long s[kBlock / 64];
    v = wave_sum_u64(v);
    if ((threadIdx.x & 63) == 0) { s[threadIdx.x >> 6] = v; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) { v += s[w]; }
    }
    __syncthreads();   // s is reused by the next call
    return v;
}

// Block-wide sum of doubles in a fixed order (a shuffle tree per wave, then
// the waves in index order): the same bits on every call. Valid in thread 0.
template <int kBlock>
__device__ __forceinline__ double block_sum_f64(double v)
{
    __shared__ double s[kBlock / 64];
    for (int off = 32; off >= 1; off >>= 1) { v += __shfl_down(v, off, 64); }
    if ((threadIdx.x & 63) == 0) { s[threadIdx.x >> 6] = v; }
    __syncthreads();
    double r = s[0];
    for (int w = 1; w < kBlock / 64; ++w) { r += s[w]; }
    __syncthreads();   // s is reused by the next call
    return r;
}

// Global-memory accesses through address-space-1 pointers: where a pointer
// comes out of a table in memory the compiler cannot infer its space and
// would emit flat instructions (which also count against the LDS counter and
// serialise the waits).
#define GUST_GLOBAL __attribute__((address_space(1)))

template <typename T>
__device__ __forceinline__ T ldg(const T* p)
{
    return __builtin_nontemporal_load((const GUST_GLOBAL T*)(p));
}

template <typename T>
__device__ __forceinline__ T gld(const T* p)
{
    return *((const GUST_GLOBAL T*)(p));
}

template <typename T>
__device__ __forceinline__ void gst(T* p, T v)
{
    *((GUST_GLOBAL T*)(p)) = v;
}

}   // namespace gust_amd::dev
