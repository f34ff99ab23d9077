// common.hip -- host plumbing of libgust_amd (errors, workspaces, device info).
#include "common.hpp"

#include <map>
#include <vector>
#include <tuple>

namespace gust_amd {

namespace {
thread_local std::string t_err;
}

void set_error(const char* fmt, ...)
{
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
}

void clear_error() { t_err.clear(); }

const char* loc_words(int loc)
{
    switch (loc) {
    case GUST_LOC_CELL: return "cell centres";
    case GUST_LOC_XFACE: return "x-faces";
    case GUST_LOC_YFACE: return "y-faces";
    case GUST_LOC_ZFACE: return "z-faces";
    default: return "?";
    }
}

void check_field(const char* op, const char* what, const gust_array4* v, const gust_geometry& g, int loc,
                 int ngrow)
{
    const IBox vb = loc_box(g, loc);
    if (!covers(v, vb, 1)) {
        GUST_FAIL("gust: %s: %s is not on the mesh's BoxArray and DistributionMapping at the %s", op, what,
                  loc_words(loc));
    }
    if (!covers(v, grow(vb, ngrow, ngrow, ngrow), 1)) {
        GUST_FAIL("gust: %s: %s needs at least %d ghost cells", op, what, ngrow);
    }
}

void check_nodal_heights(const char* op, const gust_array4* z, const gust_geometry& g)
{
    IBox nodes;
    for (int d = 0; d < 3; ++d) {
        nodes.lo[d] = g.lo[d] - 1;
        nodes.hi[d] = g.hi[d] + 2;
    }
    if (!covers(z, nodes, 1)) {
        GUST_FAIL("gust: %s: z_nd must be the mesh's node heights (nodal, same layout, one ghost node)", op);
    }
}

namespace {

struct Buf {
    void* p = nullptr;
    size_t cap = 0;
};

std::mutex g_ws_mutex;
std::map<std::tuple<int, hipStream_t, int>, Buf> g_ws;

// Pinned host buffers are per host thread: threads driving their own
// streams read back concurrently without sharing one.
struct PinnedSet {
    std::map<int, Buf> bufs;
    ~PinnedSet()
    {
        for (auto& kv : bufs) {
            if (kv.second.p) { (void)hipHostFree(kv.second.p); }
        }
    }
};
thread_local PinnedSet t_pinned;

}   // namespace

void* Workspace::get(hipStream_t s, int slot, size_t bytes)
{
    int dev = 0;
    GUST_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_ws_mutex);
    Buf& b = g_ws[std::make_tuple(dev, s, slot)];
    if (b.cap < bytes) {
        if (b.p) {
            GUST_HIP(hipStreamSynchronize(s));
            GUST_HIP(hipFree(b.p));
            b.p = nullptr;
            b.cap = 0;
        }
        const size_t want = bytes + bytes / 8 + 256;
        GUST_HIP(hipMalloc(&b.p, want));
        b.cap = want;
    }
    return b.p;
}

void* Workspace::get_host_pinned(int slot, size_t bytes)
{
    Buf& b = t_pinned.bufs[slot];
    if (b.cap < bytes) {
        if (b.p) {
            // No copy into this buffer can be pending: every use of it is
            // followed by a synchronisation of its own stream before the
            // thread goes on, so no device-wide wait is needed here.
            GUST_HIP(hipHostFree(b.p));
            b.p = nullptr;
            b.cap = 0;
        }
        const size_t want = bytes + 256;
        GUST_HIP(hipHostMalloc(&b.p, want, hipHostMallocDefault));
        b.cap = want;
    }
    return b.p;
}

namespace {
std::mutex g_hooks_mutex;
std::vector<StreamCacheHook>& hooks()
{
    static std::vector<StreamCacheHook> h;
    return h;
}
}   // namespace

void register_stream_cache(StreamCacheHook h)
{
    std::lock_guard<std::mutex> lk(g_hooks_mutex);
    hooks().push_back(h);
}

size_t Workspace::bytes_held(hipStream_t s)
{
    int dev = 0;
    GUST_HIP(hipGetDevice(&dev));
    size_t total = 0;
    {
        std::lock_guard<std::mutex> lk(g_ws_mutex);
        for (const auto& kv : g_ws) {
            if (std::get<0>(kv.first) == dev && std::get<1>(kv.first) == s) { total += kv.second.cap; }
        }
    }
    std::vector<StreamCacheHook> hs;
    {
        std::lock_guard<std::mutex> lk(g_hooks_mutex);
        hs = hooks();
    }
    for (const auto& h : hs) { total += h.bytes(dev, s); }
    return total;
}

void Workspace::release(hipStream_t s)
{
    int dev = 0;
    GUST_HIP(hipGetDevice(&dev));
    GUST_HIP(hipStreamSynchronize(s));
    {
        std::lock_guard<std::mutex> lk(g_ws_mutex);
        for (auto it = g_ws.begin(); it != g_ws.end();) {
            if (std::get<0>(it->first) == dev && std::get<1>(it->first) == s) {
                if (it->second.p) { GUST_HIP(hipFree(it->second.p)); }
                it = g_ws.erase(it);
            } else {
                ++it;
            }
        }
    }
    std::vector<StreamCacheHook> hs;
    {
        std::lock_guard<std::mutex> lk(g_hooks_mutex);
        hs = hooks();
    }
    for (const auto& h : hs) { h.drop(dev, s); }
}

const DeviceInfo& device_info()
{
    static DeviceInfo info = [] {
        DeviceInfo d;
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess) {
            hipDeviceProp_t p;
            if (hipGetDeviceProperties(&p, dev) == hipSuccess) {
                d.cus = p.multiProcessorCount;
                d.lds_per_block = p.sharedMemPerBlock;
            }
        }
        return d;
    }();
    return info;
}

}   // namespace gust_amd

extern "C" {

const char* gust_last_error(void) { return gust_amd::t_err.c_str(); }

const char* gust_version(void) { return "gust_amd 0.1 (gfx950)"; }

int gust_stream_bytes_held(void* stream, uint64_t* bytes)
{
    return gust_amd::guarded([&] {
        if (!bytes) { GUST_FAIL("gust: stream_bytes_held: bytes is null"); }
        *bytes = gust_amd::Workspace::bytes_held((hipStream_t)stream);
    });
}

int gust_release_stream(void* stream)
{
    return gust_amd::guarded([&] { gust_amd::Workspace::release((hipStream_t)stream); });
}

}
