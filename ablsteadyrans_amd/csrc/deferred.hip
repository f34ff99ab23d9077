// deferred.hip -- the slot initialisation of deferred.hpp's batches.
#include "deferred.hpp"

namespace gust_amd {

namespace {

__global__ void deferred_init_kernel(unsigned long long* p)
{
    const int t = threadIdx.x;
    p[t] = t < Deferred::kZero ? 0ull : ~0ull;
}

}   // namespace

// A fill kernel, not a memset: a batch may be recorded into a captured graph.
void Deferred::init()
{
    if (inited_) { return; }
    constexpr int n = kZero + kOnes;
    dev_ = static_cast<unsigned long long*>(Workspace::get(st_, 126, n * sizeof(unsigned long long)));
    host_ = static_cast<unsigned long long*>(Workspace::get_host_pinned(126, n * sizeof(unsigned long long)));
    deferred_init_kernel<<<1, n, 0, st_>>>(dev_);
    GUST_HIP(hipGetLastError());
    inited_ = true;
}

}   // namespace gust_amd
