// stage.hpp -- the SIMPLEC stage's state on a tile (simplec_stage.hip), shared
// with the Newton stage (newton.hip), which builds its preconditioner and its
// residual evaluations on the same fields, mesh views and Poisson solver.
#pragma once

#include "cgrid.hpp"
#include "deferred.hpp"

#include <exception>

namespace gust_amd {

// A device field owned by the stage.
struct DevField {
    double* p = nullptr;
    gust_array4 v{};
    DevField() = default;
    DevField(const DevField&) = delete;
    DevField& operator=(const DevField&) = delete;
    ~DevField()
    {
        if (p) { (void)hipFree(p); }
    }
    // Zeroed on the stream that will use the field (stream order, not the
    // null stream's: the library's streams need not wait for it).
    void alloc(const IBox& b, int ncomp, hipStream_t st)
    {
        const int64_t n = b.npts();
        GUST_HIP(hipMalloc(&p, size_t(n) * ncomp * sizeof(double)));
        GUST_HIP(hipMemsetAsync(p, 0, size_t(n) * ncomp * sizeof(double), st));
        v.p = p;
        v.jstride = b.len(0);
        v.kstride = int64_t(b.len(0)) * b.len(1);
        v.nstride = n;
        for (int d = 0; d < 3; ++d) {
            v.begin[d] = b.lo[d];
            v.end[d] = b.hi[d] + 1;
        }
        v.ncomp = ncomp;
    }
};

// The locations of the six state fields (u, v, w, p, theta, k).
constexpr int kStageLocs[6] = {GUST_LOC_XFACE, GUST_LOC_YFACE, GUST_LOC_ZFACE, GUST_LOC_CELL, GUST_LOC_CELL,
                               GUST_LOC_CELL};

struct SimplecStage {
    gust_geometry g{};
    gust_array4 z{}, J{}, ax{}, ay{}, az{};
    gust_mesh mesh{};
    gust_physics phys{};
    gust_simplec_options opt{};
    gust_poisson* solver = nullptr;
    int32_t pbc[6]{};
    bool wall = false;
    DevField R[6], cf[7], ss, F[3], dtf[3], A[3], At, Ak, mask[3], ap[3], d[3], conv0, conv1, dtau;
    gust_closure_fields cfv{};
    double res[6]{}, mon[3]{};
    int32_t nmon = 0;
    // One iteration captured as a HIP graph (its kernels and the copy of
    // its checks), the checks' callbacks and the info they fill.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    Deferred::Recorded rec;
    gust_simplec_info ginfo{};

    void drop_graph()
    {
        if (exec) { (void)hipGraphExecDestroy(exec); }
        if (graph) { (void)hipGraphDestroy(graph); }
        exec = nullptr;
        graph = nullptr;
        rec = Deferred::Recorded{};
    }

    ~SimplecStage()
    {
        drop_graph();
        if (solver) { gust_poisson_destroy(solver); }
    }

    gust_residuals residuals() const
    {
        gust_residuals r;
        for (int n = 0; n < 6; ++n) { r.R[n] = &R[n].v; }
        return r;
    }
};


// One SIMPLEC iteration of the stage on q (its checks and reported values
// read back together at its end); defined in simplec_stage.hip.
void simplec_iterate(SimplecStage& s, gust_state* q, const gust_array4* dtau, gust_simplec_info* info,
                     void* stream);

// A failed entry point of the library (a nonzero status) as an exception.
void check_rc(int rc);

}   // namespace gust_amd
