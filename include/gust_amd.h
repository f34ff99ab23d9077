/*
 * gust_amd.h -- C-ABI of the MI355X-native SIMPLEC hot path of GUST
 * (hgopalan/ablsteadyrans).
 *
 * Every entry point replaces one reference interface, cited per declaration
 * as file:line in the reference tree. The AMReX-side shim that binds them
 * (MultiFab/MFIter -> gust_array4 per box) is shown in INTEGRATION.md.
 *
 * Conventions
 *  - Status: 0 = success, nonzero = failure; gust_last_error() then returns
 *    the reference's message text ("gust: <op>: ..."), per calling thread.
 *    Nothing here throws or aborts.
 *  - Data: gust_array4 mirrors amrex::Array4 (i fastest, begin inclusive,
 *    end exclusive, ghosts included in [begin, end)). Pointers are device
 *    pointers valid on `stream`; element type is double unless the entry
 *    point ends in _f32 (float).
 *  - A tile is one box spanning its whole domain; gust_geometry names the
 *    cell-centred domain box, the periodic axes (x/y) and the spacing.
 *  - Streams: every entry point enqueues on `stream` (a hipStream_t). Entry
 *    points returning a host scalar synchronise that stream (as the
 *    reference's reductions do); the *_async variants do not.
 *  - Handles and per-stream workspaces are not thread-safe: one caller per
 *    stream at a time.
 */
#ifndef GUST_AMD_H_
#define GUST_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* amrex::Array4<Real> (AMReX_Array4.H): p, jstride, kstride, nstride,
 * begin, end, ncomp. */
typedef struct gust_array4 {
    void* p;
    int64_t jstride, kstride, nstride;
    int32_t begin[3], end[3];
    int32_t ncomp;
} gust_array4;

/* The parts of amrex::Geometry the hot path reads: the cell-centred domain
 * box (inclusive), periodicity in x and y, and dx, dy, dzeta. */
typedef struct gust_geometry {
    int32_t lo[3], hi[3];
    int32_t periodic[3];
    double dx[3];
} gust_geometry;

/* Field locations, Source/CVFluxes.H:22 (enum class Loc). */
enum { GUST_LOC_CELL = 0, GUST_LOC_XFACE = 1, GUST_LOC_YFACE = 2, GUST_LOC_ZFACE = 4 };

/* Coefficient components of A, Source/ImplicitOperator.H:57 (ImplicitCoef). */
enum { GUST_COEF_P = 0, GUST_COEF_XM, GUST_COEF_XP, GUST_COEF_YM, GUST_COEF_YP, GUST_COEF_ZM, GUST_COEF_ZP,
       GUST_N_COEFS };

/* Source/LineSweeps.H:80-104 (SweepKind, LineGhosts, LineOrder,
 * LineSweepOptions, max_line_overlap). */
enum { GUST_SWEEP_ZJACOBI = 0, GUST_SWEEP_LUSGS = 1 };
enum { GUST_GHOSTS_LAGGED = 0, GUST_GHOSTS_ZERO = 1 };
enum { GUST_ORDER_SEQUENTIAL = 0, GUST_ORDER_DIAGONAL = 1 };
#define GUST_MAX_LINE_OVERLAP 8

typedef struct gust_line_options {
    int32_t kind;       /* GUST_SWEEP_* */
    int32_t ghosts;     /* GUST_GHOSTS_* */
    int32_t overlap;    /* 0..GUST_MAX_LINE_OVERLAP, zero ghosts only */
    int32_t order;      /* GUST_ORDER_*; both give bit-identical results */
    int32_t alternate;  /* 0 or 1: odd sweeps take the j-reversed pair */
} gust_line_options;

/* One line system of a batch: a field at `loc` on a tile. */
typedef struct gust_line_system {
    int32_t loc;
    gust_geometry geom;
    gust_array4 A;      /* GUST_N_COEFS components at loc, no ghosts needed */
    gust_array4 ap;     /* 1 component at loc */
    gust_array4 rhs;    /* 1 component at loc */
    gust_array4 delta;  /* 1 component at loc, >= 1 ghost cell; output */
} gust_line_system;

/* The message of the calling thread's last failure ("" if none). */
const char* gust_last_error(void);

/* Library version string. */
const char* gust_version(void);

/* Device memory the library holds for one stream on the current device
 * (scratch workspaces and per-stream table caches), in bytes. No counterpart
 * in the reference (AMReX's arena owns its scratch); for callers that budget
 * HBM per stream, such as a server admitting requests. */
int gust_stream_bytes_held(void* stream, uint64_t* bytes);

/* Synchronises the stream and frees every buffer and cache the library holds
 * for it on the current device (the next call on that stream allocates
 * afresh). No counterpart in the reference; for callers that retire streams
 * (a server's worker exiting). Handles created on the stream (Poisson
 * solvers, SIMPLEC stages) keep their own memory until destroyed. */
int gust_release_stream(void* stream);

/* z-line Jacobi sweeps of one system.
 * Replaces: Real zline_jacobi(Loc, const MultiFab& A, const MultiFab& ap,
 *   const MultiFab& rhs, int sweeps, const Mesh&, const Geometry&,
 *   MultiFab& delta)  -- Source/Predictor.H:267-268, Source/Predictor.cpp:87-127.
 * delta is zeroed (ghosts included) and receives the iterate; *change =
 * max |change over the last sweep| (synchronises the stream). */
int gust_zline_jacobi(int loc, const gust_array4* A, const gust_array4* ap, const gust_array4* rhs, int sweeps,
                      const gust_geometry* geom, gust_array4* delta, double* change, void* stream);

/* gust_zline_jacobi that allocates nothing: no hipMalloc, hipFree, pinned
 * host memory or allocating library routine. No counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture
 * set-up, tight serving loops, harnesses that forbid allocation). Same
 * arguments, results and refusals as gust_zline_jacobi, bit for bit. Its
 * scratch (the change accumulator, the system table and the iterate) is a
 * fixed 64 MiB array per device held until the process ends; systems that
 * need more (about 8 bytes per unknown plus 512), or more than 320 levels,
 * are refused with an error. Calls that overlap on one device share that
 * array and race; synchronises `stream` before returning. */
int gust_zline_jacobi_noalloc(int loc, const gust_array4* A, const gust_array4* ap, const gust_array4* rhs,
                              int sweeps, const gust_geometry* geom, gust_array4* delta, double* change, void* stream);

/* The same on float data (the mixed-precision preconditioner path);
 * *change is returned as double. */
int gust_zline_jacobi_f32(int loc, const gust_array4* A, const gust_array4* ap, const gust_array4* rhs, int sweeps,
                          const gust_geometry* geom, gust_array4* delta, double* change, void* stream);

/* Batched z-line Jacobi: nsys systems (any locations and tiles) in one
 * launch per sweep. changes_dev (device, nsys doubles, may be NULL) receives
 * each system's last-sweep change; nothing is synchronised.
 * Replaces: the per-box ParallelFor loop of zline_jacobi
 *   (Source/Predictor.cpp:97-114) over all boxes and the five variables of
 *   an iteration (Source/Predictor.cpp:400, Source/Simplec.cpp:500-502). */
int gust_zline_jacobi_batched_async(int nsys, const gust_line_system* sys, int sweeps, double* changes_dev,
                                    void* stream);
int gust_zline_jacobi_batched_async_f32(int nsys, const gust_line_system* sys, int sweeps, double* changes_dev,
                                        void* stream);

/* A batched Jacobi plan: the systems' device table, scratch iterate and
 * change accumulators, built once (allocates; synchronises `stream` once).
 * Running a plan enqueues kernels only -- no allocation, copy, memset or
 * host synchronisation -- so it can be captured in a HIP graph and replayed
 * every iteration. The views' memory must outlive the plan.
 * Replaces: the same per-box sweep loop as gust_zline_jacobi_batched_async,
 *   with the setup hoisted out of the iteration (the reference re-allocates
 *   its three scratch MultiFabs per call, Source/Predictor.cpp:92-93).
 * dtype: 0 = double data, 1 = float data. */
typedef struct gust_jacobi_plan gust_jacobi_plan;
int gust_jacobi_plan_create(int nsys, const gust_line_system* sys, int dtype, gust_jacobi_plan** plan,
                            void* stream);
/* Zero delta's ghosts, `sweeps` sweeps, the seam sync; changes_dev as in
 * gust_zline_jacobi_batched_async. */
int gust_jacobi_plan_run(gust_jacobi_plan* plan, int sweeps, double* changes_dev, void* stream);
/* One sweep kernel alone (first: the input iterate is zero; last: record the
 * change; into_scratch: read delta and write the scratch iterate, else the
 * reverse). For timing one sweep and for callers composing their own loop. */
int gust_jacobi_plan_sweep(gust_jacobi_plan* plan, int first, int last, int into_scratch, void* stream);
/* Unknowns and 64-column tiles of the plan. */
int gust_jacobi_plan_info(const gust_jacobi_plan* plan, int64_t* unknowns, int32_t* tiles);
/* The z-line Jacobi kernel schedule for plans made after the call (process
 * wide): 0 autotuned (the default: the first line solve of a set of system
 * shapes and sweep count times every candidate -- the LDS-staged kernel,
 * the streaming kernel with c' in LDS at prefetch depths 2, 4, 8, the
 * streaming kernel with c' in global memory, the register kernel with e'
 * held in registers and c' in LDS, in registers (FP32 or short FP64
 * columns) or split between them (FP64 columns of 66-97 levels) -- over all
 * its sweeps with HIP events and keeps the fastest for those shapes; under
 * graph capture an untimed shape takes rule 3), 1 streaming always, 2
 * staged wherever a tile fits the LDS, 3 the fixed rule (fewer than 1024
 * tiles staged, more streamed), 4 the register kernel (columns of up to 97
 * levels; taller ones stream). Plan handles (gust_jacobi_plan_create) tune
 * under 0 at their first run outside graph capture.
 * Every schedule gives the same bits.
 * No counterpart in the reference (a launch-shape choice of this library). */
int gust_jacobi_set_schedule(int32_t schedule);

/* The autotuner's records: *count = their number; record index (in key
 * order) gives the dtype (0 f64, 1 f32), sweeps, the number of systems, the
 * chosen candidate, the eight candidates' times in ms (-1: did not apply)
 * and the choice's name. gust_jacobi_tune_reset forgets them.
 * No counterpart in the reference. */
int gust_jacobi_tuned(int32_t index, int32_t* count, int32_t* dtype, int32_t* sweeps, int32_t* nsys,
                      int32_t* choice, float* ms, char* name, int64_t cap);
int gust_jacobi_tune_reset(void);
void gust_jacobi_plan_destroy(gust_jacobi_plan* plan);

/* Line LU-SGS sweeps of one system with the given options.
 * Replaces: Real zline_lusgs(Loc, A, ap, rhs, int sweeps, const Mesh&,
 *   const Geometry&, MultiFab& delta, const LineSweepOptions&)
 *   -- Source/LineSweeps.H:159-161, Source/LineSweeps.cpp:415-428.
 * The sequential and diagonal orders give bit-identical results; on the
 * GPU both run as diagonal wavefronts inside one workgroup per system. */
int gust_zline_lusgs(int loc, const gust_array4* A, const gust_array4* ap, const gust_array4* rhs, int sweeps,
                     const gust_geometry* geom, gust_array4* delta, const gust_line_options* opt, double* change,
                     void* stream);

/* Dispatcher. Replaces: Real line_sweeps(const LineSweepOptions&, Loc, ...)
 *   -- Source/Predictor.H:257-259, Source/Predictor.cpp:79-85. */
int gust_line_sweeps(const gust_line_options* opt, int loc, const gust_array4* A, const gust_array4* ap,
                     const gust_array4* rhs, int sweeps, const gust_geometry* geom, gust_array4* delta,
                     double* change, void* stream);

/* The option checks, Source/LineSweeps.cpp:315-329 (check_line_options):
 * 0 if valid, else nonzero with the message naming the steady.<key>. */
int gust_check_line_options(const gust_line_options* opt);

/* ------------------------------------------------------------------------
 * The terrain Poisson solve (pressure correction, initializer projection).
 *
 * Boundary conditions: 6 codes in amrex::Orientation order xlo, ylo, zlo,
 * xhi, yhi, zhi (Source/Poisson.H:20-34, PoissonFace/PoissonBC). */
enum { GUST_PF_PERIODIC = 0, GUST_PF_NEUMANN = 1, GUST_PF_DIRICHLET = 2 };

/* Source/Poisson.H:40-69 (PoissonOptions), with the multigrid's own knobs.
 * precision: 0 = FP64 V-cycle, 1 = FP32 V-cycle under FP64 GMRES. */
typedef struct gust_poisson_options {
    double reltol;         /* solver.poisson_reltol, default 1e-10 */
    double abstol;         /* solver.poisson_abstol, default 0 */
    int32_t max_iter;      /* solver.poisson_max_iter, default 200 */
    int32_t restart;       /* solver.poisson_restart, default 50 */
    int32_t verbose;
    int32_t precision;
    int32_t pre_sweeps, post_sweeps, coarse_sweeps;   /* V(2,2), 16 on the coarsest level */
    int32_t max_levels;    /* multigrid levels at most (0: coarsen while x or y halves) */
    int32_t fixed_steps;   /* 0: GMRES(restart) to the tolerance, reading each Arnoldi step on the host;
                              k in 1..200: k Arnoldi steps on the device, no host reads (a captured graph
                              can hold the solve), the tolerance checked at the end */
    int32_t seam_copy;     /* a level whose periodic seam joins two columns of one colour (an odd periodic
                              size): 1 (default), its seam columns read those neighbours from a copy, one
                              launch per half sweep; 0, a pass of its own forms each colour's neighbour
                              terms first. The same arithmetic, the same bits. */
} gust_poisson_options;

/* Source/Poisson.H:96-103 (FaceCoeffs): optional face coefficients (null =
 * 1), each a face field on the domain's faces; precond_ratio is accepted for
 * the interface and checked, the multigrid does not need it. */
typedef struct gust_face_coeffs {
    const gust_array4* x;
    const gust_array4* y;
    const gust_array4* z;
    const gust_array4* z_cross;
    double precond_ratio;
} gust_face_coeffs;

/* Source/Poisson.H:105-117 (PoissonResult). */
typedef struct gust_poisson_result {
    int32_t iterations;
    double initial_residual, final_residual;
    int32_t singular;
    double removed_mean;
} gust_poisson_result;

typedef struct gust_poisson gust_poisson;

/* The documented defaults (Source/Poisson.H:45-65) and V(2,2). */
int gust_poisson_default_options(gust_poisson_options* opt);

/* Replaces: PoissonSolver(const Mesh&, const Geometry&, const MultiFab&
 *   z_phys_nd, const PoissonOptions&) -- Source/Poisson.H:138-139,
 *   Source/Poisson.cpp:395-403. z_nd: node heights with one ghost node;
 *   detJ, ax, ay, az: the mesh metrics on the cells / faces. The solver keeps
 *   its own copy of the mesh and builds the multigrid levels once. */
int gust_poisson_create(const gust_geometry* geom, const gust_array4* z_nd, const gust_array4* detJ,
                        const gust_array4* ax, const gust_array4* ay, const gust_array4* az,
                        const gust_poisson_options* opt, gust_poisson** solver, void* stream);
void gust_poisson_destroy(gust_poisson* solver);
/* Multigrid levels built (diagnostic). */
int gust_poisson_levels(const gust_poisson* solver, int32_t* levels);

/* Replaces: PoissonResult PoissonSolver::solve(MultiFab& rhs, MultiFab& phi,
 *   const FaceCoeffs&, const PoissonBC&, Array<MultiFab,3>* fluxes)
 *   -- Source/Poisson.H:151-152, Source/Poisson.cpp:540-623.
 * rhs: cells (modified when the problem is singular: its J-weighted mean is
 * removed); phi: cells with one ghost cell, zero initial guess, returned with
 * its ghosts filled for bc; fx, fy, fz (all or none): the coefficient-
 * weighted fluxes F = -g on the faces. Fails, naming
 * solver.poisson_max_iter, when the solve does not converge. Synchronises. */
int gust_poisson_solve(gust_poisson* solver, gust_array4* rhs, gust_array4* phi, const gust_face_coeffs* coeffs,
                       const int32_t* bc, gust_array4* fx, gust_array4* fy, gust_array4* fz,
                       gust_poisson_result* result, void* stream);

/* Replaces: PoissonSolver::apply(MultiFab& Aphi, MultiFab& phi, const
 *   FaceCoeffs&, const PoissonBC&) -- Source/Poisson.H:156, Source/Poisson.cpp:517-529
 *   (ERF's terrpoisson_adotx with the zero-flux mask, the Dirichlet ground,
 *   the z_cross correction). phi's ghosts are filled. */
int gust_poisson_apply(gust_poisson* solver, gust_array4* Aphi, gust_array4* phi, const gust_face_coeffs* coeffs,
                       const int32_t* bc, void* stream);

/* Replaces: PoissonSolver::fill_ghosts -- Source/Poisson.H:161, Source/Poisson.cpp:531-538. */
int gust_poisson_fill_ghosts(gust_poisson* solver, gust_array4* phi, const int32_t* bc, void* stream);

/* The fluxes of a given phi, as solve returns them.
 * Replaces: TerrainPoisson::getFluxes + the coefficient multiply and
 *   add_cross_flux of solve -- extern_erf/ERF_TerrainPoisson.cpp:212-264,
 *   Source/Poisson.cpp:609-619. */
int gust_poisson_fluxes(gust_poisson* solver, gust_array4* phi, const gust_face_coeffs* coeffs, const int32_t* bc,
                        gust_array4* fx, gust_array4* fy, gust_array4* fz, void* stream);

/* One multigrid V-cycle x = M^{-1} b on compact device vectors (nx*ny*nz
 * doubles, i fastest); setup = 1 first takes the coefficients. The
 * preconditioner of gust_poisson_solve, exposed for timing and tests.
 * Replaces: the MLMG cycle inside GMRESMLMG (AMReX, third party),
 *   called from Source/Poisson.cpp:681-687. */
int gust_poisson_vcycle(gust_poisson* solver, const double* b, double* x, const gust_face_coeffs* coeffs,
                        const int32_t* bc, int setup, void* stream);

/* Replaces: void face_divergence(const Mesh&, const Geometry&, const
 *   Array<const MultiFab*,3>& F, MultiFab& div) -- Source/Poisson.H:209-210,
 *   Source/Poisson.cpp:733-764. */
int gust_face_divergence(const gust_geometry* geom, const gust_array4* ax, const gust_array4* ay,
                         const gust_array4* az, const gust_array4* detJ, const gust_array4* fx,
                         const gust_array4* fy, const gust_array4* fz, gust_array4* div, void* stream);

/* ---- Mass flux on the terrain-following C-grid (Source/MassFlux.H) ------
 * u on x-faces, v on y-faces, w and Omega on z-faces, p on cells; z_nd the
 * node heights with one ghost node; the mesh's ax, ay, az, detJ. Ghost
 * points that an operator reads are the caller's to fill, as in the
 * reference. */

/* Omega = w - z_x u_f - z_y v_f on the valid z-faces grown by ngrow (0|1);
 * u, v need ngrow + 1 ghosts, w and omega ngrow.
 * Replaces: void omega_from_w(const MultiFab& u, v, w, const MultiFab& z_nd,
 *   const Mesh&, const Geometry&, MultiFab& omega, int ngrow) --
 *   Source/MassFlux.H:49-53, Source/MassFlux.cpp:51-74. */
int gust_omega_from_w(const gust_geometry* geom, const gust_array4* u, const gust_array4* v, const gust_array4* w,
                      const gust_array4* z_nd, gust_array4* omega, int32_t ngrow, void* stream);

/* w = Omega + z_x u_f + z_y v_f on the valid z-faces (u, v: one ghost).
 * Replaces: void w_from_omega(...) -- Source/MassFlux.H:55-58,
 *   Source/MassFlux.cpp:76-99. */
int gust_w_from_omega(const gust_geometry* geom, const gust_array4* u, const gust_array4* v,
                      const gust_array4* omega, const gust_array4* z_nd, gust_array4* w, void* stream);

/* F_x = a_x u, F_y = a_y v, F_z = Omega on the valid faces and one ghost
 * face (u, v: two ghosts, w: one; F: one).
 * Replaces: void transport_fluxes(const MultiFab& u, v, w, z_nd, const Mesh&,
 *   const Geometry&, Array<MultiFab,3>& F) -- Source/MassFlux.H:60-66,
 *   Source/MassFlux.cpp:101-127. */
int gust_transport_fluxes(const gust_geometry* geom, const gust_array4* u, const gust_array4* v,
                          const gust_array4* w, const gust_array4* z_nd, const gust_array4* ax,
                          const gust_array4* ay, gust_array4* Fx, gust_array4* Fy, gust_array4* Fz, void* stream);

/* div = face_divergence(u, v, Omega) on the valid cells (u, v: one ghost).
 * Replaces: void velocity_divergence(const MultiFab& u, v, w, z_nd,
 *   const Mesh&, const Geometry&, MultiFab& div) -- Source/MassFlux.H:68-72,
 *   Source/MassFlux.cpp:129-136. */
int gust_velocity_divergence(const gust_geometry* geom, const gust_array4* u, const gust_array4* v,
                             const gust_array4* w, const gust_array4* z_nd, const gust_array4* ax,
                             const gust_array4* ay, const gust_array4* az, const gust_array4* detJ,
                             gust_array4* div, void* stream);

/* (G p)_x, (G p)_y, (G p)_z on the valid faces (p: one ghost; detJ: one).
 * Replaces: void pressure_gradient(const MultiFab& p, const MultiFab& z_nd,
 *   const Mesh&, const Geometry&, const Array<MultiFab*,3>& Gp) --
 *   Source/MassFlux.H:74-76, Source/MassFlux.cpp:138-179. */
int gust_pressure_gradient(const gust_geometry* geom, const gust_array4* p, const gust_array4* z_nd,
                           const gust_array4* detJ, const gust_array4* ax, const gust_array4* ay, gust_array4* gx,
                           gust_array4* gy, gust_array4* gz, void* stream);

/* ---- SIMPLEC steps (Source/Simplec.H) ---------------------------------- */

/* Source/Simplec.H:116-123 (PressureCorrectionResult). */
typedef struct gust_pressure_correction_result {
    gust_poisson_result poisson;
    double divergence_before;   /* max |div u*| */
    double divergence_after;    /* max |div u| after the correction */
    double max_dp;              /* max |alpha_p p'| */
    double max_du, max_dv, max_dw;
} gust_pressure_correction_result;

/* The pressure correction on (u, v, w, p) in place: beta = 1/d on faces with
 * d > 0 (keep = 1; d = 0: beta = 1, keep = 0; d must be finite and >= 0),
 * rhs = div u* (- source), the Poisson solve for phi with fluxes
 * F = -beta grad phi, p += alpha_p phi, u, v, Omega += keep F, w from Omega,
 * seams and periodic ghosts of u, v, w refreshed, div after. u, v, w and p
 * need one ghost (filled). solver: a gust_poisson on the same mesh (its
 * tolerance is the correction's). source may be null. Aborts (returns
 * nonzero) on a net source in the singular case unless project_net_source.
 * Replaces: PressureCorrectionResult pressure_correction(State& q, const
 *   PredictorFields& d, PoissonSolver&, const PoissonBC&, Real alpha_p, const
 *   Mesh&, const Geometry&, SimplecWorkspace&, const MultiFab* source, bool
 *   project_net_source) -- Source/Simplec.H:139-142, Source/Simplec.cpp:134-220. */
int gust_pressure_correction(gust_poisson* solver, const gust_geometry* geom, gust_array4* u, gust_array4* v,
                             gust_array4* w, gust_array4* p, const gust_array4* z_nd, const gust_array4* ax,
                             const gust_array4* ay, const gust_array4* az, const gust_array4* detJ,
                             const gust_array4* d_u, const gust_array4* d_v, const gust_array4* d_w,
                             const int32_t* bc, double alpha_p, const gust_array4* source,
                             int32_t project_net_source, gust_pressure_correction_result* result, void* stream);

/* ---- The implicit operator A1 (Source/ImplicitOperator.H) -------------- */

/* A1's seven coefficients (GUST_COEF_* components of A at loc, valid
 * points): first-order bounded upwind convection from the transport fluxes
 * F (one ghost) plus the principal diffusion with the face mean of the cell
 * diffusivity g (one ghost); face_mean 0 arithmetic, 1 logarithmic (z-faces).
 * Replaces: void implicit_coefficients(Loc, const Array<const MultiFab*,3>& F,
 *   const MultiFab& g, const MultiFab& z_nd, const Mesh&, const Geometry&,
 *   MultiFab& A, FaceMean) -- Source/ImplicitOperator.H:59-65,
 *   Source/ImplicitOperator.cpp:95-121. */
int gust_implicit_coefficients(int32_t loc, const gust_geometry* geom, const gust_array4* Fx, const gust_array4* Fy,
                               const gust_array4* Fz, const gust_array4* g, const gust_array4* z_nd,
                               const gust_array4* detJ, int32_t face_mean, gust_array4* A, void* stream);

/* out = A1 q on the valid points at loc (q: one ghost).
 * Replaces: void apply_implicit(Loc, const MultiFab& A, const MultiFab& q,
 *   const Mesh&, MultiFab& out) -- Source/ImplicitOperator.H:69-70,
 *   Source/ImplicitOperator.cpp:123-135. */
int gust_apply_implicit(int32_t loc, const gust_geometry* geom, const gust_array4* A, const gust_array4* q,
                        gust_array4* out, void* stream);

/* ---- Convection and diffusion over control volumes ----------------------
 * (Source/Convection.H, Source/Diffusion.H, Source/CVFluxes.H). A field at
 * loc owns the control volume centred on its points; G[d] are the fluxes on
 * that volume's d-faces (location loc ^ (1 << d)) covering cv_flux_box
 * (d-centres [lo-1, hi+1] where loc is nodal along d, else d-nodes
 * [lo, hi+1]). Limiter: 0 van Leer, 1 van Albada, 2 upwind
 * (Source/Limiter_K.H:19). face_mean: 0 arithmetic, 1 logarithmic. */
enum { GUST_LIMITER_VAN_LEER = 0, GUST_LIMITER_VAN_ALBADA = 1, GUST_LIMITER_UPWIND = 2 };

/* Convective fluxes F q_face (MUSCL, upwind side; q: two ghosts; F: one).
 * Replaces: void convective_fluxes(Loc, const MultiFab& q, const
 *   Array<const MultiFab*,3>& F, const ConvectionOptions&, const Mesh&,
 *   Array<MultiFab,3>& G) -- Source/Convection.H:81-84, Convection.cpp:117-132. */
int gust_convective_fluxes(int32_t loc, const gust_geometry* geom, const gust_array4* q, const gust_array4* Fx,
                           const gust_array4* Fy, const gust_array4* Fz, int32_t limiter, double smooth_eps,
                           gust_array4* Gx, gust_array4* Gy, gust_array4* Gz, void* stream);

/* out = (1/J) [d_xi G_x + d_eta G_y + d_zeta G_z] on the valid points.
 * Replaces: void flux_divergence(Loc, const Array<MultiFab,3>& G, const Mesh&,
 *   const Geometry&, MultiFab& out) -- Source/CVFluxes.H:46-53, CVFluxes.cpp:69-80. */
int gust_flux_divergence(int32_t loc, const gust_geometry* geom, const gust_array4* Gx, const gust_array4* Gy,
                         const gust_array4* Gz, const gust_array4* detJ, gust_array4* out, void* stream);

/* C(q) (bounded: C(q) - q C(1)) in one fused pass.
 * Replaces: void convection(Loc, const MultiFab& q, const Array<const
 *   MultiFab*,3>& F, const ConvectionOptions&, const Mesh&, const Geometry&,
 *   MultiFab& out) -- Source/Convection.H:86-90, Convection.cpp:134-141. */
int gust_convection(int32_t loc, const gust_geometry* geom, const gust_array4* q, const gust_array4* Fx,
                    const gust_array4* Fy, const gust_array4* Fz, const gust_array4* detJ, int32_t limiter,
                    double smooth_eps, int32_t bounded, gust_array4* out, void* stream);

/* gamma (grad q . A) on the cell faces (q, gamma: one ghost).
 * Replaces: void scalar_diffusive_fluxes(...) -- Source/Diffusion.H:40-42,
 *   Diffusion.cpp:63-77. */
int gust_scalar_diffusive_fluxes(const gust_geometry* geom, const gust_array4* q, const gust_array4* gamma,
                                 const gust_array4* z_nd, int32_t face_mean, gust_array4* Gx, gust_array4* Gy,
                                 gust_array4* Gz, void* stream);

/* D(q) on the cells, fused.
 * Replaces: void scalar_diffusion(...) -- Source/Diffusion.H:45-46,
 *   Diffusion.cpp:79-85. */
int gust_scalar_diffusion(const gust_geometry* geom, const gust_array4* q, const gust_array4* gamma,
                          const gust_array4* z_nd, const gust_array4* detJ, int32_t face_mean, gust_array4* out,
                          void* stream);

/* tau_c . A of c-momentum on the control-volume faces of component c (u, v,
 * w: two ghosts; nu: one).
 * Replaces: void momentum_diffusive_fluxes(...) -- Source/Diffusion.H:51-54,
 *   Diffusion.cpp:87-109. */
int gust_momentum_diffusive_fluxes(int32_t c, const gust_geometry* geom, const gust_array4* u, const gust_array4* v,
                                   const gust_array4* w, const gust_array4* nu, const gust_array4* z_nd,
                                   int32_t face_mean, gust_array4* Gx, gust_array4* Gy, gust_array4* Gz,
                                   void* stream);

/* D_c(u) on the c-faces for c = 0, 1, 2, fused.
 * Replaces: void momentum_diffusion(...) -- Source/Diffusion.H:57-60,
 *   Diffusion.cpp:111-119. */
int gust_momentum_diffusion(const gust_geometry* geom, const gust_array4* u, const gust_array4* v,
                            const gust_array4* w, const gust_array4* nu, const gust_array4* z_nd,
                            const gust_array4* detJ, int32_t face_mean, gust_array4* out_u, gust_array4* out_v,
                            gust_array4* out_w, void* stream);

/* One field per velocity component: x-faces, y-faces, z-faces. */
typedef struct gust_faces {
    gust_array4* c[3];
} gust_faces;

/* The momentum predictor's settings (Source/Predictor.H:242-252,
 * Source/Simplec.H:26-40): line sweeps, simplec.rotation_relax,
 * simplec.w_vnn_cap and _slope_margin, the Coriolis parameter f (0: off;
 * coriolis_parameter(in.coriolis)), gravity (0: off) and theta0. */
typedef struct gust_predictor_options {
    int32_t sweeps;
    double rotation_relax;
    double w_vnn_cap, w_vnn_slope_margin;
    double coriolis_f;
    double gravity, theta0;
    gust_line_options sweep;
    /* 0: the line sweeps in FP64 (the reference); 1: mixed precision, the
     * sweeps' operands (A, a_P, rhs) rounded to FP32 and solved in FP32 (z-line
     * Jacobi or line LU-SGS), the change added to the FP64 field (a defect
     * correction: R stays FP64, so the converged field is the FP64 one). */
    int32_t sweep_precision;
} gust_predictor_options;

/* Source/Predictor.H:214-220 (MomentumPredictorResult). */
typedef struct gust_momentum_predictor_result {
    double max_du, max_dv, max_dw;
    double last_sweep_du, last_sweep_dv, last_sweep_dw;
} gust_momentum_predictor_result;

/* The momentum predictor: for u, v, w in turn, a_P = keep/dtau + A_P,
 * d = keep (a_P - sum a_nb) (keep = 1 - the Dirichlet-row mask), the w
 * rows' VNN cap (plain or slope-aware), rhs = -R, the point-implicit rotation
 * (per-level mean of the other residual at the inertial step, the
 * fluctuation at the row's own step) and buoyancy, the line sweeps, and
 * q += delta. A holds each component's implicit rows (implicit_rows: A1 with
 * the boundary conditions' folds; 7 components); mask holds the Dirichlet
 * rows (null: none; required with rotation or buoyancy). ap and d receive the
 * SIMPLEC fields the pressure correction uses.
 * Replaces: MomentumPredictorResult momentum_predictor(State& q, const
 *   ResidualFields& R, const Array<const MultiFab*,3>& F, const
 *   ResidualInputs&, const Array<MultiFab,3>& dtau_faces, const Mesh&, const
 *   Geometry&, int sweeps, PredictorFields& ap, const SurfaceState*, const
 *   ClosureFields*, Real rotation_relax, const LineSweepOptions&, Real
 *   w_vnn_cap, Real w_vnn_slope_margin) -- Source/Predictor.H:242-252,
 *   Source/Predictor.cpp:129-405. */
int gust_momentum_predictor(const gust_geometry* geom, gust_faces* vel, const gust_array4* theta,
                            const gust_array4* z_nd, const gust_faces* R, const gust_array4* R_theta,
                            const gust_faces* A, const gust_array4* A_theta, const gust_faces* dtau,
                            const gust_faces* mask, const gust_predictor_options* opt, gust_faces* ap, gust_faces* d,
                            gust_momentum_predictor_result* result, void* stream);

/* Equations of the scalar sweep (Source/Residual.H, Equation). */
enum { GUST_EQ_THETA = 0, GUST_EQ_TKE = 1 };

/* Source/Simplec.H:144-149 (ScalarSweepResult). */
typedef struct gust_scalar_sweep_result {
    double max_dq;
    double last_sweep;
    int64_t clipped;
} gust_scalar_sweep_result;

/* One z-line step of theta or k in delta form: (1/dtau + A1) dq = -R, the
 * sweeps, field += dq; k clipped at floor when clip (counted). A: the
 * equation's implicit rows (7 components).
 * Replaces: ScalarSweepResult scalar_sweep(Equation, MultiFab& field, const
 *   MultiFab& R_eq, const Array<const MultiFab*,3>& F, const MultiFab& z_nd,
 *   const ResidualInputs&, const MultiFab& dtau, const Mesh&, const Geometry&,
 *   int sweeps, Real floor, const SurfaceState*, const ClosureFields*,
 *   SimplecWorkspace&, const LineSweepOptions&, bool clip) --
 *   Source/Simplec.H:158-163, Source/Simplec.cpp:222-281. */
int gust_scalar_sweep(const gust_geometry* geom, int32_t eq, gust_array4* field, const gust_array4* R_eq,
                      const gust_array4* A, const gust_array4* dtau, int32_t sweeps, double floor, int32_t clip,
                      const gust_line_options* sweep, gust_scalar_sweep_result* result, void* stream);

/* ---- Pseudo-time (Source/PseudoTime.H, Source/Predictor.H) ------------- */

/* dtau = cfl / max(|u_c|/dx + |v_c|/dy + |Omega_c|/(J dzeta) + 2 nu (1/dx^2 +
 * 1/dy^2), 1e-12) on the valid cells; u, v need one ghost; nu may be null.
 * Replaces: void local_dtau(const State& q, const MultiFab* nu, const Mesh&,
 *   const Geometry&, Real cfl, MultiFab& dtau) -- Source/PseudoTime.H:35-36,
 *   Source/PseudoTime.cpp:12-51. */
int gust_local_dtau(const gust_geometry* geom, const gust_array4* u, const gust_array4* v, const gust_array4* w,
                    const gust_array4* z_nd, const gust_array4* detJ, const gust_array4* nu, double cfl,
                    gust_array4* dtau, void* stream);

/* min(cfl_max, growth cfl, cfl0 r^-p) with the reference's guards (host).
 * Replaces: Real ser_cfl(Real cfl, Real r, Real cfl0, Real cfl_max, Real p,
 *   Real growth) -- Source/PseudoTime.H:47-48, Source/PseudoTime.cpp:53-61. */
int gust_ser_cfl(double cfl, double r, double cfl0, double cfl_max, double p, double growth, double* out);

/* Face means of a cell dtau (one-sided on non-periodic boundary faces);
 * refuses a dtau_cc that is not finite and > 0 everywhere (synchronises).
 * Replaces: void face_dtau(const MultiFab& dtau_cc, const Mesh&, const
 *   Geometry&, Array<MultiFab,3>& dtau_faces) -- Source/Predictor.cpp:30-77. */
int gust_face_dtau(const gust_geometry* geom, const gust_array4* dtau_cc, gust_array4* dtau_x, gust_array4* dtau_y,
                   gust_array4* dtau_z, void* stream);

/* ---- Reductions with rank- and order-independent bits ------------------ */

/* Sum of component comp of a cell-centred field over the tile's valid cells:
 * one device thread per (i, j) column adds its values in k order; the host
 * adds the column sums in storage order (i fastest, then j). A MultiFab of
 * several boxes calls this per box and adds the box sums in box-index order
 * after the all-reduce of one slot per box (the shim's part). Synchronises.
 * Replaces: Real deterministic_sum(const MultiFab&, int comp) --
 *   Source/Reductions.H:19, Source/Reductions.cpp:11-53 (one box's slot). */
int gust_deterministic_sum(const gust_geometry* geom, const gust_array4* a, int32_t comp, double* sum,
                           void* stream);

/* Maximum of component comp over the valid points at loc: NaN if any valid
 * value is NaN, zero reported as +0. Synchronises.
 * Replaces: Real global_max(const MultiFab&, int comp) --
 *   Source/Reductions.H:24, Source/Reductions.cpp:55-66. */
int gust_global_max(const gust_geometry* geom, int32_t loc, const gust_array4* a, int32_t comp, double* max,
                    void* stream);

/* Over the owned points at loc (on a periodic axis along which loc is
 * nodal, the hi face is the lo face's copy and is skipped): sums[0] = sum
 * |(q - offset) / s|, sums[1] = sum (v / s)^2, sums[2] = the point count;
 * each column in k order on the device, the columns in storage order on
 * the host. offset may be null. Synchronises.
 * Replaces: template <bool OFFSET> void scaled_sums(const MultiFab& q, const
 *   MultiFab& v, const MultiFab* offset, Real s, const Periodicity&,
 *   Vector<double>& box_sums) -- Source/Residual.cpp:101-150 (one box). */
int gust_scaled_sums(const gust_geometry* geom, int32_t loc, const gust_array4* q, const gust_array4* v,
                     const gust_array4* offset, double s, double sums[3], void* stream);

/* Typical field magnitudes, Source/Residual.H:72-83 (JvpScales; defaults
 * 10 m/s, 10 K, 0.1 m^2/s^2, 100 m^2/s^2, sqrt(DBL_EPSILON)). */
typedef struct gust_jvp_scales {
    double velocity, theta, tke, pressure, step;
} gust_jvp_scales;

/* scale[n] for u, v, w, p, theta, k: U^2/H three times, U/H, U theta_s/H,
 * U k_s/H, U = velocity.
 * Replaces: residual_scales(const JvpScales&, Real H, Real scale[6]) --
 *   Source/Simplec.cpp:374-382. */
int gust_residual_scales(const gust_jvp_scales* s, double height, double scale[6]);

/* The depth of the z-levels, or the box height when the two agree to
 * 1000 eps of the box height.
 * Replaces: Real residual_scale_height(const Mesh&, const Geometry&) --
 *   Source/Mesh.cpp:332-340. */
int gust_residual_scale_height(const double* zlevels, int32_t nlevels, double prob_lo_z, double prob_hi_z,
                               double* height);

/* The six residual fields without ghosts: u, v, w (x-, y-, z-faces), p,
 * theta, k (cells); Source/Residual.H:61-68 (ResidualFields). */
typedef struct gust_residuals {
    const gust_array4* R[6];
} gust_residuals;

/* res[n] = sqrt(sum R_n^2) / sqrt(npts_n) / scale[n] over the valid points
 * at the field's location (both copies of a periodic face, as numPts);
 * mon[0] = max |u|, |v|, |w| over their valid faces; mon[1] = the mean of k
 * over the cells; mon[2] = the mean of u* over the ground slab (nx x ny
 * cells at the domain's lowest k) when ustar is not null. *nmon = 2 or 3.
 * The sums use a fixed reduction tree: the same tile gives the same bits on
 * every call. Synchronises.
 * Replaces: void stage_norms(const ResidualFields&, const State&, const
 *   SurfaceState*, const JvpScales&, Real height, const Geometry&,
 *   vector<Real>& res, vector<Real>& mon) -- Source/Simplec.cpp:385-409. */
int gust_stage_norms(const gust_geometry* geom, const gust_residuals* R, const gust_array4* u, const gust_array4* v,
                     const gust_array4* w, const gust_array4* tke, const gust_array4* ustar,
                     const gust_jvp_scales* s, double height, double res[6], double mon[3], int32_t* nmon,
                     void* stream);

/* The share of cells whose residuals (u, v, w at both of the cell's faces)
 * are all at most threshold[n] scale[n], and the largest height above the
 * ground (node-mean centre minus node-mean ground of the column) of the
 * other cells, 0 if none. z_nd needs the cells' nodes. Synchronises.
 * Replaces: bool SimplecStage::confidence(const vector<Real>& threshold,
 *   Real& share, Real& height) -- Source/Simplec.cpp:411-445, with
 *   height_above_ground -- Source/Closure.cpp:261-293. */
int gust_confidence(const gust_geometry* geom, const gust_residuals* R, const gust_array4* z_nd,
                    const gust_jvp_scales* s, double height, const double threshold[6], double* share, double* top,
                    void* stream);

/* ---- The residual and its physics (periodic sides) ---------------------- */

/* Face types, Source/BoundaryConditions.H:76 (FaceType); indexed like
 * amrex::Orientation: xlo, ylo, zlo, xhi, yhi, zhi. */
enum { GUST_FACE_PERIODIC = 0, GUST_FACE_INFLOW, GUST_FACE_OUTFLOW, GUST_FACE_SLIP_WALL, GUST_FACE_NO_SLIP_WALL,
       GUST_FACE_SURFACE_LAYER, GUST_FACE_GEOSTROPHIC };

/* The k-L closure's options, Source/Closure.H:78-105 (ClosureOptions:
 * Cmu0 0.5562, Cb 0.35, Rt_crit -1, Rt_min -3, max_geom_lscale 30 m,
 * tke_floor 0 = machine epsilon, sigma_k 1, nu_molecular 0, alpha_molecular
 * 0, kappa 0.41). nut_relax 1, no stratification gate, no realizability cap. */
typedef struct gust_closure_options {
    double Cmu0, Cb, Rt_crit, Rt_min, max_geom_lscale, tke_floor, sigma_k, nu_molecular, alpha_molecular, kappa;
} gust_closure_options;

/* The wall model, Source/SurfaceLayer.H:100-146 (SurfaceLayerOptions):
 * neutral stability, the reference point at the first cell centre.
 * wall_k: 0 dirichlet, 1 source; denominator_floor: 0 log_law, 1 one. */
typedef struct gust_surface_options {
    double z0, Cmu0, kappa;
    int32_t wall_k, denominator_floor;
} gust_surface_options;

/* Source/Forcing.H:40-96 (BuoyancyOptions, CoriolisOptions, the kinematic
 * pressure-gradient force). */
typedef struct gust_forcing {
    double gravity, theta0;
    int32_t coriolis, coriolis_3d, geostrophic;
    double latitude, rotation_period, ug, vg;
    double pressure_gradient[3];
} gust_forcing;

enum { GUST_INFLOW_LOG_LAW = 0, GUST_INFLOW_POWER_LAW = 1, GUST_INFLOW_TABLE = 2 };

/* Source/InflowProfile.H:42-66 (InflowProfileOptions; defaults from
 * gust_inflow_default_profile) with the file kind's table given in memory
 * (the caller reads "z u v theta [k]"; or a converged column's table,
 * InflowProfile.H:81-89): n rows, z strictly increasing from 0, theta > 0,
 * k >= 0 or k = NULL (k from the table's shear); host arrays, copied by each
 * call. unbounded: bc.inflow_tke_length = unbounded (else the closure's cap
 * l_max = turb.max_geom_lscale). The sounding theta_ref(z) (bc.theta_profile_z,
 * bc.theta_profile): 1 to 64 points, heights increasing. */
typedef struct gust_inflow_profile {
    int32_t kind;
    double ustar, z0;
    double speed_ref, z_ref, alpha;
    double direction;     /* degrees from the x axis */
    double Cmu0, kappa;
    int32_t unbounded;
    double l_max;
    int32_t n;
    const double* z;
    const double* u;
    const double* v;
    const double* theta;
    const double* k;
    int32_t n_theta_ref;
    double theta_ref_z[64];
    double theta_ref[64];
} gust_inflow_profile;

/* The Rayleigh and sponge layers, Source/BoundaryConditions.H:53-57, 107-121
 * (bc.rayleigh_*, bc.sponge_*; both off by default): above rayleigh_zdamp (> 0
 * enables) R += c (q - q_ref), c = rayleigh_dampcoef sin^2(pi/2 (z - zdamp) /
 * (ztop - zdamp)), for the fields rayleigh_fields (u, v, w, theta, k) with
 * z absolute or above the ground (rayleigh_above_ground); within
 * sponge_width (> 0 enables) of an inflow or outflow face c = sponge_strength
 * (1 - s/width)^2 for every field; q_ref the inflow profile (w: 0); c is
 * added to the implicit diagonal. ztop: the mesh's top level; prob_lo,
 * prob_hi: the domain's x and y extent (m). */
typedef struct gust_damping {
    double rayleigh_zdamp, rayleigh_dampcoef;
    int32_t rayleigh_fields[5];
    int32_t rayleigh_above_ground;
    double sponge_width, sponge_strength;
    double ztop;
    double prob_lo[2], prob_hi[2];
} gust_damping;

/* The residual's inputs, Source/Residual.H:90-128 (ResidualInputs) on this
 * scope: the closure (use_closure) or given cell diffusivities with one
 * ghost (nu, gamma_theta, gamma_tke); the wall model on a surface_layer
 * ground; the forcing; the boundary faces (use_bc; GUST_FACE_* in
 * orientation order, periodic exactly where the geometry is); theta_ref(z)
 * as up to 64 (height, theta) points; the convection limiters
 * (GUST_LIMITER_*) and van Albada eps of momentum and theta; the
 * diffusivities' face mean (0 arithmetic, 1 logarithmic); the inflow profile
 * an inflow face takes (Source/BoundaryConditions.H:82-90, bc.inflow_*). */
typedef struct gust_physics {
    int32_t use_closure;
    gust_closure_options closure;
    const gust_array4* nu;
    const gust_array4* gamma_theta;
    const gust_array4* gamma_tke;
    gust_surface_options surface;
    gust_forcing forcing;
    int32_t use_bc;
    int32_t face[6];
    double zhi_theta_grad;
    double wall_velocity[6][3];
    int32_t n_theta_ref;
    double theta_ref_z[64];
    double theta_ref[64];
    int32_t momentum_limiter, theta_limiter;
    double momentum_eps, theta_eps;
    int32_t face_mean;
    gust_inflow_profile inflow;
    gust_damping damping;
} gust_physics;

/* The six state fields u, v, w (faces), p, theta, k (cells), each with two
 * ghost layers (Source/State.H:13). */
typedef struct gust_state {
    gust_array4* f[6];
} gust_state;

/* The mesh's nodal heights and metrics, each with one ghost layer. */
typedef struct gust_mesh {
    const gust_array4* z_nd;
    const gust_array4* detJ;
    const gust_array4* ax;
    const gust_array4* ay;
    const gust_array4* az;
} gust_mesh;

/* What the closure gives: nu, gamma_theta, gamma_tke with one ghost cell;
 * Ps, Pb, eps, dSdk on the valid cells (Source/Closure.H:112-127). */
typedef struct gust_closure_fields {
    gust_array4* nu;
    gust_array4* gamma_theta;
    gust_array4* gamma_tke;
    gust_array4* Ps;
    gust_array4* Pb;
    gust_array4* eps;
    gust_array4* dSdk;
} gust_closure_fields;

/* Components of the surface-state slab (Source/SurfaceLayer.H:163-174):
 * one cell thick at the ground, one lateral ghost column. */
enum { GUST_SS_USTAR = 0, GUST_SS_SPEED, GUST_SS_D0, GUST_SS_TAU0, GUST_SS_TAU1, GUST_SS_TAU2, GUST_SS_DRAG,
       GUST_SS_KWALL, GUST_SS_AREA, GUST_SS_INVL, GUST_SS_N };

/* Cells and one ghost cell: the node-mean centre height minus the
 * node-mean ground of the column; the vertical ghosts copy the first and
 * last cells.
 * Replaces: void height_above_ground(const MultiFab& z_nd, MultiFab& dist)
 *   -- Source/Closure.H:141, Source/Closure.cpp:261-293. */
int gust_height_above_ground(const gust_geometry* geom, const gust_array4* z_nd, gust_array4* dist, void* stream);

/* The closure's fields at the state (ghosts filled) with the wall distance
 * dist (one ghost) and the ground roughness z0.
 * Replaces: void Closure::evaluate(const State&, const MultiFab& dist,
 *   const Roughness&, const BuoyancyOptions&, const Mesh&, const Geometry&,
 *   ClosureFields&) const -- Source/Closure.H:148-149,
 *   Source/Closure.cpp:366-466. */
int gust_closure_evaluate(const gust_geometry* geom, const gust_state* q, const gust_array4* z_nd,
                          const gust_array4* dist, double z0, const gust_closure_options* opt, double gravity,
                          double theta0, gust_closure_fields* out, void* stream);

/* The surface state of every ground column and one lateral ghost column
 * into ss (GUST_SS_N components); fails if a u* is not positive and finite
 * (synchronises).
 * Replaces: void compute_surface_state(const State&, const
 *   SurfaceLayerOptions&, const BuoyancyOptions&, const Mesh&, const
 *   Geometry&, SurfaceState&) -- Source/SurfaceLayer.H:178-179,
 *   Source/SurfaceLayer.cpp:355-428. */
int gust_surface_state(const gust_geometry* geom, const gust_array4* u, const gust_array4* v,
                       const gust_array4* z_nd, const gust_surface_options* opt, gust_array4* ss, void* stream);

/* Every ghost of the state: the periodic images along x and y, the z faces'
 * rules (the wall law's gradient on a surface_layer ground, w = Omega-free
 * on the ground and the top). ss (may be null) receives the surface state.
 * Replaces: void BoundaryConditions::fill_ghosts(State&, const
 *   SurfaceLayerOptions*, const BuoyancyOptions&, bool homogeneous,
 *   SurfaceState*) const -- Source/BoundaryConditions.H:163-164,
 *   Source/BoundaryConditions.cpp:597-666 (not homogeneous). */
int gust_fill_ghosts(const gust_geometry* geom, gust_state* q, const gust_mesh* mesh, const gust_physics* phys,
                     gust_array4* ss, void* stream);

/* R(q) with the state's ghosts filled: R_c = C - D + G p - S (momentum),
 * R_p = D(U), R_theta = C - D, R_k = bounded upwind C - D - S_k, the wall
 * model's ground fluxes and wall-k rows, the Dirichlet rows of w, and the
 * periodic seam sync of R_u, R_v. cf and ss (may be null) receive the
 * closure's fields and the surface state.
 * Replaces: void residual(const State&, const ResidualInputs&, const Mesh&,
 *   const Geometry&, ResidualFields&, ClosureFields*) --
 *   Source/Residual.H:133-134, Source/Residual.cpp:159-256. */
int gust_residual(const gust_geometry* geom, const gust_state* q, const gust_mesh* mesh, const gust_physics* phys,
                  gust_residuals* R, gust_closure_fields* cf, gust_array4* ss, void* stream);

/* The implicit rows of equation eq (GUST_EQ_U..GUST_EQ_TKE) at its
 * location: A1 from the transport fluxes F (one ghost) and the equation's
 * diffusivity, the Patankar diagonal dSdk for k, the wall model's rows and
 * the z faces' folds (7 components, valid points). cf: the closure's
 * fields (null: the given diffusivities); ss: the surface state (needed
 * with a surface_layer ground).
 * Replaces: void implicit_rows(const Array<const MultiFab*,3>& F, const
 *   MultiFab& z_nd, const ResidualInputs&, Equation, const Mesh&, const
 *   Geometry&, MultiFab& A, const SurfaceState*, const ClosureFields*) --
 *   Source/Residual.H:157-159, Source/Residual.cpp:272-304. */
enum { GUST_EQ_U = 10, GUST_EQ_V = 11, GUST_EQ_W = 12 };
int gust_implicit_rows(int32_t eq, const gust_geometry* geom, const gust_faces* F, const gust_mesh* mesh,
                       const gust_physics* phys, const gust_closure_fields* cf, const gust_array4* ss,
                       gust_array4* A, void* stream);

/* ---- The SIMPLEC stage and the steady driver ---------------------------- */

/* Source/Simplec.H:71-83 (SimplecOptions: pressure_reltol 1e-3, alpha_p 1,
 * rotation_relax 0.3, w_vnn_cap 0, w_vnn_cap_slope_margin 4) and the stage's
 * other settings: steady.zline_sweeps and the sweep options, turb.tke_floor
 * (the clip of k), the residual scales (JvpScales), the residual scale
 * height (residual_scale_height), the Poisson controls (their reltol is
 * replaced by pressure_reltol; precision 1 runs the V-cycles in FP32). */
typedef struct gust_simplec_options {
    double pressure_reltol, alpha_p, rotation_relax, w_vnn_cap, w_vnn_cap_slope_margin;
    int32_t sweeps;
    gust_line_options sweep;
    double tke_floor;
    gust_jvp_scales scales;
    double height;
    gust_poisson_options poisson;
    /* gust_predictor_options.sweep_precision for every sweep of the stage
     * (momentum and scalars); with poisson.precision = 1 the stage's
     * mixed-precision form. */
    int32_t sweep_precision;
} gust_simplec_options;

/* Source/Simplec.H:178-184 (SimplecIterationInfo). */
typedef struct gust_simplec_info {
    gust_momentum_predictor_result predictor;
    gust_pressure_correction_result pressure;
    gust_scalar_sweep_result theta, tke;
} gust_simplec_info;

typedef struct gust_simplec gust_simplec;

/* Replaces: SimplecOptions{} -- Source/Simplec.H:71-83. */
int gust_simplec_default_options(gust_simplec_options* opt);

/* The stage on a tile: its fields (R, the closure's fields, the surface
 * state, the transport fluxes, the implicit rows, ap and d, the Dirichlet
 * masks) and its Poisson solver, allocated here and kept. mesh's views and
 * phys's pointees are referenced and must outlive the stage.
 * Replaces: SimplecStage(const SimplecOptions&, int sweeps, const
 *   ResidualInputs&, Closure*, const PoissonOptions&, const Mesh&, const
 *   Geometry&, const MultiFab& z_nd, Real tke_floor, const BoxArray&, const
 *   DistributionMapping&, const LineSweepOptions&) -- Source/Simplec.H:197-200,
 *   Source/Simplec.cpp:293-336. */
int gust_simplec_create(const gust_geometry* geom, const gust_mesh* mesh, const gust_physics* phys,
                        const gust_simplec_options* opt, gust_simplec** out, void* stream);
int gust_simplec_destroy(gust_simplec* s);

/* The boundary conditions' ghosts at q.
 * Replaces: void SimplecStage::init(State&) -- Source/Simplec.cpp:338-359. */
int gust_simplec_init(gust_simplec* s, gust_state* q, void* stream);

/* The closure's history forgotten (its fields zero, as on a fresh stage), so
 * that a new run on this stage starts as a fresh one: the driver's first step
 * limit reads the diffusivity of the last closure fields. Synchronises.
 * Replaces: constructing a new SimplecStage for the next run (the reference's
 *   application builds one per run, Source/Simplec.cpp:293-336). */
int gust_simplec_reset(gust_simplec* s, void* stream);

/* One SIMPLEC iteration with the cell pseudo-time step dtau: ghosts and
 * surface state, R(q) and the norms, the momentum predictor, the ghosts,
 * the pressure correction, theta's convection at the corrected transport,
 * the theta and k sweeps. info may be null. Synchronises.
 * Replaces: void SimplecStage::iterate(State&, const MultiFab& dtau) --
 *   Source/Simplec.H:204, Source/Simplec.cpp:452-506. */
int gust_simplec_iterate(gust_simplec* s, gust_state* q, const gust_array4* dtau, gust_simplec_info* info,
                         void* stream);

/* The scaled residual norms and monitors of the last iteration's R(q).
 * Replaces: residual_norms() and monitors() -- Source/Simplec.H:205-206. */
int gust_simplec_norms(const gust_simplec* s, double res[6], double mon[3], int32_t* nmon);

/* A view of the momentum diffusivity the driver's dtau reads (the closure's
 * nu at the last residual, zero before the first; or the given nu).
 * Replaces: const MultiFab& SimplecStage::nu() const -- Source/Simplec.H:214. */
int gust_simplec_nu(const gust_simplec* s, gust_array4* nu);

/* ---- Steady criterion (host only) --------------------------------------- */

/* Source/SteadyCriterion.H:13-33 (SteadyParams). */
typedef struct gust_steady_params {
    double residual_tol;
    double residual_drop_orders;
    int32_t residual_ref_iterations;
    double monitor_rel_tol;
    int32_t monitor_window;
} gust_steady_params;

/* Source/SteadyCriterion.H:50-55 (SteadyStatus). */
enum { GUST_STEADY_NOT_STEADY = 0, GUST_STEADY_STEADY = 1, GUST_STEADY_DIVERGED = 2 };

/* Source/SteadyCriterion.H:57-64 (SteadyVerdict); the reasons are returned
 * as text, one line each. */
typedef struct gust_steady_result {
    int32_t status;
    int32_t iteration;
    int32_t residuals_met;
    int32_t monitors_met;
    int32_t nreasons;
    int64_t reasons_len;   /* bytes of the full text, without the NUL */
} gust_steady_result;

typedef struct gust_steady gust_steady;

/* Replaces: SteadyParams{} -- Source/SteadyCriterion.H:16-32. */
int gust_steady_default_params(gust_steady_params* p);

/* One message per violated range, one line each, into buf (NUL-terminated,
 * truncated to cap); *count = the number of violations.
 * Replaces: vector<string> validate(const SteadyParams&) --
 *   Source/SteadyCriterion.cpp:23-48. */
int gust_steady_validate(const gust_steady_params* p, int32_t* count, char* buf, int64_t cap);

/* Fails (the reference aborts) if p fails validation, a monitor scale is not
 * finite and > 0, or there is no residual.
 * Replaces: SteadyCriterion(const SteadyParams&, vector<string>
 *   residual_names, vector<MonitorSpec>) -- Source/SteadyCriterion.cpp:50-68. */
int gust_steady_create(const gust_steady_params* p, int32_t nres, const char* const* residual_names, int32_t nmon,
                       const char* const* monitor_names, const double* monitor_scales, gust_steady** out);
int gust_steady_destroy(gust_steady* c);

/* Replaces: void SteadyCriterion::add(const vector<Real>& residuals, const
 *   vector<Real>& monitors) -- Source/SteadyCriterion.cpp:96-131. */
int gust_steady_add(gust_steady* c, const double* residuals, int32_t nres, const double* monitors, int32_t nmon);

/* reasons: the reasons' text, one per line (may be null with cap 0).
 * Replaces: SteadyVerdict SteadyCriterion::verdict() const --
 *   Source/SteadyCriterion.cpp:133-227. */
int gust_steady_verdict(const gust_steady* c, gust_steady_result* v, char* reasons, int64_t cap);

/* Replaces: int SteadyCriterion::iterations() const -- Source/SteadyCriterion.H:105. */
int gust_steady_iterations(const gust_steady* c, int32_t* n);

/* out[q] = min(residual_tol, 10^-drop max(reference_q, residual_tol)).
 * Replaces: vector<Real> SteadyCriterion::residual_thresholds() const --
 *   Source/SteadyCriterion.cpp:229-238. */
int gust_steady_residual_thresholds(const gust_steady* c, double* out, int32_t n);

/* out[q] = the residual references (the largest of the first
 * residual_ref_iterations values, or the seed raised by them).
 * Replaces: SteadyCriterion::history().res_ref -- Source/SteadyCriterion.H:69-78. */
int gust_steady_references(const gust_steady* c, double* out, int32_t n);

/* Replaces: void SteadyCriterion::seed_reference(const vector<Real>&) --
 *   Source/SteadyCriterion.cpp:85-94. */
int gust_steady_seed_reference(gust_steady* c, const double* ref, int32_t n);

/* The recorded history (Source/SteadyCriterion.H:69-78, SteadyHistory) as a
 * byte image: int32 count, nonfinite, len + bytes of the nonfinite text,
 * int32 nres, res_ref[nres], res_last[nres] (double), int32 nmon, then per
 * monitor int32 n and n doubles, oldest first; little-endian. *size = the
 * image's size; nothing is written if cap is smaller.
 * Replaces: const SteadyHistory& SteadyCriterion::history() const and
 *   void restore(const SteadyHistory&) -- Source/SteadyCriterion.cpp:70-83. */
int gust_steady_history_save(const gust_steady* c, void* buf, int64_t cap, int64_t* size);
int gust_steady_history_restore(gust_steady* c, const void* buf, int64_t size);

/* r_n after the last add: the previous value when the history holds a NaN or
 * infinity, 1 during the reference hold (iterations <= hold), else max_q
 * |res_q| / max(reference_q, residual_tol).
 * Replaces: the r_n block of SteadyVerdict Driver::step(State&) --
 *   Source/Driver.cpp:421-437. */
int gust_steady_relative_residual(const gust_steady* c, const double* residuals, int32_t nres, int32_t hold,
                                  double previous, double* r);

/* cfl0 during the hold, else ser_cfl(cfl, r, cfl0, cfl_max, p, growth).
 * Replaces: Real Driver::next_cfl(Real r) const -- Source/Driver.cpp:325-334. */
int gust_steady_next_cfl(const gust_steady* c, int32_t hold, double cfl, double r, double cfl0, double cfl_max,
                         double ser_exponent, double max_growth, double* out);

/* ---- Grid sequencing's stop and checks (host only) ------------------------ */

/* Largest mesh.hierarchy_levels (Source/MeshHierarchy.H:48). */
#define GUST_MAX_HIERARCHY_LEVELS 10

/* Source/Sequencing.H:85-92 (SequencingOptions). given: the keys given, as
 * bits 1 sequencing_drop_orders, 2 sequencing_monitor_tol, 4 sequencing_hold,
 * 8 sequencing_cfl_start. */
typedef struct gust_sequencing_options {
    int32_t enabled;        /* steady.sequencing (0 | 1; 0) */
    double drop_orders;     /* steady.sequencing_drop_orders (2) */
    double monitor_tol;     /* steady.sequencing_monitor_tol (0: off) */
    int32_t hold;           /* steady.sequencing_hold (0 | 1; 0) */
    int32_t cfl_previous;   /* steady.sequencing_cfl_start: 0 cfl0, 1 previous */
    uint32_t given;
} gust_sequencing_options;

/* Replaces: SequencingOptions{} -- Source/Sequencing.H:85-92. */
int gust_sequencing_default_options(gust_sequencing_options* o);

/* One message per problem, each naming its key, one line each into buf
 * (NUL-terminated, truncated to cap); *count = the number of problems.
 * levels = mesh.hierarchy_levels, nx, ny = mesh.n_cell x and y.
 * Replaces: vector<string> check_sequencing_options(const SequencingOptions&,
 *   const SteadyParams&, const MeshInputs&) -- Source/Sequencing.cpp:118-176. */
int gust_check_sequencing_options(const gust_sequencing_options* o, const gust_steady_params* strict,
                                  int32_t levels, int32_t nx, int32_t ny, int32_t* count, char* buf, int64_t cap);

/* The loose stop after the last iteration the criterion recorded: *stop = 1
 * and the branch that held in reason (NUL-terminated, truncated to cap), or
 * *stop = 0. hold_iterations: the level's reference hold (0 on a released
 * seeded level).
 * Replaces: LooseStop loose_stop(const SteadyHistory&, const SteadyParams&,
 *   const vector<MonitorSpec>&, const SequencingOptions&, int) --
 *   Source/Sequencing.cpp:179-234. */
int gust_loose_stop(const gust_steady* c, const gust_sequencing_options* o, int32_t hold_iterations, int32_t* stop,
                    char* reason, int64_t cap);

/* ---- The steady driver ---------------------------------------------------- */

/* Source/Driver.H:58-66 (SteadyOptions: max_iterations 10000, cfl0 1, cfl_max 10,
 * ser_exponent 1, ser_max_growth 2) and the criterion. */
typedef struct gust_driver_options {
    int32_t max_iterations;
    double cfl0, cfl_max, ser_exponent, ser_max_growth;
    gust_steady_params criterion;
} gust_driver_options;

typedef struct gust_driver_result {
    int32_t iterations;
    int32_t status;             /* GUST_STEADY_* */
    double cfl;                 /* the CFL of the next iteration */
    double relative_residual;   /* r_n of the last iteration */
    double res[6];
    double mon[3];
    int32_t nmon;
} gust_driver_result;

/* The steady loop on one stage from q: local dtau at the current CFL, the
 * iteration, the criterion, r_n and the SER CFL, until Steady, Diverged or
 * max_iterations. Synchronises every iteration.
 * Replaces: DriverResult Driver::run(State&) with one SIMPLEC stage --
 *   Source/Driver.cpp:403-492 (step, verdict, run). */
int gust_steady_solve(gust_simplec* s, gust_state* q, const gust_driver_options* opt, gust_driver_result* result,
                      void* stream);

/* ---- One SIMPLEC iteration as a HIP graph ----------------------------------
 * No counterpart in the reference (its iteration is a sequence of launches
 * and host reads): the iteration's launches, its checks and the copy of
 * their values recorded once and replayed, one launch and one wait per
 * iteration. The pressure solve runs a fixed number of GMRES steps on the
 * device (gust_poisson_options.fixed_steps); its tolerance is still checked
 * (an error after the replay when the steps were too few). */

/* The stage's pressure solver runs k fixed device GMRES steps (k >= 1) or,
 * k = 0, GMRES to the tolerance with host reads (the default). Drops any
 * captured iteration. */
int gust_poisson_set_fixed_steps(gust_poisson* p, int32_t steps);
/* soft = 1: a fixed-step solve that misses its tolerance is counted, not an
 * error (a caller that adapts the step count); *misses (may be null) = the
 * count so far. */
int gust_poisson_fixed_steps_misses(gust_poisson* p, int32_t soft, int64_t* misses);
int gust_simplec_set_pressure_steps(gust_simplec* s, int32_t steps);

/* Records one iteration on (q, dtau) into the stage's graph without running
 * it. Before it, run one iteration with gust_simplec_iterate on the same
 * stream, state and dtau and with the pressure steps fixed, so that every
 * buffer and table exists; the graph holds q's and dtau's addresses. */
int gust_simplec_capture(gust_simplec* s, gust_state* q, const gust_array4* dtau, void* stream);

/* Runs the captured iteration: one graph launch, one wait, then the checks
 * (the first failed one is the error, with the reference's message), the
 * norms (gust_simplec_norms) and info (may be null). */
int gust_simplec_replay(gust_simplec* s, gust_simplec_info* info, void* stream);

/* gust_steady_solve with the iteration captured and replayed: the first
 * iteration runs eagerly with GMRES to the tolerance and its GMRES count + 1
 * becomes the fixed pressure steps (pressure_steps > 0 fixes them instead);
 * the next runs eagerly with them and is captured; from then on each
 * iteration is one replay (local dtau at the current CFL is the one launch
 * outside the graph). A replay whose pressure solve misses its tolerance
 * keeps its (less converged) correction, and the steps grow by 2 and the
 * iteration is captured again. */
int gust_steady_solve_graphed(gust_simplec* s, gust_state* q, const gust_driver_options* opt, int32_t pressure_steps,
                              gust_driver_result* result, void* stream);

/* How a level of grid sequencing starts and stops (the driver's seeding,
 * Source/Driver.cpp:275-313, and the level loop of
 * Source/Sequencing.cpp:361-390). loose: the loose stop's options (null: the
 * strict criterion alone, as level 0); seeded: seed_ref holds the coarsest
 * level's residual references (seed_reference); release_hold: no reference
 * hold (needs seeded); start_cfl > 0: the first CFL (no hold; raised to cfl0). */
typedef struct gust_level_control {
    const gust_sequencing_options* loose;
    int32_t seeded;
    double seed_ref[6];
    int32_t release_hold;
    double start_cfl;
} gust_level_control;

/* stop: GUST_LEVEL_* (how the loop ended); res_ref: the criterion's
 * references at the end (the coarsest level's seed the finer ones). */
enum { GUST_LEVEL_MAX_ITERATIONS = 0, GUST_LEVEL_STEADY = 1, GUST_LEVEL_LOOSE = 2, GUST_LEVEL_DIVERGED = 3 };
typedef struct gust_level_result {
    gust_driver_result driver;
    int32_t stop;
    int32_t hold;   /* the level's reference hold (Driver::hold_iterations) */
    double res_ref[6];
    char reason[256];
} gust_level_result;

/* The steady loop of gust_steady_solve on one level of grid sequencing: the
 * seeding and start of lc, then until Steady, Diverged, the loose stop or
 * max_iterations (the CFL of the next iteration is computed before the
 * loose stop is tested, as the reference's step does). A diverged level
 * returns status Diverged (the caller aborts, naming the level).
 * Replaces: Driver::seed_reference / release_reference_hold / start_cfl /
 *   step and the level loop of sequence_coarse_levels --
 *   Source/Driver.cpp:275-313, 403-468, Source/Sequencing.cpp:361-390. */
int gust_steady_solve_level(gust_simplec* s, gust_state* q, const gust_driver_options* opt,
                            const gust_level_control* lc, gust_level_result* result, void* stream);

/* ---- Grid transfers (Source/Transfer.H) -----------------------------------
 * Level lc is coarser than level lf by ratio = 2^(lc - lf) in x and y; z is
 * never coarsened. The geometries are the two levels' (the same lo, hi z;
 * x, y coarsened by ratio). */

/* fine's valid points of u, v, w, p, theta, k from coarse's, bilinear per
 * zeta-level at each field's own staggered point: linear extrapolation in
 * the half cell next to a non-periodic face, the periodic images across a
 * seam, constant along a non-periodic cell-centred direction one coarse cell
 * wide. coarse's ghosts are not read; fine's are not filled.
 * Replaces: void prolong_fields(const MeshHierarchy&, int lc, const State&
 *   coarse, int lf, State& fine) -- Source/Transfer.cpp:324-352. */
int gust_prolong_fields(const gust_geometry* fine_geom, const gust_geometry* coarse_geom, int32_t ratio,
                        const gust_state* coarse, gust_state* fine, void* stream);

/* coarse's valid points from fine: theta, k, p the detJ-weighted means, u, v
 * the face fluxes over the coarse areas, w from the restricted vertical flux
 * Omega (omega, when not null, receives Omega_c on the coarse z-faces); the
 * slope term of w reads u, v with periodic images and zero-gradient rows
 * below the ground and above the top. fine's u and v need one ghost filled.
 * Replaces: void restrict_state(const MeshHierarchy&, int lf, const State&
 *   fine, int lc, State& coarse, MultiFab* omega) -- Source/Transfer.cpp:263-322. */
int gust_restrict_state(const gust_geometry* fine_geom, const gust_mesh* fine_mesh, const gust_state* fine,
                        const gust_geometry* coarse_geom, const gust_mesh* coarse_mesh, int32_t ratio,
                        gust_state* coarse, gust_array4* omega, void* stream);

typedef struct gust_prolong_result {
    gust_pressure_correction_result projection;   /* divergence before and after, the Poisson solve */
    int64_t tke_clipped;                          /* fine cells raised to the k floor */
} gust_prolong_result;

/* The prolongation into the fine level of stage: gust_prolong_fields, k
 * clipped at tke_floor, the stage's boundary conditions (its ghost fill with
 * the wall state), then the projection: the pressure correction with d = 1
 * off the Dirichlet rows (0 on them), solver (a gust_poisson on the fine
 * mesh) and the stage's pressure boundary types, alpha 1; p is restored
 * afterwards.
 * Replaces: ProlongResult prolong_state(const MeshHierarchy&, int lc, const
 *   State& coarse, int lf, State& fine, const BoundaryConditions&, const
 *   SurfaceLayerOptions*, const BuoyancyOptions&, PoissonSolver&, Real
 *   tke_floor) -- Source/Transfer.cpp:354-403. */
int gust_prolong_state(gust_simplec* fine_stage, gust_poisson* solver, const gust_geometry* coarse_geom,
                       int32_t ratio, const gust_state* coarse, gust_state* fine, double tke_floor,
                       gust_prolong_result* result, void* stream);

/* ---- The state's files (Source/StateIO.H) ---------------------------------
 * prefix + the step as five digits names the directory (created). prob_lo,
 * prob_hi: the domain's corners (3 each). max_grid_size > 0 splits the
 * domain in x and y into boxes of at most that many cells (AMReX's maxSize);
 * 0 writes one box. name (may be null) receives the directory's name. */

/* An AMReX plotfile (HyperCLaw-V1.1): the cell variables u, v, w (face
 * means), p, theta, tke, z (the mean of the cell's eight node heights) and,
 * with detJ, detJ; the node heights with their ghost nodes as the VisMF
 * MultiFab Level_0/z_phys_nd.
 * Replaces: string write_plotfile(const string& prefix, const State&, const
 *   Geometry&, const MultiFab* detJ) -- Source/StateIO.cpp:41-85. */
int gust_write_plotfile(const char* prefix, int32_t step, const gust_geometry* geom, const double* prob_lo,
                        const double* prob_hi, const gust_state* q, const gust_array4* z_nd, const gust_array4* detJ,
                        int32_t max_grid_size, char* name, int64_t cap, void* stream);

/* A checkpoint: the text Header (magic and version, real bytes, step, mesh
 * in hex floats, periodicity, BoxArray) and every field as a VisMF MultiFab
 * with its ghosts (u, v, w, p, theta, tke: two; z_phys_nd: one).
 * Replaces: string write_checkpoint(const string& prefix, State&, const
 *   MeshInputs&) -- Source/StateIO.cpp:87-113. */
int gust_write_checkpoint(const char* prefix, int32_t step, const gust_geometry* geom, const double* prob_lo,
                          const double* prob_hi, const gust_state* q, const gust_array4* z_nd, int32_t max_grid_size,
                          char* name, int64_t cap, void* stream);

/* Reads a checkpoint into q (valid values exactly, from any box layout) and,
 * when z_nd is not null, the node heights with their ghost nodes; *step =
 * its step. Fails, naming io.restart, on a corrupt Header, another
 * precision, or a mesh (n_cell, prob_lo, prob_hi, periodicity) other than
 * geom's and prob_lo/prob_hi.
 * Replaces: void read_checkpoint(const string& dir, State&, const MeshInputs&)
 *   -- Source/StateIO.cpp:115-211. */
int gust_read_checkpoint(const char* dir, const gust_geometry* geom, const double* prob_lo, const double* prob_hi,
                         gust_state* q, gust_array4* z_nd, int32_t* step, void* stream);

/* ---- The mass-consistent initializer --------------------------------------
 * Source/Initializer.H, Source/InflowProfile.H: the first guess from an
 * inflow profile by height above the ground, then one anisotropic
 * projection by the terrain Poisson solver. */

/* The reference's defaults: log law, u* 0.4, z0 0.1, speed 10 at 10 m with
 * alpha 0.14, direction 0, Cmu0 0.5562, kappa 0.41, the closure's cap at 30 m,
 * theta_ref 300 K. */
int gust_inflow_default_profile(gust_inflow_profile* p);

/* 0 if valid, else nonzero with the first problem, naming its bc.* key.
 * Replaces: check_inflow_profile_options -- Source/InflowProfile.cpp:14-47,
 *   and the table's and the sounding's checks of the InflowProfile
 *   constructors (InflowProfile.cpp:49-145, Source/Forcing.cpp:30-49). */
int gust_check_inflow_profile(const gust_inflow_profile* p);

/* The profile at n heights above the ground (device arrays; any output may
 * be NULL): the wind, theta, the equilibrium k and nu_t.
 * Replaces: InflowProfile::View::eval -- Source/InflowProfile.H:149-200,
 *   with the view's constants of Source/InflowProfile.cpp:178-205. */
int gust_inflow_eval(const gust_inflow_profile* p, int32_t n, const double* z, double* u, double* v, double* theta,
                     double* k, double* nut, void* stream);

/* Source/Initializer.H:63-98 (InitOptions), the profile first guess:
 * first_guess 0 none | 1 profile (init.first_guess); projection (1);
 * z_agl_min (0), alpha_h, alpha_v (1), k_decay_height (0), obrien (0);
 * column_theta: theta from the table by height above the ground instead of
 * the sounding by absolute height (a stratified column's table). */
typedef struct gust_init_options {
    int32_t first_guess;
    int32_t projection;
    double z_agl_min;
    double alpha_h, alpha_v;
    double k_decay_height;
    int32_t obrien;
    int32_t column_theta;
} gust_init_options;

int gust_init_default_options(gust_init_options* o);

/* Replaces: check_init_options (the profile's keys) -- Source/Initializer.cpp:109-179. */
int gust_check_init_options(const gust_init_options* o);

/* Source/Initializer.H:110-120 (InitResult). */
typedef struct gust_init_result {
    gust_poisson_result poisson;
    double divergence_before, divergence_after;   /* 2-norms of velocity_divergence (1/s) */
    double obrien_top_residual;                   /* max |Omega_top| removed (m/s) */
} gust_init_result;

/* The correction's boundary types for the face types faces[6] (GUST_FACE_*,
 * orientation order): Dirichlet on inflow and outflow faces, Neumann
 * elsewhere, periodic where the geometry is.
 * Replaces: PoissonBC initializer_poisson_bc(const BoundaryInputs&, const
 *   Geometry&) -- Source/Initializer.H:108, Source/Initializer.cpp:181-195. */
int gust_initializer_poisson_bc(const gust_geometry* geom, const int32_t* faces, int32_t* pbc);

/* Omega on the valid z-faces from the column continuity of (u, v) from the
 * ground up, the top residual E removed with (h / depth)^2; *top_residual =
 * max |E| before the removal. Synchronises.
 * Replaces: Real obrien_omega(const MultiFab& u, v, z_nd, const Mesh&,
 *   const Geometry&, MultiFab& omega) -- Source/Initializer.H:156-157,
 *   Source/Initializer.cpp:197-267. */
int gust_obrien_omega(const gust_geometry* geom, const gust_mesh* mesh, const gust_array4* u, const gust_array4* v,
                      gust_array4* omega, double* top_residual, void* stream);

/* Steps 1-3 of the initializer on q (two ghosts): u, v from the profile at
 * every point, w = 0 but for the ground faces' kinematic value, k the
 * profile's with the decay above k_decay_height, theta the sounding (one
 * ghost cell) or the table's, p' = 0, then O'Brien when asked;
 * *obrien_residual its top residual. Synchronises.
 * Replaces: Real first_guess_state(const InitOptions&, const BoundaryInputs&,
 *   const Mesh&, const Geometry&, State&, ...) -- Source/Initializer.H:133-135,
 *   Source/Initializer.cpp:269-335. */
int gust_first_guess_state(const gust_geometry* geom, const gust_mesh* mesh, const gust_inflow_profile* prof,
                           const gust_init_options* opt, gust_state* q, double* obrien_residual, void* stream);

/* The first guess, then (projection = 1) the correction div(Sigma grad
 * lambda) = div(u0) by solver (created on this mesh by the caller, its
 * options the solver.poisson_* controls), (u, v, Omega) += F, the periodic
 * images of u and v, w from Omega. lambda: cells with one ghost cell, or
 * NULL. The state's other ghost cells are not filled. Synchronises.
 * Replaces: InitResult initialize_state(const InitOptions&, const
 *   BoundaryInputs&, const PoissonOptions&, const Mesh&, const Geometry&,
 *   State&, MultiFab* lambda, ...) -- Source/Initializer.H:146-149,
 *   Source/Initializer.cpp:344-418. */
int gust_initialize_state(const gust_geometry* geom, const gust_mesh* mesh, const gust_inflow_profile* prof,
                          const gust_init_options* opt, const int32_t* faces, gust_poisson* solver, gust_state* q,
                          gust_array4* lambda, gust_init_result* result, void* stream);

/* ---- The terrain-following mesh of a tile (Source/Mesh.H) --------------- */

/* Build a tile's BTF mesh on the device from its ground node heights.
 * Replaces: Mesh::build_from_ground with terrain.smoothing = 0
 *   (Source/Mesh.cpp:107-140, the ghost columns 37-98) and the metrics of
 *   extern_erf/ERF_TerrainMetrics.cpp:330-356 (BTF) and 679-807 (make_J,
 *   make_areas, make_zcc).
 * ground: host, (nx + 3) * (ny + 3) node heights, node (i, j) at
 *   [(j + 1) * (nx + 3) + (i + 1)] for i in -1 .. nx + 1 (what the terrain
 *   gives at each node, ghost nodes included; the periodic images and the
 *   clamped edge columns are taken here). zlevels: host, nz + 1 increasing.
 * Outputs (device, one ghost layer): z_nd on the nodes, detJ and zcc (may be
 *   null) on the cells, ax, ay, az on the faces, bitwise the host builder's.
 * Refuses terrain at or above the top level and inverted cells with the
 *   reference's messages. */
int gust_build_mesh(const gust_geometry* geom, const double* ground, const double* zlevels, gust_array4* z_nd,
                    gust_array4* detJ, gust_array4* ax, gust_array4* ay, gust_array4* az, gust_array4* zcc,
                    void* stream);

/* ---- The Newton-Krylov finish (Source/Newton.H, Source/Continuation.H,
 * Source/NewtonPreconditioner.H, Source/NewtonStage.H) ------------------- */

/* The fill of a perturbation: every ghost of the state as gust_fill_ghosts,
 * with each prescribed value, inflow profile and z gradient (the wall law's
 * too) replaced by 0; no surface state.
 * Replaces: BoundaryConditions::fill_ghosts(q, surface, buoyancy,
 *   homogeneous = true) -- Source/BoundaryConditions.H:163-164,
 *   Source/BoundaryConditions.cpp:418-540, 597-666. */
int gust_fill_ghosts_homogeneous(const gust_geometry* geom, gust_state* q, const gust_mesh* mesh,
                                 const gust_physics* phys, void* stream);

/* The Newton stage's inputs. Replaces: NewtonOptions -- Source/Newton.H:78-99;
 * ContinuationOptions -- Source/Continuation.H:63-79; the driver's
 * steady.newton_cfl_max, steady.newton_preconditioner, steady.cfl_max (the
 * preconditioner's CFL without a step, the continuation's first CFL) --
 * Source/Driver.H:78-97. */
typedef struct gust_newton_options {
    int32_t gmres_restart;     /* steady.newton_gmres_restart, 30 (5..200) */
    int32_t gmres_max_iters;   /* steady.newton_gmres_max_iters, 2000 (>= 1) */
    double fd_eps_scale;       /* steady.newton_fd_eps_scale, sqrt(machine epsilon) */
    double krylov_rtol;        /* steady.newton_krylov_rtol, 1e-2 (0 < rtol < 1) */
    int32_t preconditioner;    /* steady.newton_preconditioner: 0 identity, 1 simplec (default) */
    double cfl_max;            /* steady.newton_cfl_max, 128: the continuation's cap and SER cfl0 */
    double start_cfl;          /* steady.cfl_max: the continuation's first CFL and the preconditioner's
                                  CFL when a step has none */
    double ser_exponent, ser_max_growth;   /* steady.ser_exponent (1), steady.ser_max_growth (2) */
    double armijo_c;           /* steady.newton_armijo_c, 1e-4 */
    double backtrack;          /* steady.newton_backtrack, 0.5 */
    double lambda_min;         /* steady.newton_lambda_min, 1e-2 */
    double cfl_cut;            /* steady.newton_cfl_cut, 0.25 */
    int32_t max_retries;       /* steady.newton_max_retries, 4 */
    int32_t dtau_inf_at_cap;   /* steady.newton_dtau_inf_at_cap, 1 */
    double ptc_growth;         /* steady.newton_ptc_growth, 1 (off) */
    int32_t pc_denominator;    /* the SIMPLEC preconditioner's pressure denominators
                                  (PressureDenominator, Source/NewtonPreconditioner.H:116-118):
                                  1 diagonal, d = a_P (default); 0 stage, the SIMPLEC
                                  d = a_P - sum a_nb, with which apply(-R(q)) is one
                                  iteration of the stage */
} gust_newton_options;

/* Source/Newton.H:236-252 (NewtonStepInfo). */
typedef struct gust_newton_step_info {
    int32_t gmres_iterations;
    int32_t gmres_status;      /* 0 converged, 1 the iteration cap */
    double residual_norm;      /* |T^-1 R(q)| before the step */
    double krylov_residual, krylov_relative;
    double eps_min, eps_max;
    int32_t jvp_calls;
    int32_t applied;
} gust_newton_step_info;

/* Source/Continuation.H:118-126 (ContinuationState). */
typedef struct gust_continuation_state {
    int32_t steps;
    double cfl;
    int32_t infinite;
    double r0, r_last;
} gust_continuation_state;

/* Source/Continuation.H:133-151 (ContinuationRecord); outcome 0 accepted,
 * 1 skipped, 2 rejected. */
typedef struct gust_continuation_record {
    int32_t step, attempt;
    double residual, residual_new, cfl;
    int32_t infinite;
    double lambda;
    int32_t backtracks, gmres_iterations, gmres_status;
    double krylov_relative;
    int32_t outcome, by_growth;
} gust_continuation_record;

typedef struct gust_newton gust_newton;

/* The documented defaults. */
int gust_newton_default_options(gust_newton_options* opt);
/* One message per violated range (naming steady.<key>), one line each, into
 * buf (NUL-terminated, truncated to cap); *count = their number.
 * Replaces: check_newton_options -- Source/Newton.cpp:161-193;
 *   check_continuation_options -- Source/Continuation.cpp. */
int gust_check_newton_options(const gust_newton_options* opt, char* buf, int64_t cap, int32_t* count);

/* J v by the forward difference of the full residual (Source/Residual.cpp:
 * 324-396): eps = step (1 + mean|q/s|) / rms(v/s) over the six fields
 * (theta about the stage's theta_ref), q + eps v on every point, its ghost
 * cells refilled, Jv = (R(q + eps v) - R0) / eps on the valid points. q:
 * ghosts filled, R0 = R(q); v: ghosts filled (homogeneously). *eps = 0 and
 * Jv = 0 when v = 0. The stage gives the mesh, the residual inputs and the
 * scales (step: scales.step; a Newton stage uses its fd_eps_scale).
 * Replaces: Real jvp(const State& q, const State& v, const ResidualInputs&,
 *   const Mesh&, const Geometry&, ResidualFields& Jv, const ResidualFields*
 *   R0, bool refill = true) -- Source/Residual.H (jvp), Residual.cpp:324-396. */
int gust_jvp(gust_simplec* stage, const gust_state* q, const gust_state* v, const gust_residuals* R0,
             gust_residuals* Jv, double* eps, void* stream);

/* The Newton-Krylov core, the continuation and the stage on a SIMPLEC
 * stage's tile: its mesh, residual inputs, scales, sweeps, pressure solver
 * and Dirichlet rows (the stage must outlive it). Allocates FGMRES(restart)'s
 * 2 restart + 1 vectors of the six fields and the work states.
 * Replaces: NewtonKrylov, Continuation, SimplecPreconditioner and NewtonStage's
 *   constructors -- Source/Newton.H:277-285, Continuation.H:163-168,
 *   NewtonPreconditioner.H:151-155, NewtonStage.H:99-102. */
int gust_newton_create(gust_simplec* stage, const gust_newton_options* opt, gust_newton** out, void* stream);
void gust_newton_destroy(gust_newton* n);

/* q += dq from (D + J) dq = -R(q) (D = 1/dtau on the rows with a pseudo-time
 * term; dtau null: plain Newton), FGMRES in the scaled variables with the
 * options' preconditioner; q's ghosts filled before and after; the update
 * skipped when |R| or the Krylov residual is not finite or the Krylov
 * residual exceeds |R|. Synchronises.
 * Replaces: NewtonStepInfo NewtonKrylov::step(State&, const MultiFab* dtau,
 *   NewtonPreconditioner*) -- Source/Newton.H:287-292, Newton.cpp:531-539. */
int gust_newton_step(gust_newton* n, gust_state* q, const gust_array4* dtau, gust_newton_step_info* info,
                     void* stream);
/* |T^-1 R(fill(q))|_2 (q's ghosts filled).
 * Replaces: NewtonKrylov::residual_norm -- Source/Newton.cpp:553-565. */
int gust_newton_residual_norm(gust_newton* n, gust_state* q, double* r, void* stream);

/* The SIMPLEC preconditioner of the stage (created with preconditioner 1)
 * on its own: setup at q (ghosts filled) with the cell step dtau (null: the
 * local step at start_cfl), then dq = apply(r) on the six fields' valid
 * points (r's periodic seam copies equal; dq's take the owner's, the low
 * face). The Newton step calls both itself; these are for callers that
 * precondition their own Krylov solves.
 * Replaces: SimplecPreconditioner::setup(const State&, const MultiFab* dtau),
 *   SimplecPreconditioner::apply(NewtonVector& dq, const NewtonVector& r) --
 *   Source/NewtonPreconditioner.H:141-144, NewtonPreconditioner.cpp:114-224. */
int gust_newton_pc_setup(gust_newton* n, const gust_state* q, const gust_array4* dtau, void* stream);
int gust_newton_pc_apply(gust_newton* n, const gust_residuals* r, gust_residuals* dq, void* stream);

/* One continuation step with its retries: local dtau at the state's CFL (or
 * none at infinite dtau), the Newton direction, the Armijo line search (k
 * clipped at the stage's floor), the SER update; *status 0 accepted, 1
 * failed (q unchanged). *last: the last attempt's record (may be null).
 * Replaces: ContinuationStatus Continuation::step(State&, NewtonPreconditioner*)
 *   -- Source/Continuation.H:171-174, Continuation.cpp:339-427. */
int gust_continuation_step(gust_newton* n, gust_state* q, int32_t* status, gust_continuation_record* last,
                           void* stream);
/* The continuation state: read, or restore (the CFL brought to at most the
 * cap). Replaces: Continuation::state / restore -- Continuation.H:180-188. */
int gust_continuation_get(const gust_newton* n, gust_continuation_state* st);
int gust_continuation_restore(gust_newton* n, const gust_continuation_state* st);

/* The driver's switch rules between the SIMPLEC and Newton stages
 * (Source/Driver.H:67-97). */
typedef struct gust_switch_options {
    int32_t switch_window;        /* steady.switch_window, 20 */
    double switch_residual;       /* steady.switch_residual, 1e-4 */
    double switch_back_growth;    /* steady.switch_back_growth, 10 */
    double switch_monitor_tol;    /* steady.switch_monitor_tol, 1e-2 (0: off) */
    int32_t newton_fallback_failures;   /* steady.newton_fallback_failures, 2 */
} gust_switch_options;

/* What a run with the Newton stage did. */
typedef struct gust_newton_run_result {
    gust_driver_result driver;
    int32_t simplec_iterations, newton_iterations;
    int32_t switches, first_switch;    /* the iteration of the first switch to Newton, -1 if none */
    int32_t newton_steps;              /* accepted continuation steps */
    int32_t stage;                     /* the stage at the end: 0 SIMPLEC, 1 Newton */
    int32_t handed_back;
} gust_newton_run_result;

/* steady.stage = simplec_newton: the driver loop over the SIMPLEC stage and
 * the Newton stage with the switch rules, the hysteresis and the fallbacks,
 * until Steady, Diverged or max_iterations. Synchronises every iteration.
 * Replaces: Driver::run with the stages SIMPLEC and Newton --
 *   Source/Driver.cpp:335-400, 403-492; NewtonStage::init / iterate /
 *   leave -- Source/NewtonStage.cpp:91-207. */
int gust_steady_solve_newton(gust_simplec* s, gust_newton* n, gust_state* q, const gust_driver_options* opt,
                             const gust_switch_options* sw, gust_newton_run_result* result, void* stream);
int gust_switch_default_options(gust_switch_options* sw);

#ifdef __cplusplus
}
#endif

#endif
