/*
 * gust_oracle.h -- the FP64 CPU oracle of the SIMPLEC hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * solver's arithmetic (GUST, hgopalan/ablsteadyrans), written operation by
 * operation in the reference's order so that the HIP library can be checked
 * against it. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it; the product library never links or calls it.
 *
 * Every function cites the reference file:line it restates. Each one is
 * pinned by checks in tests/test_oracle_*.py against closed forms, dense
 * solves, invariants and the reference's own fixtures; a function without
 * such a check says "parity unpinned" next to its declaration.
 *
 * Data: Array4-compatible views (AMReX Array4 semantics, i fastest):
 *   element (i,j,k,n) at p[(i-begin[0]) + (j-begin[1])*jstride
 *                           + (k-begin[2])*kstride + n*nstride],
 *   begin inclusive, end exclusive (the box of the view, ghosts included).
 * A tile is one box spanning its whole domain (x/y periodic or not).
 */
#ifndef GUST_ORACLE_H_
#define GUST_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double* p;
    int64_t jstride, kstride, nstride;
    int32_t begin[3], end[3];
    int32_t ncomp;
} oa4;

typedef struct {
    int32_t lo[3], hi[3];    /* the cell-centred domain box (inclusive) */
    int32_t periodic[3];     /* 1 = periodic along that axis (z never) */
    double dx[3];            /* dx, dy, dzeta */
} ogeom;

/* Field locations (Source/CVFluxes.H:22): bit d set = nodal along d. */
enum { OLOC_CELL = 0, OLOC_XFACE = 1, OLOC_YFACE = 2, OLOC_ZFACE = 4 };

/* Coefficient components (Source/ImplicitOperator.H:57). */
enum { OC_P = 0, OC_XM, OC_XP, OC_YM, OC_YP, OC_ZM, OC_ZP, OC_N };

/* Line sweep options (Source/LineSweeps.H:80-104). */
enum { OGHOSTS_LAGGED = 0, OGHOSTS_ZERO = 1 };
enum { OORDER_SEQUENTIAL = 0, OORDER_DIAGONAL = 1 };
#define O_MAX_LINE_OVERLAP 8

const char* oracle_last_error(void);

/* The exact column solve (Source/LineSolve_K.H:22-45). d is read at the
 * x/y neighbours, c/e are Thomas scratch at loc, out receives the column. */
void oracle_solve_column(const oa4* a, const oa4* dg, const oa4* r, const oa4* d, oa4* c, oa4* e, oa4* out,
                         int i, int j, int klo, int khi);

/* z-line Jacobi (Source/Predictor.cpp:87-127 with check_line_system,
 * Source/LineSweeps.cpp:391-413). delta needs one ghost cell; it is zeroed,
 * swept, and the periodic seam copies synchronised (the low copy wins).
 * Returns 0 and the last sweep's max |change| in *change, or -1. */
int oracle_zline_jacobi(int loc, const oa4* A, const oa4* ap, const oa4* rhs, int sweeps, const ogeom* g,
                        oa4* delta, double* change);
/* The same on `threads` OpenMP threads (the columns split; the same bits). */
int oracle_zline_jacobi_threads(int loc, const oa4* A, const oa4* ap, const oa4* rhs, int sweeps, const ogeom* g,
                                oa4* delta, double* change, int threads);

/* Line LU-SGS (Source/LineSweeps.cpp:181-305, 415-428): lagged or zero
 * ghosts, overlap, sequential or diagonal order, alternate 0/1.
 * The diagonal order is executed plane by plane (LineSweeps.cpp:66-94). */
int oracle_zline_lusgs(int loc, const oa4* A, const oa4* ap, const oa4* rhs, int sweeps, const ogeom* g,
                       oa4* delta, int ghosts, int overlap, int order, int alternate, double* change);

/* Number of column-pass kernel launches the reference would make
 * (line_sweep_launches, Source/LineSweeps.H:106-111) for the last call. */
uint64_t oracle_last_lusgs_launches(void);

/* Terrain Poisson operator (oracle/poisson.c). bc: 6 codes in
 * amrex::Orientation order xlo, ylo, zlo, xhi, yhi, zhi; 0 periodic,
 * 1 neumann, 2 dirichlet (Source/Poisson.H:20-34). phi needs one ghost cell,
 * zp (nodal heights) one ghost node. */
const char* oracle_poisson_error(void);
int oracle_poisson_zero_flux_mask(const int* bc);
int oracle_poisson_fill_ghosts(oa4* phi, const ogeom* g, const int* bc);
int oracle_poisson_apply(oa4* y, oa4* phi, const oa4* ax, const oa4* ay, const oa4* az, const oa4* J,
                         const oa4* zp, const oa4* dcross, const oa4* az_mesh, const ogeom* g, const int* bc);
int oracle_poisson_fluxes(oa4* fx, oa4* fy, oa4* fz, oa4* phi, const oa4* zp, const oa4* bx, const oa4* by,
                          const oa4* bz, const oa4* dcross, const ogeom* g, const int* bc);
void oracle_face_divergence(oa4* div, const oa4* fx, const oa4* fy, const oa4* fz, const oa4* ax, const oa4* ay,
                            const oa4* az, const oa4* J, const ogeom* g);

/* Mass-flux operators of the terrain-following C-grid (oracle/massflux.c,
 * Source/MassFlux.cpp). z is z_phys_nd; views are read at the ghost points
 * the reference's operators read (the caller fills them). */
void oracle_omega_from_w(oa4* omega, const oa4* u, const oa4* v, const oa4* w, const oa4* z, const ogeom* g,
                         int ngrow);
void oracle_w_from_omega(oa4* w, const oa4* u, const oa4* v, const oa4* omega, const oa4* z, const ogeom* g);
void oracle_transport_fluxes(oa4* Fx, oa4* Fy, oa4* Fz, const oa4* u, const oa4* v, const oa4* w, const oa4* z,
                             const oa4* ax, const oa4* ay, const ogeom* g);
/* omega_scratch: valid z-faces, receives Omega. */
void oracle_velocity_divergence(oa4* div, const oa4* u, const oa4* v, const oa4* w, const oa4* z, const oa4* ax,
                                const oa4* ay, const oa4* az, const oa4* J, const ogeom* g, oa4* omega_scratch);
void oracle_pressure_gradient(oa4* gx, oa4* gy, oa4* gz, const oa4* p, const oa4* z, const oa4* J, const oa4* ax,
                              const oa4* ay, const ogeom* g);

/* Pseudo-time (oracle/pseudotime.c). local_dtau returns -1 for a cfl that is
 * not finite and > 0; face_dtau returns -1 unless dtau_cc is finite and > 0
 * everywhere. nu may be null. */
int oracle_local_dtau(oa4* dtau, const oa4* u, const oa4* v, const oa4* w, const oa4* z, const oa4* J, const oa4* nu,
                      double cfl, const ogeom* g, oa4* omega_scratch);
double oracle_ser_cfl(double cfl, double r, double cfl0, double cfl_max, double p, double growth);
int oracle_face_dtau(oa4* fx, oa4* fy, oa4* fz, const oa4* d, const ogeom* g);

/* The implicit operator A1 (oracle/implicit.c, Source/ImplicitOperator.cpp).
 * A: 7 components at loc on the valid points; F: the transport fluxes with
 * one ghost; g: cell diffusivity with one ghost; logz: the logarithmic
 * z-face mean of g. Returns -1 for an unknown loc. */
int oracle_implicit_coefficients(int loc, const oa4* fx, const oa4* fy, const oa4* fz, const oa4* g, const oa4* z,
                                 const oa4* J, const ogeom* geo, int logz, oa4* A);
int oracle_apply_implicit(int loc, const oa4* A, const oa4* q, const ogeom* geo, oa4* out);

/* Convection and diffusion over control volumes (oracle/transport.c,
 * Source/Convection.cpp, Source/Diffusion.cpp, Source/CVFluxes.cpp).
 * limiter: 0 van Leer, 1 van Albada, 2 upwind (Source/Limiter_K.H:19).
 * G[d]: at loc ^ (1 << d), covering cv_flux_box(loc, valid, d). q needs two
 * ghosts for convection, one for diffusion; F and the diffusivities one.
 * convective_fluxes returns -1 for a bad loc/limiter, -2 for a bad eps. */
int oracle_convective_fluxes(int loc, const oa4* q, const oa4* fx, const oa4* fy, const oa4* fz, int limiter,
                             double eps, const ogeom* g, oa4* G0, oa4* G1, oa4* G2);
int oracle_flux_divergence(int loc, const oa4* G0, const oa4* G1, const oa4* G2, const oa4* J, const ogeom* g,
                           oa4* out);
int oracle_subtract_transport_divergence(int loc, const oa4* q, const oa4* fx, const oa4* fy, const oa4* fz,
                                         const oa4* J, const ogeom* g, oa4* out);
int oracle_scalar_diffusive_fluxes(const oa4* q, const oa4* gamma, const oa4* z, const ogeom* g, int logz, oa4* G0,
                                   oa4* G1, oa4* G2);
int oracle_momentum_diffusive_fluxes(int c, const oa4* u, const oa4* v, const oa4* w, const oa4* nu, const oa4* z,
                                     const ogeom* g, int logz, oa4* G0, oa4* G1, oa4* G2);

/* Reductions of the control loop (oracle/reductions.c, Source/Reductions.cpp,
 * Source/Residual.cpp:101-150, Source/Simplec.cpp:374-445). lo/hi: the
 * inclusive box of the points reduced. stage_norms returns the number of
 * monitors (2, or 3 with ustar: the ground slab at k = lo[2]). */
double oracle_box_sum(const oa4* a, const int lo[3], const int hi[3]);
double oracle_global_max(const oa4* a, const int lo[3], const int hi[3]);
void oracle_scaled_sums(const oa4* q, const oa4* v, const oa4* off, double s, int loc, const ogeom* g,
                        double out[3]);
void oracle_residual_scales(double U, double theta_s, double k_s, double H, double scale[6]);
int oracle_stage_norms(const oa4* const R[6], const oa4* u, const oa4* v, const oa4* w, const oa4* tke,
                       const oa4* ustar, double U, double theta_s, double k_s, double H, const ogeom* g,
                       double res[6], double mon[3]);
void oracle_confidence(const oa4* const R[6], const oa4* z, double U, double theta_s, double k_s, double H,
                       const double threshold[6], const ogeom* g, double* share, double* top);

/* The residual's physics (oracle/physics.c). Sides periodic in x and y.
 * Face types as Source/BoundaryConditions.H:76 (FaceType). */
enum { OFACE_PERIODIC = 0, OFACE_INFLOW, OFACE_OUTFLOW, OFACE_SLIP_WALL, OFACE_NO_SLIP_WALL, OFACE_SURFACE_LAYER,
       OFACE_GEOSTROPHIC };
/* Components of the surface-state slab (Source/SurfaceLayer.H:163-174). */
enum { OSS_USTAR = 0, OSS_SPEED, OSS_D0, OSS_TAU0, OSS_TAU1, OSS_TAU2, OSS_DRAG, OSS_KWALL, OSS_AREA, OSS_INVL,
       OSS_N };

typedef struct {   /* Source/Closure.H:78-105 (ClosureOptions) */
    double Cmu0, Cb, Rt_crit, Rt_min, max_geom_lscale, tke_floor, sigma_k, nu_molecular, alpha_molecular, kappa;
} oclosure;

typedef struct {   /* Source/SurfaceLayer.H:100-146 (neutral, first centre) */
    double z0, Cmu0, kappa;
    int wall_k_dirichlet, floor_one;
} osurface;

typedef struct {   /* Source/Forcing.H:40-96 */
    double gravity, theta0;
    int coriolis, three_d, geostrophic;
    double latitude, rotation_period, ug, vg;
    double pgrad[3];
} oforcing;

typedef struct {   /* Source/InflowProfile.H:95-133 (View) */
    int kind;                     /* 0 log_law, 1 power_law, 2 table */
    double ustar, z0, speed_ref, z_ref, alpha, cosd, sind, Cmu0, kappa, inv_lmax;
    const double *tz, *tu, *tv, *tt, *tk;
    int n, has_k;
    const double *rz, *rt;        /* the sounding (analytic kinds' theta) */
    int rn;
} oprofile;

typedef struct {   /* Source/BoundaryConditions.H:107-121 (Rayleigh and sponge layers) */
    double rayleigh_zdamp, rayleigh_dampcoef;
    int rayleigh_fields[5];       /* u, v, w, theta, k */
    int rayleigh_above_ground;
    double sponge_width, sponge_strength;
    double ztop;                  /* the mesh's top level */
    double prob_lo[2], prob_hi[2];
} odamping;

typedef struct {   /* Source/BoundaryConditions.H:82-100 */
    int face[6];                  /* OFACE_*, orientation order xlo, ylo, zlo, xhi, yhi, zhi */
    double zhi_theta_grad;
    double wall_velocity[6][3];
    double ug, vg;                /* a geostrophic top's wind */
    const oprofile* profile;      /* the inflow profile (an inflow face or the damping needs it) */
    odamping damping;
} obc;

/* Source/BoundaryConditions.cpp:1022-1068: R += c (q - q_ref) on the valid
 * points of u, v, w, theta, k (q with two ghosts, R begin lo). */
void oracle_add_damping(oa4* const q[6], const oa4* z, const obc* bc, const ogeom* g, oa4* const R[6]);

/* Source/InflowProfile.H:149-200: (u, v, theta, k, nu_t) at height z above the ground. */
void oracle_profile_eval(const oprofile* p, double z, double out[5]);

void oracle_height_above_ground(const oa4* z, const ogeom* g, oa4* dist);
/* Returns -1 if the wall distance plus z0 is not > 0 somewhere. */
int oracle_closure_evaluate(const oa4* u, const oa4* v, const oa4* w, const oa4* theta, const oa4* tke, const oa4* z,
                            const oa4* dist, double z0, const oclosure* opt, double gravity, double theta0,
                            const ogeom* g, oa4* nu, oa4* gth, oa4* gk, oa4* Ps, oa4* Pb, oa4* eps, oa4* dSdk);
/* ss: OSS_N components over the ground columns and one lateral ghost column. */
int oracle_surface_state(const oa4* u, const oa4* v, const oa4* z, const osurface* o, const ogeom* g, oa4* ss);
/* q: u, v, w, p, theta, k with two ghosts; -3 if an inflow face has no profile. */
int oracle_fill_ghosts(oa4* const q[6], const oa4* z, const obc* bc, const osurface* so, const ogeom* g, oa4* ss);
int oracle_fill_ghosts_homogeneous(oa4* const q[6], const oa4* z, const obc* bc, const ogeom* g);
double oracle_coriolis_parameter(int enabled, double latitude, double rotation_period);
void oracle_fill_theta_ref(int n, const double* zs, const double* ts, const oa4* z, oa4* theta_ref);
void oracle_forcing(const oa4* u, const oa4* v, const oa4* w, const oa4* theta, const oa4* theta_ref,
                    const oforcing* f, const ogeom* g, oa4* Ru, oa4* Rv, oa4* Rw);
void oracle_set_ground_fluxes(const oa4* ss, const ogeom* g, oa4* Gu, oa4* Gv, oa4* Gt, oa4* Gk);
void oracle_apply_wall_k(const oa4* ss, const oa4* tke, int dirichlet, const ogeom* g, oa4* Rk);
void oracle_apply_rows(const oa4* u, const oa4* v, const oa4* w, const oa4* z, const obc* bc, const ogeom* g,
                       oa4* Ru, oa4* Rv, oa4* Rw);
/* Source/BoundaryConditions.cpp:849-876: 1 at the Dirichlet face points of eq at its location, else 0. */
void oracle_dirichlet_rows(int eq, const obc* bc, const ogeom* g, oa4* mask);
/* eq 0..4 = u, v, w, theta, k; dSdk may be null. */
void oracle_implicit_folds(int eq, const oa4* ss, const oa4* dSdk, const oa4* J, const oa4* zn, const obc* bc,
                           const osurface* so, const ogeom* g, oa4* A);

#ifdef __cplusplus
}
#endif

#endif
