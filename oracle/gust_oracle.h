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

#ifdef __cplusplus
}
#endif

#endif
