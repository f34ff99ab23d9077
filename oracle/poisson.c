/*
 * poisson.c -- oracle restatement of the reference's terrain Poisson operator.
 * Derived in part from ERF (BSD 3-Clause, Copyright (c) 2022 The Regents of the University of California
 * et al.); the full notice is in THIRD_PARTY_NOTICES.md at the repository root.
 *
 * TEST INFRASTRUCTURE ONLY (see gust_oracle.h). Compiled without
 * floating-point contraction, in the reference's operation order:
 *   extern_erf/ERF_TerrainPoisson_3D_K.H:22-196   face gradients (flux_x/y/z)
 *   extern_erf/ERF_TerrainPoisson_3D_K.H:212-275  one-sided Dirichlet ground
 *   extern_erf/ERF_TerrainPoisson_3D_K.H:318-530  terrpoisson_adotx
 *   extern_erf/ERF_TerrainPoisson.cpp:56-66       zero-flux face mask
 *   extern_erf/ERF_TerrainPoisson.cpp:212-264     getFluxes
 *   Source/Poisson.cpp:84-125                     fill_bc_ghosts
 *   Source/Poisson.cpp:162-190, 231-273            Dirichlet ground, z_cross
 *   Source/PoissonCross_K.H:16-58                 z_face_cross
 *   Source/Poisson.cpp:733-764                    face_divergence
 * adotx equals (ax(i+1) G(i+1) - ax(i) G(i)) dxinv + ... with G the face
 * gradients, each computed once per face here: the kernel reuses the same
 * sub-expressions between neighbouring faces, so the values are the same bits.
 */
#include "gust_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define AT(v, i, j, k)                                                                                             \
    ((v)->p[((int64_t)(i) - (v)->begin[0]) + ((int64_t)(j) - (v)->begin[1]) * (v)->jstride                        \
            + ((int64_t)(k) - (v)->begin[2]) * (v)->kstride])

static const double half = 0.5, fourth = 0.25;

/* The low x-face gradient of cell (i,j,k): ERF's px_lo (flux_x = -this). */
static double grad_x(const oa4* x, const oa4* zp, int i, int j, int k, double dxinv)
{
    double h_xi, h_zeta;
    double px_lo = (AT(x, i, j, k) - AT(x, i - 1, j, k)) * dxinv;
    double pz_lo_md_hi = half * (AT(x, i, j, k + 1) + AT(x, i - 1, j, k + 1) - AT(x, i, j, k) - AT(x, i - 1, j, k));
    h_xi = fourth * (AT(zp, i + 1, j, k + 1) - AT(zp, i - 1, j, k + 1) + AT(zp, i + 1, j + 1, k + 1)
                     - AT(zp, i - 1, j + 1, k + 1)) * dxinv;
    h_zeta = fourth * (AT(zp, i, j, k + 2) - AT(zp, i, j, k) + AT(zp, i, j + 1, k + 2) - AT(zp, i, j + 1, k));
    pz_lo_md_hi *= h_xi / h_zeta;
    double pz_lo_md_lo = half * (AT(x, i, j, k) + AT(x, i - 1, j, k) - AT(x, i, j, k - 1) - AT(x, i - 1, j, k - 1));
    h_xi = fourth * (AT(zp, i + 1, j, k) - AT(zp, i - 1, j, k) + AT(zp, i + 1, j + 1, k) - AT(zp, i - 1, j + 1, k))
           * dxinv;
    h_zeta = fourth * (AT(zp, i, j, k + 1) - AT(zp, i, j, k - 1) + AT(zp, i, j + 1, k + 1) - AT(zp, i, j + 1, k - 1));
    pz_lo_md_lo *= h_xi / h_zeta;
    px_lo -= half * (pz_lo_md_hi + pz_lo_md_lo);
    return px_lo;
}

static double grad_y(const oa4* x, const oa4* zp, int i, int j, int k, double dyinv)
{
    double h_eta, h_zeta;
    double py_lo = (AT(x, i, j, k) - AT(x, i, j - 1, k)) * dyinv;
    double pz_md_lo_hi = half * (AT(x, i, j, k + 1) + AT(x, i, j - 1, k + 1) - AT(x, i, j, k) - AT(x, i, j - 1, k));
    h_eta = fourth * (AT(zp, i, j + 1, k + 1) - AT(zp, i, j - 1, k + 1) + AT(zp, i + 1, j + 1, k + 1)
                      - AT(zp, i + 1, j - 1, k + 1)) * dyinv;
    h_zeta = fourth * (AT(zp, i, j, k + 2) - AT(zp, i, j, k) + AT(zp, i + 1, j, k + 2) - AT(zp, i + 1, j, k));
    pz_md_lo_hi *= h_eta / h_zeta;
    double pz_md_lo_lo = half * (AT(x, i, j, k) + AT(x, i, j - 1, k) - AT(x, i, j, k - 1) - AT(x, i, j - 1, k - 1));
    h_eta = fourth * (AT(zp, i, j + 1, k) - AT(zp, i, j - 1, k) + AT(zp, i + 1, j + 1, k) - AT(zp, i + 1, j - 1, k))
            * dyinv;
    h_zeta = fourth * (AT(zp, i, j, k + 1) - AT(zp, i, j, k - 1) + AT(zp, i + 1, j, k + 1) - AT(zp, i + 1, j, k - 1));
    pz_md_lo_lo *= h_eta / h_zeta;
    py_lo -= half * (pz_md_lo_hi + pz_md_lo_lo);
    return py_lo;
}

/* The low z-face gradient of cell (i,j,k) (ERF's pz_lo before the k == 0
 * zeroing). */
static double grad_z(const oa4* x, const oa4* zp, int i, int j, int k, double dxinv, double dyinv)
{
    double h_xi, h_eta, h_zeta;
    double pz_lo = (AT(x, i, j, k) - AT(x, i, j, k - 1));
    const double hzeta_inv_on_zlo =
        8.0 / ((AT(zp, i, j, k + 1) + AT(zp, i + 1, j, k + 1) + AT(zp, i, j + 1, k + 1) + AT(zp, i + 1, j + 1, k + 1))
               - (AT(zp, i, j, k - 1) + AT(zp, i + 1, j, k - 1) + AT(zp, i, j + 1, k - 1) + AT(zp, i + 1, j + 1, k - 1)));
    pz_lo *= hzeta_inv_on_zlo;
    const double px_hi_md_lo =
        half * (AT(x, i + 1, j, k) - AT(x, i, j, k) + AT(x, i + 1, j, k - 1) - AT(x, i, j, k - 1)) * dxinv;
    const double px_lo_md_lo =
        half * (AT(x, i, j, k) - AT(x, i - 1, j, k) + AT(x, i, j, k - 1) - AT(x, i - 1, j, k - 1)) * dxinv;
    const double py_md_hi_lo =
        half * (AT(x, i, j + 1, k) - AT(x, i, j, k) + AT(x, i, j + 1, k - 1) - AT(x, i, j, k - 1)) * dyinv;
    const double py_md_lo_lo =
        half * (AT(x, i, j, k) - AT(x, i, j - 1, k) + AT(x, i, j, k - 1) - AT(x, i, j - 1, k - 1)) * dyinv;
    double pz_hi_md_lo = half * (AT(x, i + 1, j, k) + AT(x, i, j, k) - AT(x, i + 1, j, k - 1) - AT(x, i, j, k - 1));
    h_xi = fourth * (AT(zp, i + 2, j, k) - AT(zp, i, j, k) + AT(zp, i + 2, j + 1, k) - AT(zp, i, j + 1, k)) * dxinv;
    h_zeta = fourth * (AT(zp, i + 1, j, k + 1) - AT(zp, i + 1, j, k - 1) + AT(zp, i + 1, j + 1, k + 1)
                       - AT(zp, i + 1, j + 1, k - 1));
    pz_hi_md_lo *= h_xi / h_zeta;
    double pz_lo_md_lo = half * (AT(x, i, j, k) + AT(x, i - 1, j, k) - AT(x, i, j, k - 1) - AT(x, i - 1, j, k - 1));
    h_xi = fourth * (AT(zp, i + 1, j, k) - AT(zp, i - 1, j, k) + AT(zp, i + 1, j + 1, k) - AT(zp, i - 1, j + 1, k))
           * dxinv;
    h_zeta = fourth * (AT(zp, i, j, k + 1) - AT(zp, i, j, k - 1) + AT(zp, i, j + 1, k + 1) - AT(zp, i, j + 1, k - 1));
    pz_lo_md_lo *= h_xi / h_zeta;
    double pz_md_hi_lo = half * (AT(x, i, j + 1, k) + AT(x, i, j, k) - AT(x, i, j + 1, k - 1) - AT(x, i, j, k - 1));
    h_eta = fourth * (AT(zp, i, j + 2, k) - AT(zp, i, j, k) + AT(zp, i + 1, j + 2, k) - AT(zp, i + 1, j, k)) * dyinv;
    h_zeta = fourth * (AT(zp, i, j + 1, k + 1) - AT(zp, i, j + 1, k - 1) + AT(zp, i + 1, j + 1, k + 1)
                       - AT(zp, i + 1, j + 1, k - 1));
    pz_md_hi_lo *= h_eta / h_zeta;
    double pz_md_lo_lo = half * (AT(x, i, j, k) + AT(x, i, j - 1, k) - AT(x, i, j, k - 1) - AT(x, i, j - 1, k - 1));
    h_eta = fourth * (AT(zp, i, j + 1, k) - AT(zp, i, j - 1, k) + AT(zp, i + 1, j + 1, k) - AT(zp, i + 1, j - 1, k))
            * dyinv;
    h_zeta = fourth * (AT(zp, i, j, k + 1) - AT(zp, i, j, k - 1) + AT(zp, i + 1, j, k + 1) - AT(zp, i + 1, j, k - 1));
    pz_md_lo_lo *= h_eta / h_zeta;
    const double h_xi_on_zlo =
        half * (AT(zp, i + 1, j + 1, k) + AT(zp, i + 1, j, k) - AT(zp, i, j + 1, k) - AT(zp, i, j, k)) * dxinv;
    const double h_eta_on_zlo =
        half * (AT(zp, i + 1, j + 1, k) + AT(zp, i, j + 1, k) - AT(zp, i + 1, j, k) - AT(zp, i, j, k)) * dyinv;
    pz_lo -= half * h_xi_on_zlo * ((px_hi_md_lo + px_lo_md_lo) - (pz_hi_md_lo + pz_lo_md_lo));
    pz_lo -= half * h_eta_on_zlo * ((py_md_hi_lo + py_md_lo_lo) - (pz_md_hi_lo + pz_md_lo_lo));
    return pz_lo;
}

/* The one-sided ground gradient with phi = 0 on the ground
 * (terrpoisson_flux_zlo_dir = -this). */
static double grad_z_dir(const oa4* x, const oa4* zp, int i, int j, int k, double dxinv, double dyinv)
{
    double h_xi, h_eta, h_zeta;
    double pz_lo = AT(x, i, j, k);
    const double hzeta_inv_on_zlo =
        8.0 / ((AT(zp, i, j, k + 1) + AT(zp, i + 1, j, k + 1) + AT(zp, i, j + 1, k + 1) + AT(zp, i + 1, j + 1, k + 1))
               - (AT(zp, i, j, k) + AT(zp, i + 1, j, k) + AT(zp, i, j + 1, k) + AT(zp, i + 1, j + 1, k)));
    pz_lo *= hzeta_inv_on_zlo;
    const double px_hi_md_lo = 0.0, px_lo_md_lo = 0.0, py_md_hi_lo = 0.0, py_md_lo_lo = 0.0;
    double pz_hi_md_lo = half * (AT(x, i + 1, j, k) + AT(x, i, j, k));
    h_xi = fourth * (AT(zp, i + 2, j, k) - AT(zp, i, j, k) + AT(zp, i + 2, j + 1, k) - AT(zp, i, j + 1, k)) * dxinv;
    h_zeta = fourth * (AT(zp, i + 1, j, k + 1) - AT(zp, i + 1, j, k) + AT(zp, i + 1, j + 1, k + 1)
                       - AT(zp, i + 1, j + 1, k));
    pz_hi_md_lo *= h_xi / h_zeta;
    double pz_lo_md_lo = half * (AT(x, i, j, k) + AT(x, i - 1, j, k));
    h_xi = fourth * (AT(zp, i + 1, j, k) - AT(zp, i - 1, j, k) + AT(zp, i + 1, j + 1, k) - AT(zp, i - 1, j + 1, k))
           * dxinv;
    h_zeta = fourth * (AT(zp, i, j, k + 1) - AT(zp, i, j, k) + AT(zp, i, j + 1, k + 1) - AT(zp, i, j + 1, k));
    pz_lo_md_lo *= h_xi / h_zeta;
    double pz_md_hi_lo = half * (AT(x, i, j + 1, k) + AT(x, i, j, k));
    h_eta = fourth * (AT(zp, i, j + 2, k) - AT(zp, i, j, k) + AT(zp, i + 1, j + 2, k) - AT(zp, i + 1, j, k)) * dyinv;
    h_zeta = fourth * (AT(zp, i, j + 1, k + 1) - AT(zp, i, j + 1, k) + AT(zp, i + 1, j + 1, k + 1)
                       - AT(zp, i + 1, j + 1, k));
    pz_md_hi_lo *= h_eta / h_zeta;
    double pz_md_lo_lo = half * (AT(x, i, j, k) + AT(x, i, j - 1, k));
    h_eta = fourth * (AT(zp, i, j + 1, k) - AT(zp, i, j - 1, k) + AT(zp, i + 1, j + 1, k) - AT(zp, i + 1, j - 1, k))
            * dyinv;
    h_zeta = fourth * (AT(zp, i, j, k + 1) - AT(zp, i, j, k) + AT(zp, i + 1, j, k + 1) - AT(zp, i + 1, j, k));
    pz_md_lo_lo *= h_eta / h_zeta;
    const double h_xi_on_zlo =
        half * (AT(zp, i + 1, j + 1, k) + AT(zp, i + 1, j, k) - AT(zp, i, j + 1, k) - AT(zp, i, j, k)) * dxinv;
    const double h_eta_on_zlo =
        half * (AT(zp, i + 1, j + 1, k) + AT(zp, i, j + 1, k) - AT(zp, i + 1, j, k) - AT(zp, i, j, k)) * dyinv;
    pz_lo -= half * h_xi_on_zlo * ((px_hi_md_lo + px_lo_md_lo) - (pz_hi_md_lo + pz_lo_md_lo));
    pz_lo -= half * h_eta_on_zlo * ((py_md_hi_lo + py_md_lo_lo) - (pz_md_hi_lo + pz_md_lo_lo));
    return pz_lo;
}

/* z_face_cross (Source/PoissonCross_K.H:16-58). */
static double z_face_cross(const oa4* x, const oa4* zp, int i, int j, int k, double dxinv, double dyinv)
{
    const double px_hi_md_lo =
        half * (AT(x, i + 1, j, k) - AT(x, i, j, k) + AT(x, i + 1, j, k - 1) - AT(x, i, j, k - 1)) * dxinv;
    const double px_lo_md_lo =
        half * (AT(x, i, j, k) - AT(x, i - 1, j, k) + AT(x, i, j, k - 1) - AT(x, i - 1, j, k - 1)) * dxinv;
    const double py_md_hi_lo =
        half * (AT(x, i, j + 1, k) - AT(x, i, j, k) + AT(x, i, j + 1, k - 1) - AT(x, i, j, k - 1)) * dyinv;
    const double py_md_lo_lo =
        half * (AT(x, i, j, k) - AT(x, i, j - 1, k) + AT(x, i, j, k - 1) - AT(x, i, j - 1, k - 1)) * dyinv;
    double pz_hi_md_lo = half * (AT(x, i + 1, j, k) + AT(x, i, j, k) - AT(x, i + 1, j, k - 1) - AT(x, i, j, k - 1));
    double h_xi = fourth * (AT(zp, i + 2, j, k) - AT(zp, i, j, k) + AT(zp, i + 2, j + 1, k) - AT(zp, i, j + 1, k)) * dxinv;
    double h_zeta = fourth * (AT(zp, i + 1, j, k + 1) - AT(zp, i + 1, j, k - 1) + AT(zp, i + 1, j + 1, k + 1)
                              - AT(zp, i + 1, j + 1, k - 1));
    pz_hi_md_lo *= h_xi / h_zeta;
    double pz_lo_md_lo = half * (AT(x, i, j, k) + AT(x, i - 1, j, k) - AT(x, i, j, k - 1) - AT(x, i - 1, j, k - 1));
    h_xi = fourth * (AT(zp, i + 1, j, k) - AT(zp, i - 1, j, k) + AT(zp, i + 1, j + 1, k) - AT(zp, i - 1, j + 1, k))
           * dxinv;
    h_zeta = fourth * (AT(zp, i, j, k + 1) - AT(zp, i, j, k - 1) + AT(zp, i, j + 1, k + 1) - AT(zp, i, j + 1, k - 1));
    pz_lo_md_lo *= h_xi / h_zeta;
    double pz_md_hi_lo = half * (AT(x, i, j + 1, k) + AT(x, i, j, k) - AT(x, i, j + 1, k - 1) - AT(x, i, j, k - 1));
    double h_eta = fourth * (AT(zp, i, j + 2, k) - AT(zp, i, j, k) + AT(zp, i + 1, j + 2, k) - AT(zp, i + 1, j, k)) * dyinv;
    h_zeta = fourth * (AT(zp, i, j + 1, k + 1) - AT(zp, i, j + 1, k - 1) + AT(zp, i + 1, j + 1, k + 1)
                       - AT(zp, i + 1, j + 1, k - 1));
    pz_md_hi_lo *= h_eta / h_zeta;
    double pz_md_lo_lo = half * (AT(x, i, j, k) + AT(x, i, j - 1, k) - AT(x, i, j, k - 1) - AT(x, i, j - 1, k - 1));
    h_eta = fourth * (AT(zp, i, j + 1, k) - AT(zp, i, j - 1, k) + AT(zp, i + 1, j + 1, k) - AT(zp, i + 1, j - 1, k))
            * dyinv;
    h_zeta = fourth * (AT(zp, i, j, k + 1) - AT(zp, i, j, k - 1) + AT(zp, i + 1, j, k + 1) - AT(zp, i + 1, j, k - 1));
    pz_md_lo_lo *= h_eta / h_zeta;
    const double h_xi_on_zlo =
        half * (AT(zp, i + 1, j + 1, k) + AT(zp, i + 1, j, k) - AT(zp, i, j + 1, k) - AT(zp, i, j, k)) * dxinv;
    const double h_eta_on_zlo =
        half * (AT(zp, i + 1, j + 1, k) + AT(zp, i, j + 1, k) - AT(zp, i + 1, j, k) - AT(zp, i, j, k)) * dyinv;
    return half * h_xi_on_zlo * ((px_hi_md_lo + px_lo_md_lo) - (pz_hi_md_lo + pz_lo_md_lo))
         + half * h_eta_on_zlo * ((py_md_hi_lo + py_md_lo_lo) - (pz_md_hi_lo + pz_md_lo_lo));
}

/* BC codes per face in amrex::Orientation order xlo, ylo, zlo, xhi, yhi, zhi. */
enum { OPF_PERIODIC = 0, OPF_NEUMANN = 1, OPF_DIRICHLET = 2 };

static char g_perr[512];
const char* oracle_poisson_error(void) { return g_perr; }

static int check_bc(const ogeom* g, const int* bc)
{
    const char* nm[3] = {"x", "y", "z"};
    for (int d = 0; d < 3; ++d) {
        for (int s = 0; s < 2; ++s) {
            const int f = bc[d + 3 * s];
            if ((g->periodic[d] != 0) != (f == OPF_PERIODIC)) {
                snprintf(g_perr, sizeof g_perr,
                         "gust: invalid inputs:\n  Poisson BC %s%s: periodic exactly when the domain is periodic in %s",
                         nm[d], s ? "hi" : "lo", nm[d]);
                return -1;
            }
        }
    }
    return 0;
}

/* fill_bc_ghosts (Source/Poisson.cpp:84-125) on the one box: periodic
 * images, then the even/odd copies x, y (over the x ghosts), z (over both). */
int oracle_poisson_fill_ghosts(oa4* phi, const ogeom* g, const int* bc)
{
    if (check_bc(g, bc)) { return -1; }
    const int lo[3] = {g->lo[0], g->lo[1], g->lo[2]}, hi[3] = {g->hi[0], g->hi[1], g->hi[2]};
    const int L[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
    /* FillBoundary: one ghost layer, periodic images. */
    for (int k = lo[2] - 1; k <= hi[2] + 1; ++k)
        for (int j = lo[1] - 1; j <= hi[1] + 1; ++j)
            for (int i = lo[0] - 1; i <= hi[0] + 1; ++i) {
                int iv[3] = {i, j, k}, inside = 1, ok = 1;
                for (int d = 0; d < 3; ++d) {
                    if (iv[d] >= lo[d] && iv[d] <= hi[d]) { continue; }
                    inside = 0;
                    if (!g->periodic[d] || d == 2) { ok = 0; break; }
                    iv[d] += iv[d] < lo[d] ? L[d] : -L[d];
                }
                if (!inside && ok) { AT(phi, i, j, k) = AT(phi, iv[0], iv[1], iv[2]); }
            }
    for (int d = 0; d < 3; ++d) {
        if (d < 2 && g->periodic[d]) { continue; }
        for (int s = 0; s < 2; ++s) {
            const int f = bc[d + 3 * s];
            const double sign = f == OPF_DIRICHLET ? -1.0 : 1.0;
            int blo[3], bhi[3];
            for (int e = 0; e < 3; ++e) {
                blo[e] = lo[e];
                bhi[e] = hi[e];
                if (e == d) { continue; }
                blo[e] -= 1;
                bhi[e] += 1;
                if (e > d && !(e < 2 && g->periodic[e])) {
                    if (blo[e] < lo[e]) { blo[e] = lo[e]; }
                    if (bhi[e] > hi[e]) { bhi[e] = hi[e]; }
                }
            }
            const int edge = s == 0 ? lo[d] : hi[d];
            blo[d] = bhi[d] = s == 0 ? edge - 1 : edge + 1;
            const int st = s == 0 ? 1 : -1;
            const int di = d == 0 ? st : 0, dj = d == 1 ? st : 0, dk = d == 2 ? st : 0;
            for (int k = blo[2]; k <= bhi[2]; ++k)
                for (int j = blo[1]; j <= bhi[1]; ++j)
                    for (int i = blo[0]; i <= bhi[0]; ++i) {
                        AT(phi, i, j, k) = sign * AT(phi, i + di, j + dj, k + dk);
                    }
        }
    }
    return 0;
}

/* The zero-flux mask (extern_erf/ERF_TerrainPoisson.cpp:56-66): every
 * non-periodic face that is not Dirichlet, except the ground. */
static int zero_flux_faces(const int* bc)
{
    int zf = 0;
    if (bc[0] == OPF_NEUMANN) { zf |= 1; }
    if (bc[3] == OPF_NEUMANN) { zf |= 2; }
    if (bc[1] == OPF_NEUMANN) { zf |= 4; }
    if (bc[4] == OPF_NEUMANN) { zf |= 8; }
    if (bc[5] == OPF_NEUMANN) { zf |= 32; }
    return zf;
}

int oracle_poisson_zero_flux_mask(const int* bc) { return zero_flux_faces(bc); }

/* A phi on the valid cells (Source/Poisson.cpp:140-190, 231-254):
 * ghosts of phi filled for bc, ERF's adotx with the zero-flux mask, the
 * Dirichlet ground face, the z_cross correction (dcross may be null).
 * ax, ay, az are the premultiplied areas a beta. */
int oracle_poisson_apply(oa4* y, oa4* phi, const oa4* ax, const oa4* ay, const oa4* az, const oa4* J,
                         const oa4* zp, const oa4* dcross, const oa4* az_mesh, const ogeom* g, const int* bc)
{
    if (oracle_poisson_fill_ghosts(phi, g, bc)) { return -1; }
    const double dxinv = 1.0 / g->dx[0], dyinv = 1.0 / g->dx[1], dzinv = 1.0 / g->dx[2];
    const int zf = zero_flux_faces(bc);
    const int dir_ground = bc[2] == OPF_DIRICHLET;
    const int top_open = bc[5] != OPF_NEUMANN;
    const int k0 = g->lo[2], ktop = g->hi[2] + 1;
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                double px_hi = grad_x(phi, zp, i + 1, j, k, dxinv);
                double px_lo = grad_x(phi, zp, i, j, k, dxinv);
                double py_hi = grad_y(phi, zp, i, j + 1, k, dyinv);
                double py_lo = grad_y(phi, zp, i, j, k, dyinv);
                double pz_hi = grad_z(phi, zp, i, j, k + 1, dxinv, dyinv);
                double pz_lo = grad_z(phi, zp, i, j, k, dxinv, dyinv);
                if (k == 0) { pz_lo = 0.0; }
                if ((zf & 1) && i == g->lo[0]) { px_lo = 0.0; }
                if ((zf & 2) && i == g->hi[0]) { px_hi = 0.0; }
                if ((zf & 4) && j == g->lo[1]) { py_lo = 0.0; }
                if ((zf & 8) && j == g->hi[1]) { py_hi = 0.0; }
                if ((zf & 32) && k == g->hi[2]) { pz_hi = 0.0; }
                double v = (AT(ax, i + 1, j, k) * px_hi - AT(ax, i, j, k) * px_lo) * dxinv
                           + (AT(ay, i, j + 1, k) * py_hi - AT(ay, i, j, k) * py_lo) * dyinv
                           + (AT(az, i, j, k + 1) * pz_hi - AT(az, i, j, k) * pz_lo) * dzinv;
                v /= AT(J, i, j, k);
                if (dir_ground && k == k0) {
                    /* y += az * flux_zlo_dir * dzinv / J, flux = -gradient */
                    v += AT(az, i, j, k) * (-grad_z_dir(phi, zp, i, j, k, dxinv, dyinv)) * dzinv / AT(J, i, j, k);
                }
                if (dcross) {
                    double fhi = 0.0, flo = 0.0;
                    if (!(k + 1 == k0 || (k + 1 == ktop && !top_open))) {
                        fhi = AT(az_mesh, i, j, k + 1) * AT(dcross, i, j, k + 1) * z_face_cross(phi, zp, i, j, k + 1, dxinv, dyinv);
                    }
                    if (!(k == k0 || (k == ktop && !top_open))) {
                        flo = AT(az_mesh, i, j, k) * AT(dcross, i, j, k) * z_face_cross(phi, zp, i, j, k, dxinv, dyinv);
                    }
                    v += (fhi - flo) * dzinv / AT(J, i, j, k);
                }
                AT(y, i, j, k) = v;
            }
    return 0;
}

/* The fluxes F = -gradient on every face of the valid box, zero on the
 * masked faces (getFluxes, extern_erf/ERF_TerrainPoisson.cpp:212-264, and
 * fluxes_dirichlet_ground, Source/Poisson.cpp:192-227), times the face
 * coefficients bx/by/bz (null: 1), minus the z_cross part
 * (add_cross_flux, Source/Poisson.cpp:256-273; dcross may be null).
 * phi's ghosts are filled for bc. */
int oracle_poisson_fluxes(oa4* fx, oa4* fy, oa4* fz, oa4* phi, const oa4* zp, const oa4* bx, const oa4* by,
                          const oa4* bz, const oa4* dcross, const ogeom* g, const int* bc)
{
    if (oracle_poisson_fill_ghosts(phi, g, bc)) { return -1; }
    const double dxinv = 1.0 / g->dx[0], dyinv = 1.0 / g->dx[1];
    const int zf = zero_flux_faces(bc);
    const int dir_ground = bc[2] == OPF_DIRICHLET;
    const int top_open = bc[5] != OPF_NEUMANN;
    const int k0 = g->lo[2], ktop = g->hi[2] + 1;
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0] + 1; ++i) {
                const int zero = ((zf & 1) && i == g->lo[0]) || ((zf & 2) && i == g->hi[0] + 1);
                double f = zero ? 0.0 : -grad_x(phi, zp, i, j, k, dxinv);
                if (bx) { f *= AT(bx, i, j, k); }
                AT(fx, i, j, k) = f;
            }
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1]; j <= g->hi[1] + 1; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                const int zero = ((zf & 4) && j == g->lo[1]) || ((zf & 8) && j == g->hi[1] + 1);
                double f = zero ? 0.0 : -grad_y(phi, zp, i, j, k, dyinv);
                if (by) { f *= AT(by, i, j, k); }
                AT(fy, i, j, k) = f;
            }
    for (int k = g->lo[2]; k <= g->hi[2] + 1; ++k)
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                double f;
                if ((zf & 32) && k == g->hi[2] + 1) {
                    f = 0.0;
                } else if (dir_ground && k == k0) {
                    f = -grad_z_dir(phi, zp, i, j, k, dxinv, dyinv);
                } else {
                    f = -grad_z(phi, zp, i, j, k, dxinv, dyinv);
                    if (k == 0) { f = -0.0; }   /* flux_z returns -pz_lo with pz_lo = 0 at k == 0 */
                }
                if (bz) { f *= AT(bz, i, j, k); }
                if (dcross && !(k == k0 || (k == ktop && !top_open))) {
                    f -= AT(dcross, i, j, k) * z_face_cross(phi, zp, i, j, k, dxinv, dyinv);
                }
                AT(fz, i, j, k) = f;
            }
    return 0;
}

/* face_divergence (Source/Poisson.cpp:733-764). */
void oracle_face_divergence(oa4* div, const oa4* fx, const oa4* fy, const oa4* fz, const oa4* ax, const oa4* ay,
                            const oa4* az, const oa4* J, const ogeom* g)
{
    const double dxinv = 1.0 / g->dx[0], dyinv = 1.0 / g->dx[1], dzinv = 1.0 / g->dx[2];
    for (int k = g->lo[2]; k <= g->hi[2]; ++k)
        for (int j = g->lo[1]; j <= g->hi[1]; ++j)
            for (int i = g->lo[0]; i <= g->hi[0]; ++i) {
                AT(div, i, j, k) = ((AT(ax, i + 1, j, k) * AT(fx, i + 1, j, k) - AT(ax, i, j, k) * AT(fx, i, j, k)) * dxinv
                                    + (AT(ay, i, j + 1, k) * AT(fy, i, j + 1, k) - AT(ay, i, j, k) * AT(fy, i, j, k)) * dyinv
                                    + (AT(az, i, j, k + 1) * AT(fz, i, j, k + 1) - AT(az, i, j, k) * AT(fz, i, j, k)) * dzinv)
                                   / AT(J, i, j, k);
            }
}
