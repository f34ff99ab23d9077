/*
 * line_sweeps.c -- oracle restatement of the reference's z-line solvers.
 *
 * TEST INFRASTRUCTURE ONLY (see gust_oracle.h). Compiled with
 * -ffp-contract=off: the reference builds ISO C++20 (CMakeLists.txt:21-24),
 * where GCC does not contract a*b+c into fused multiply-adds, so every
 * expression below rounds as the reference's CPU build does.
 *
 * Restated from:
 *   Source/LineSolve_K.H:22-45      solve_column (Thomas with x/y neighbours)
 *   Source/Predictor.cpp:87-127     zline_jacobi
 *   Source/LineSweeps.cpp:38-155    sequential / diagonal passes (+ j-reversed)
 *   Source/LineSweeps.cpp:181-305   lusgs_lagged, lusgs_zero
 *   Source/LineSweeps.cpp:315-413   option checks, ownership, sweep region
 *   Source/LineSweeps.cpp:415-428   zline_lusgs
 * The domain is one box (a tile): AMReX FillBoundary then fills the periodic
 * images of that box into its own ghosts, and OverrideSync reconciles the two
 * copies of a periodic seam face.
 */
#include "gust_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[1024];
static uint64_t g_launches;

const char* oracle_last_error(void) { return g_err; }
uint64_t oracle_last_lusgs_launches(void) { return g_launches; }

static int fail(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return -1;
}

#define AT(v, i, j, k, n)                                                                          \
    ((v)->p[((int64_t)(i) - (v)->begin[0]) + ((int64_t)(j) - (v)->begin[1]) * (v)->jstride        \
            + ((int64_t)(k) - (v)->begin[2]) * (v)->kstride + (int64_t)(n) * (v)->nstride])

typedef struct { int lo[3], hi[3]; } obox;   /* inclusive */

static obox loc_box(const ogeom* g, int loc)
{
    obox b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = g->lo[d];
        b.hi[d] = g->hi[d] + ((loc >> d) & 1);
    }
    return b;
}

static int box_len(const obox* b, int d) { return b->hi[d] - b->lo[d] + 1; }

static int box_contains(const obox* b, int i, int j, int k)
{
    return i >= b->lo[0] && i <= b->hi[0] && j >= b->lo[1] && j <= b->hi[1] && k >= b->lo[2] && k <= b->hi[2];
}

/* A temporary FAB over box b with ncomp components (zero-filled). */
static oa4 fab_new(obox b, int ncomp)
{
    oa4 v;
    int64_t nx = box_len(&b, 0), ny = box_len(&b, 1), nz = box_len(&b, 2);
    v.p = (double*)calloc((size_t)(nx * ny * nz * ncomp), sizeof(double));
    v.jstride = nx;
    v.kstride = nx * ny;
    v.nstride = nx * ny * nz;
    for (int d = 0; d < 3; ++d) { v.begin[d] = b.lo[d]; v.end[d] = b.hi[d] + 1; }
    v.ncomp = ncomp;
    return v;
}

static void fab_free(oa4* v) { free(v->p); v->p = NULL; }

static obox view_box(const oa4* v)
{
    obox b;
    for (int d = 0; d < 3; ++d) { b.lo[d] = v->begin[d]; b.hi[d] = v->end[d] - 1; }
    return b;
}

static int view_covers(const oa4* v, const obox* b, int ncomp)
{
    if (!v || !v->p || v->ncomp < ncomp) { return 0; }
    for (int d = 0; d < 3; ++d) {
        if (v->begin[d] > b->lo[d] || v->end[d] - 1 < b->hi[d]) { return 0; }
    }
    return 1;
}

static void set_all(oa4* v, double val)
{
    const obox b = view_box(v);
    for (int n = 0; n < v->ncomp; ++n)
        for (int k = b.lo[2]; k <= b.hi[2]; ++k)
            for (int j = b.lo[1]; j <= b.hi[1]; ++j)
                for (int i = b.lo[0]; i <= b.hi[0]; ++i) { AT(v, i, j, k, n) = val; }
}

static void copy_box(oa4* dst, const oa4* src, const obox* b, int ncomp)
{
    for (int n = 0; n < ncomp; ++n)
        for (int k = b->lo[2]; k <= b->hi[2]; ++k)
            for (int j = b->lo[1]; j <= b->hi[1]; ++j)
                for (int i = b->lo[0]; i <= b->hi[0]; ++i) { AT(dst, i, j, k, n) = AT(src, i, j, k, n); }
}

/* AMReX FillBoundary(periodicity) on a one-box domain: every ghost point of
 * v outside the valid box V whose out-of-range axes are all periodic takes
 * the value of its periodic image inside V (shifts by the domain's cell
 * count). Ghosts across a non-periodic face are not touched. */
static void fill_boundary(oa4* v, const obox* V, const ogeom* g)
{
    const obox b = view_box(v);
    const int L[3] = {g->hi[0] - g->lo[0] + 1, g->hi[1] - g->lo[1] + 1, g->hi[2] - g->lo[2] + 1};
    for (int n = 0; n < v->ncomp; ++n)
        for (int k = b.lo[2]; k <= b.hi[2]; ++k)
            for (int j = b.lo[1]; j <= b.hi[1]; ++j)
                for (int i = b.lo[0]; i <= b.hi[0]; ++i) {
                    if (box_contains(V, i, j, k)) { continue; }
                    int iv[3] = {i, j, k}, ok = 1;
                    for (int d = 0; d < 3 && ok; ++d) {
                        if (iv[d] >= V->lo[d] && iv[d] <= V->hi[d]) { continue; }
                        if (d == 2 || !g->periodic[d]) { ok = 0; break; }
                        while (iv[d] < V->lo[d]) { iv[d] += L[d]; }
                        while (iv[d] > V->hi[d]) { iv[d] -= L[d]; }
                    }
                    if (ok) { AT(v, i, j, k, n) = AT(v, iv[0], iv[1], iv[2], n); }
                }
}

/* OverrideSync of a one-box domain: along a periodic axis in which loc is
 * nodal, the seam face exists twice (index lo and lo + L). low_owner = 1:
 * AMReX's default owner mask (the unshifted low copy owns; the high copy is
 * overwritten). low_owner = 0: the mask of line_owned_box
 * (Source/LineSweeps.cpp:361-374: the domain's high face owns). */
static void override_sync(oa4* v, int loc, const obox* V, const ogeom* g, int low_owner)
{
    for (int d = 0; d < 2; ++d) {
        if (!((loc >> d) & 1) || !g->periodic[d]) { continue; }
        const int lo = V->lo[d], hi = V->hi[d];
        obox face = *V;
        face.lo[d] = face.hi[d] = 0;
        for (int k = face.lo[2]; k <= face.hi[2]; ++k)
            for (int j = face.lo[1]; j <= face.hi[1]; ++j)
                for (int i = face.lo[0]; i <= face.hi[0]; ++i) {
                    int a[3] = {i, j, k}, b[3] = {i, j, k};
                    a[d] = lo;
                    b[d] = hi;
                    if (low_owner) { AT(v, b[0], b[1], b[2], 0) = AT(v, a[0], a[1], a[2], 0); }
                    else           { AT(v, a[0], a[1], a[2], 0) = AT(v, b[0], b[1], b[2], 0); }
                }
    }
}

/* Source/LineSolve_K.H:22-45. */
void oracle_solve_column(const oa4* a, const oa4* dg, const oa4* r, const oa4* d, oa4* c, oa4* e, oa4* out,
                         int i, int j, int klo, int khi)
{
    for (int k = klo; k <= khi; ++k) {
        const double rk = AT(r, i, j, k, 0) + AT(a, i, j, k, OC_XM) * AT(d, i - 1, j, k, 0)
                          + AT(a, i, j, k, OC_XP) * AT(d, i + 1, j, k, 0)
                          + AT(a, i, j, k, OC_YM) * AT(d, i, j - 1, k, 0)
                          + AT(a, i, j, k, OC_YP) * AT(d, i, j + 1, k, 0);
        const double lower = -AT(a, i, j, k, OC_ZM), upper = -AT(a, i, j, k, OC_ZP);
        if (k == klo) {
            const double m = AT(dg, i, j, k, 0);
            AT(c, i, j, k, 0) = upper / m;
            AT(e, i, j, k, 0) = rk / m;
        } else {
            const double m = AT(dg, i, j, k, 0) - lower * AT(c, i, j, k - 1, 0);
            AT(c, i, j, k, 0) = upper / m;
            AT(e, i, j, k, 0) = (rk - lower * AT(e, i, j, k - 1, 0)) / m;
        }
    }
    AT(out, i, j, khi, 0) = AT(e, i, j, khi, 0);
    for (int k = khi - 1; k >= klo; --k) {
        AT(out, i, j, k, 0) = AT(e, i, j, k, 0) - AT(c, i, j, k, 0) * AT(out, i, j, k + 1, 0);
    }
}

static const char* loc_word(int loc)
{
    switch (loc) {
    case OLOC_CELL: return "cell centres";
    case OLOC_XFACE: return "x-faces";
    case OLOC_YFACE: return "y-faces";
    case OLOC_ZFACE: return "z-faces";
    default: return "?";
    }
}

/* check_line_system (Source/LineSweeps.cpp:391-413) for a one-box domain:
 * sweeps >= 1, A with the 7 coefficients over the valid box at loc, ap and
 * rhs over it, delta over it grown by one; delta zeroed. */
static int check_line_system(const char* op, int loc, const oa4* A, const oa4* ap, const oa4* rhs, int sweeps,
                             const ogeom* g, oa4* delta, obox* V)
{
    if (sweeps < 1) { return fail("gust: %s: sweeps must be >= 1 (got %d)", op, sweeps); }
    if (loc != OLOC_CELL && loc != OLOC_XFACE && loc != OLOC_YFACE && loc != OLOC_ZFACE) {
        return fail("gust: %s: unknown location %d", op, loc);
    }
    for (int d = 0; d < 3; ++d) {
        if (g->hi[d] < g->lo[d]) { return fail("gust: %s: empty domain", op); }
    }
    if (g->periodic[2]) { return fail("gust: %s: the domain cannot be periodic in z", op); }
    *V = loc_box(g, loc);
    if (!A || A->ncomp != OC_N) { return fail("gust: %s: A needs the implicit operator's components", op); }
    if (!view_covers(A, V, OC_N)) { return fail("gust: %s: A is not on the valid box at the %s", op, loc_word(loc)); }
    if (!view_covers(ap, V, 1)) { return fail("gust: %s: ap is not on the valid box at the %s", op, loc_word(loc)); }
    if (!view_covers(rhs, V, 1)) { return fail("gust: %s: rhs is not on the valid box at the %s", op, loc_word(loc)); }
    obox G = *V;
    for (int d = 0; d < 3; ++d) { G.lo[d] -= 1; G.hi[d] += 1; }
    if (!view_covers(delta, &G, 1)) {
        return fail("gust: %s: delta needs at least 1 ghost cells at the %s", op, loc_word(loc));
    }
    set_all(delta, 0.0);
    return 0;
}

/* The z-line Jacobi sweeps on `threads` host threads: within a sweep every
 * column reads the previous iterate only, so the columns split across the
 * threads (OpenMP) with the serial order's bits; the change is a maximum
 * (exact in any order). */
static int zline_jacobi_impl(int loc, const oa4* A, const oa4* ap, const oa4* rhs, int sweeps, const ogeom* g,
                             oa4* delta, double* change, int threads)
{
    obox V;
    if (check_line_system("zline_jacobi", loc, A, ap, rhs, sweeps, g, delta, &V)) { return -1; }
    if (threads < 1) { threads = 1; }
    oa4 c = fab_new(V, 1), e = fab_new(V, 1), next = fab_new(V, 1);
    double ch = 0;
    const int ilo = V.lo[0], jlo = V.lo[1], nxv = V.hi[0] - V.lo[0] + 1, nyv = V.hi[1] - V.lo[1] + 1;
    const long ncol = (long)nxv * nyv;
    for (int s = 0; s < sweeps; ++s) {
        fill_boundary(delta, &V, g);
#pragma omp parallel for num_threads(threads) schedule(static) if (threads > 1)
        for (long t = 0; t < ncol; ++t) {
            const int i = ilo + (int)(t % nxv), j = jlo + (int)(t / nxv);
            oracle_solve_column(A, ap, rhs, delta, &c, &e, &next, i, j, V.lo[2], V.hi[2]);
        }
        if (s == sweeps - 1) {
            double m = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(max : m) if (threads > 1)
            for (long t = 0; t < ncol; ++t) {
                const int i = ilo + (int)(t % nxv), j = jlo + (int)(t / nxv);
                for (int k = V.lo[2]; k <= V.hi[2]; ++k) {
                    const double dlt = fabs(AT(&next, i, j, k, 0) - AT(delta, i, j, k, 0));
                    if (dlt > m) { m = dlt; }
                }
            }
            ch = m;
        }
        copy_box(delta, &next, &V, 1);
    }
    override_sync(delta, loc, &V, g, 1);
    fab_free(&c);
    fab_free(&e);
    fab_free(&next);
    if (change) { *change = ch; }
    return 0;
}

int oracle_zline_jacobi(int loc, const oa4* A, const oa4* ap, const oa4* rhs, int sweeps, const ogeom* g,
                        oa4* delta, double* change)
{
    return zline_jacobi_impl(loc, A, ap, rhs, sweeps, g, delta, change, 1);
}

int oracle_zline_jacobi_threads(int loc, const oa4* A, const oa4* ap, const oa4* rhs, int sweeps, const ogeom* g,
                                oa4* delta, double* change, int threads)
{
    return zline_jacobi_impl(loc, A, ap, rhs, sweeps, g, delta, change, threads);
}

/* One pass over the columns of S, in place in d (Source/LineSweeps.cpp:38-155).
 * forward/jrev pick the pair; order picks sequential or diagonal planes. */
static void line_pass(const obox* S, const oa4* a, const oa4* dg, const oa4* r, oa4* d, oa4* c, oa4* e, int order,
                      int forward, int jrev)
{
    const int ilo = S->lo[0], ihi = S->hi[0], jlo = S->lo[1], jhi = S->hi[1], klo = S->lo[2], khi = S->hi[2];
    if (order == OORDER_SEQUENTIAL) {
        ++g_launches;
        if (!jrev && forward) {
            for (int j = jlo; j <= jhi; ++j)
                for (int i = ilo; i <= ihi; ++i) { oracle_solve_column(a, dg, r, d, c, e, d, i, j, klo, khi); }
        } else if (!jrev) {
            for (int j = jhi; j >= jlo; --j)
                for (int i = ihi; i >= ilo; --i) { oracle_solve_column(a, dg, r, d, c, e, d, i, j, klo, khi); }
        } else if (forward) {
            for (int j = jhi; j >= jlo; --j)
                for (int i = ilo; i <= ihi; ++i) { oracle_solve_column(a, dg, r, d, c, e, d, i, j, klo, khi); }
        } else {
            for (int j = jlo; j <= jhi; ++j)
                for (int i = ihi; i >= ilo; --i) { oracle_solve_column(a, dg, r, d, c, e, d, i, j, klo, khi); }
        }
        return;
    }
    const int nx = ihi - ilo + 1, ny = jhi - jlo + 1, last = nx + ny - 2;
    for (int t = 0; t <= last; ++t) {
        const int s = forward ? t : last - t;
        const int amin = s - (ny - 1) > 0 ? s - (ny - 1) : 0, amax = s < nx - 1 ? s : nx - 1;
        ++g_launches;
        for (int m = 0; m <= amax - amin; ++m) {
            const int ia = amin + m;
            const int i = ilo + ia, j = jrev ? jhi - (s - ia) : jlo + s - ia;
            oracle_solve_column(a, dg, r, d, c, e, d, i, j, klo, khi);
        }
    }
}

/* Source/LineSweeps.cpp:315-329. */
static int check_line_options(int ghosts, int overlap, int order, int alternate)
{
    if (overlap < 0 || overlap > O_MAX_LINE_OVERLAP) {
        return fail("gust: zline_lusgs: invalid options:\n  steady.line_overlap = %d: must be an integer >= 0 and <= %d",
                    overlap, O_MAX_LINE_OVERLAP);
    }
    if (overlap > 0 && ghosts != OGHOSTS_ZERO) {
        return fail("gust: zline_lusgs: invalid options:\n  steady.line_overlap = %d: needs steady.line_ghosts = zero "
                    "(the lagged ghosts sweep the valid box alone)", overlap);
    }
    if (alternate != 0 && alternate != 1) {
        return fail("gust: zline_lusgs: invalid options:\n  steady.line_alternate = %d: must be 0 or 1", alternate);
    }
    if (ghosts != OGHOSTS_LAGGED && ghosts != OGHOSTS_ZERO) {
        return fail("gust: zline_lusgs: invalid options:\n  steady.line_ghosts = %d: must be lagged or zero", ghosts);
    }
    if (order != OORDER_SEQUENTIAL && order != OORDER_DIAGONAL) {
        return fail("gust: zline_lusgs: invalid options:\n  steady.line_order = %d: must be sequential or diagonal", order);
    }
    return 0;
}

static double max_abs_diff(const oa4* a, const oa4* b, const obox* B)
{
    double m = 0;
    for (int k = B->lo[2]; k <= B->hi[2]; ++k)
        for (int j = B->lo[1]; j <= B->hi[1]; ++j)
            for (int i = B->lo[0]; i <= B->hi[0]; ++i) {
                const double v = fabs(AT(a, i, j, k, 0) - AT(b, i, j, k, 0));
                if (v > m) { m = v; }
            }
    return m;
}

/* Source/LineSweeps.cpp:181-210. */
static double lusgs_lagged(const oa4* A, const oa4* ap, const oa4* rhs, int sweeps, int order, int alternate,
                           const ogeom* g, const obox* V, oa4* delta)
{
    oa4 c = fab_new(*V, 1), e = fab_new(*V, 1), before = fab_new(*V, 1);
    double change = 0;
    for (int s = 0; s < sweeps; ++s) {
        for (int f = 1; f >= 0; --f) {
            if (s == sweeps - 1 && !f) { copy_box(&before, delta, V, 1); }
            fill_boundary(delta, V, g);
            line_pass(V, A, ap, rhs, delta, &c, &e, order, f, alternate && (s % 2 == 1));
        }
        if (s == sweeps - 1) { change = max_abs_diff(&before, delta, V); }
    }
    fab_free(&c);
    fab_free(&e);
    fab_free(&before);
    return change;
}

/* line_owned_box (Source/LineSweeps.cpp:361-374) for the one box. */
static obox owned_box(const obox* V, const ogeom* g, int loc)
{
    obox o = *V;
    for (int d = 0; d < 2; ++d) {
        if (!((loc >> d) & 1)) { continue; }
        if (V->hi[d] != g->hi[d] + 1) { o.hi[d] -= 1; }
        if (g->periodic[d] && V->lo[d] == g->lo[d]) { o.lo[d] += 1; }
    }
    return o;
}

/* line_overlap_width (Source/LineSweeps.cpp:337-341). */
static void overlap_width(int overlap, const ogeom* g, int ng[3])
{
    ng[0] = (g->hi[0] - g->lo[0] + 1) > 1 ? overlap : 0;
    ng[1] = (g->hi[1] - g->lo[1] + 1) > 1 ? overlap : 0;
    ng[2] = 0;
}

/* Source/LineSweeps.cpp:215-305. */
static double lusgs_zero(int loc, const oa4* A, const oa4* ap, const oa4* rhs, int sweeps, int order, int alternate,
                         int overlap, const ogeom* g, const obox* V, oa4* delta)
{
    int ng[3];
    overlap_width(overlap, g, ng);
    const int fold[2] = {g->periodic[0] && (g->hi[0] - g->lo[0] + 1) == 1,
                         g->periodic[1] && (g->hi[1] - g->lo[1] + 1) == 1};
    oa4 Ag = {0}, apg = {0}, rg = {0};
    const oa4 *As = A, *aps = ap, *rs = rhs;
    const int grown = ng[0] > 0 || ng[1] > 0 || fold[0] || fold[1];
    if (grown) {
        obox Gb = *V;
        for (int d = 0; d < 3; ++d) { Gb.lo[d] -= ng[d]; Gb.hi[d] += ng[d]; }
        Ag = fab_new(Gb, OC_N);
        apg = fab_new(Gb, 1);
        rg = fab_new(Gb, 1);
        copy_box(&Ag, A, V, OC_N);
        copy_box(&apg, ap, V, 1);
        copy_box(&rg, rhs, V, 1);
        fill_boundary(&Ag, V, g);
        fill_boundary(&apg, V, g);
        fill_boundary(&rg, V, g);
        for (int d = 0; d < 2; ++d) {
            if (!fold[d]) { continue; }
            const int cm = d == 0 ? OC_XM : OC_YM, cp = d == 0 ? OC_XP : OC_YP;
            for (int k = Gb.lo[2]; k <= Gb.hi[2]; ++k)
                for (int j = Gb.lo[1]; j <= Gb.hi[1]; ++j)
                    for (int i = Gb.lo[0]; i <= Gb.hi[0]; ++i) {
                        AT(&apg, i, j, k, 0) -= AT(&Ag, i, j, k, cm) + AT(&Ag, i, j, k, cp);
                        AT(&Ag, i, j, k, cm) = 0;
                        AT(&Ag, i, j, k, cp) = 0;
                    }
        }
        As = &Ag;
        aps = &apg;
        rs = &rg;
    }
    obox Wb = *V, Cb = *V;
    for (int d = 0; d < 3; ++d) {
        Wb.lo[d] -= ng[d] + 1;
        Wb.hi[d] += ng[d] + 1;
        Cb.lo[d] -= ng[d];
        Cb.hi[d] += ng[d];
    }
    oa4 work = fab_new(Wb, 1), c = fab_new(Cb, 1), e = fab_new(Cb, 1), before = fab_new(*V, 1);
    const obox own = owned_box(V, g, loc);
    /* line_sweep_box (Source/LineSweeps.cpp:376-389) */
    obox S = own;
    const obox dom = loc_box(g, loc);
    for (int d = 0; d < 2; ++d) {
        S.lo[d] -= ng[d];
        S.hi[d] += ng[d];
        if (!g->periodic[d]) {
            if (S.lo[d] < dom.lo[d]) { S.lo[d] = dom.lo[d]; }
            if (S.hi[d] > dom.hi[d]) { S.hi[d] = dom.hi[d]; }
        }
    }
    for (int s = 0; s < sweeps; ++s) {
        for (int f = 1; f >= 0; --f) {
            if (s == sweeps - 1 && !f) { copy_box(&before, &work, V, 1); }
            line_pass(&S, As, aps, rs, &work, &c, &e, order, f, alternate && (s % 2 == 1));
        }
    }
    double change = 0;
    for (int k = V->lo[2]; k <= V->hi[2]; ++k)
        for (int j = V->lo[1]; j <= V->hi[1]; ++j)
            for (int i = V->lo[0]; i <= V->hi[0]; ++i) {
                if (!box_contains(&own, i, j, k)) { continue; }
                const double v = fabs(AT(&before, i, j, k, 0) - AT(&work, i, j, k, 0));
                if (v > change) { change = v; }
            }
    copy_box(delta, &work, V, 1);
    if (loc == OLOC_XFACE || loc == OLOC_YFACE) { override_sync(delta, loc, V, g, 0); }
    fab_free(&work);
    fab_free(&c);
    fab_free(&e);
    fab_free(&before);
    if (grown) {
        fab_free(&Ag);
        fab_free(&apg);
        fab_free(&rg);
    }
    return change;
}

int oracle_zline_lusgs(int loc, const oa4* A, const oa4* ap, const oa4* rhs, int sweeps, const ogeom* g,
                       oa4* delta, int ghosts, int overlap, int order, int alternate, double* change)
{
    obox V;
    g_launches = 0;
    if (check_line_system("zline_lusgs", loc, A, ap, rhs, sweeps, g, delta, &V)) { return -1; }
    if (check_line_options(ghosts, overlap, order, alternate)) { return -1; }
    /* check_line_overlap_layout (Source/LineSweeps.cpp:343-359): one box, so
     * the narrowest box is the domain along each axis the overlap applies to. */
    int ng[3];
    overlap_width(overlap, g, ng);
    int narrowest = 0x7fffffff;
    for (int d = 0; d < 2; ++d) {
        const int len = g->hi[d] - g->lo[d] + 1;
        if (ng[d] > 0 && len < narrowest) { narrowest = len; }
    }
    if (overlap > 0 && narrowest != 0x7fffffff && overlap > narrowest) {
        return fail("gust: zline_lusgs: invalid options:\n  steady.line_overlap = %d: must be <= the narrowest box's "
                    "width in x and y (%d cells; mesh.max_grid_size)", overlap, narrowest);
    }
    double ch;
    if (ghosts == OGHOSTS_ZERO) {
        ch = lusgs_zero(loc, A, ap, rhs, sweeps, order, alternate, overlap, g, &V, delta);
    } else {
        ch = lusgs_lagged(A, ap, rhs, sweeps, order, alternate, g, &V, delta);
        override_sync(delta, loc, &V, g, 1);
    }
    if (change) { *change = ch; }
    return 0;
}
