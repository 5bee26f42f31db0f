/*
 * nunatak_oracle.c — CPU restatement of the reference velocity solve. TEST INFRASTRUCTURE ONLY.
 *
 * This file is the parity oracle of libnunatak_hip.so. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, and only as the checker (or the timed CPU baseline), never as a product path.
 *
 * It restates the Julia reference (Nunatak.jl on Oceananigans) kernel by kernel, in the reference's own structure:
 * fields are halo'd parent arrays, halos are filled after every write the way Oceananigans' fill_halo_regions!
 * does for the locations used here (Center on Bounded: copy of the adjacent interior value; Periodic: wrap), and
 * every expression keeps the reference's evaluation order (Julia evaluates a*b*c left to right and never contracts
 * to FMA), so compile this file with -ffp-contract=off. Each function cites the reference file:line it follows.
 *
 * Parity status of each part (what pins it) is listed in tests/test_oracle_*.py and in the repository's design
 * notes. Known unpinned details, all at the last-ulp level: the nesting order of the two-dimensional
 * interpolations ℑxy (taken as ℑy∘ℑx), Julia's pow/Float64^Int versus libm pow, and the order of Julia's pairwise
 * `sum` in the dot products.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/nunatak_hip.h"

#define EXPORT __attribute__((visibility("default")))

/* OpenMP (the timed CPU baseline): loops over grid rows run in parallel when a level is large enough to pay for
 * it. Every parallel loop writes disjoint points with the same per-point arithmetic, and the pairwise sums split at
 * the same indices as the serial recursion, so results are bit for bit those of one thread. */
#ifdef _OPENMP
#include <omp.h>
#endif
#define OMP_BIG(g) ((long)(g)->Nx * (long)(g)->Ny >= 16384)

enum { C = 0, F = 1 };

typedef struct {
    int lx, ly, sigma;
} nko_loc;

static const nko_loc LOCS[NK_FIELD_COUNT] = {
    [NK_FIELD_U] = {F, C, 0},      [NK_FIELD_V] = {C, F, 0},       [NK_FIELD_H] = {C, C, 0},
    [NK_FIELD_BED] = {C, C, 0},    [NK_FIELD_S] = {C, C, 0},       [NK_FIELD_PHI] = {C, C, 0},
    [NK_FIELD_PHIU] = {F, C, 0},   [NK_FIELD_PHIV] = {C, F, 0},    [NK_FIELD_N] = {C, C, 0},
    [NK_FIELD_MU] = {C, C, 0},     [NK_FIELD_TAUU] = {F, C, 0},    [NK_FIELD_TAUV] = {C, F, 0},
    [NK_FIELD_ETAH] = {C, C, 0},   [NK_FIELD_ETAHFF] = {F, F, 0},  [NK_FIELD_BETAU] = {F, C, 0},
    [NK_FIELD_BETAV] = {C, F, 0},  [NK_FIELD_F2] = {C, C, 0},      [NK_FIELD_ETASIGMA] = {C, C, 1},
    [NK_FIELD_BETABU] = {F, C, 0}, [NK_FIELD_BETABV] = {C, F, 0},  [NK_FIELD_UB] = {F, C, 0},
    [NK_FIELD_VB] = {C, F, 0},     [NK_FIELD_USIGMA] = {F, C, 1},  [NK_FIELD_VSIGMA] = {C, F, 1},
    [NK_FIELD_A] = {C, C, 0},      [NK_FIELD_C] = {C, C, 0},       [NK_FIELD_TEMPERATURE] = {C, C, 1},
    [NK_FIELD_PU] = {F, C, 0},     [NK_FIELD_PV] = {C, F, 0},      [NK_FIELD_QU] = {F, C, 0},
    [NK_FIELD_QV] = {C, F, 0},     [NK_FIELD_UBAR] = {F, C, 0},    [NK_FIELD_VBAR] = {C, F, 0},
    [NK_FIELD_LAMBDAU] = {F, C, 0},[NK_FIELD_LAMBDAV] = {C, F, 0}, [NK_FIELD_CBAR] = {C, C, 0},
    [NK_FIELD_ABAR] = {C, C, 0},   [NK_FIELD_USIGMABAR] = {F, C, 1}, [NK_FIELD_VSIGMABAR] = {C, F, 1},
    [NK_FIELD_ENTHALPY] = {C, C, 1}, [NK_FIELD_WATER_FRACTION] = {C, C, 1},
    [NK_FIELD_SMB] = {C, C, 0},    [NK_FIELD_BASAL_MELT] = {C, C, 0}, [NK_FIELD_GROUNDED_MELT] = {C, C, 0},
    [NK_FIELD_HPREV] = {C, C, 0},  [NK_FIELD_GH] = {C, C, 0},       [NK_FIELD_HCLIP] = {C, C, 0},
    [NK_FIELD_HW] = {C, C, 0},     [NK_FIELD_HN] = {C, C, 0},       [NK_FIELD_GEOTHERMAL] = {C, C, 0},
    [NK_FIELD_TSURF] = {C, C, 0},  [NK_FIELD_TOCEAN] = {C, C, 0},   [NK_FIELD_STRAIN_HEATING] = {C, C, 1},
    [NK_FIELD_HBAR] = {C, C, 0},   [NK_FIELD_BEDBAR] = {C, C, 0},
};

/* ------------------------------------------------------------------------------------------------------------ */
/* Grid and fields                                                                                               */
/* ------------------------------------------------------------------------------------------------------------ */

typedef struct {
    int Nx, Ny, tx, ty, H, PX, PY;
    double dx, dy;
    size_t plane;
} og;  /* one grid (a multigrid level has its own) */

#define IDX(g, i, j) ((size_t)((j) - 1 + (g)->H) * (size_t)(g)->PX + (size_t)((i) - 1 + (g)->H))
#define AT(g, f, i, j) ((f)[IDX(g, i, j)])

static void og_init(og *g, int Nx, int Ny, int tx, int ty, double dx, double dy) {
    g->Nx = Nx; g->Ny = Ny; g->tx = tx; g->ty = ty; g->dx = dx; g->dy = dy;
    g->H = 3;
    g->PX = Nx + 1 + 2 * g->H;
    g->PY = Ny + 1 + 2 * g->H;
    g->plane = (size_t)g->PX * (size_t)g->PY;
}

static double *og_alloc(const og *g, int nz) { return calloc(g->plane * (size_t)(nz < 1 ? 1 : nz), sizeof(double)); }

/* Interior extent of a location: Face on a Bounded direction has N + 1 points (Oceananigans `size`). */
static int nx_of(const og *g, int lx) { return (lx == F && g->tx) ? g->Nx + 1 : g->Nx; }
static int ny_of(const og *g, int ly) { return (ly == F && g->ty) ? g->Ny + 1 : g->Ny; }

static int wrap(int i, int N) { int r = (i - 1) % N; if (r < 0) r += N; return r + 1; }
static int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

/* fill_halo_regions! for one field: Periodic wraps; Bounded Center copies the adjacent interior value
 * (zero gradient, relied on at src/StressBalances/depth_integrated.jl:41-43); a Face field on its own Bounded
 * direction copies the boundary value (those halo points are peripheral and never enter a result). */
static void fill_halo(const og *g, double *f, int lx, int ly, int nz) {
    int nx = nx_of(g, lx), ny = ny_of(g, ly);
    for (int k = 0; k < (nz < 1 ? 1 : nz); k++) {
        double *p = f + (size_t)k * g->plane;
        for (int j = 1; j <= ny; j++) {
            for (int i = 1 - g->H; i <= g->Nx + 1 + g->H; i++) {
                if (i >= 1 && i <= nx) continue;
                int src = g->tx ? clampi(i, 1, nx) : wrap(i, g->Nx);
                AT(g, p, i, j) = AT(g, p, src, j);
            }
        }
        for (int j = 1 - g->H; j <= g->Ny + 1 + g->H; j++) {
            if (j >= 1 && j <= ny) continue;
            int src = g->ty ? clampi(j, 1, ny) : wrap(j, g->Ny);
            for (int i = 1 - g->H; i <= g->Nx + 1 + g->H; i++) AT(g, p, i, j) = AT(g, p, i, src);
        }
    }
}

/* peripheral_node(i, j, k, grid, LX, LY, Center()): a Face on a Bounded direction is peripheral on the boundary
 * and beyond, a Center outside the domain (Oceananigans Grids/inactive_node.jl); never on Periodic. */
static int periph(const og *g, int i, int j, int lx, int ly) {
    int px = g->tx && (lx == F ? (i <= 1 || i >= g->Nx + 1) : (i < 1 || i > g->Nx));
    int py = g->ty && (ly == F ? (j <= 1 || j >= g->Ny + 1) : (j < 1 || j > g->Ny));
    return px || py;
}

/* Julia Base.max / min / clamp for Float64 (signed zeros as Julia treats them). */
static double jmax(double x, double y) { double d = x - y; return (isnan(x) || isnan(y)) ? d : (signbit(d) ? y : x); }
static double jmin(double x, double y) { double d = x - y; return (isnan(x) || isnan(y)) ? d : (signbit(d) ? x : y); }
static double jclamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }

/* ------------------------------------------------------------------------------------------------------------ */
/* SSA operator (src/StressBalances/ssa_operator.jl:47-89, strain_rates.jl)                                                      */
/* ------------------------------------------------------------------------------------------------------------ */

/* ice_freeᶠᶜᶜ / ice_freeᶜᶠᶜ (src/Models/smooth_ice_mask.jl:72-75); μ or η̄H may be absent (NULL). */
static int ice_free_u(const og *g, int i, int j, const double *m) { return m && (AT(g, m, i - 1, j) + AT(g, m, i, j) == 0.0); }
static int ice_free_v(const og *g, int i, int j, const double *m) { return m && (AT(g, m, i, j - 1) + AT(g, m, i, j) == 0.0); }

/* inactiveᶠᶜᶜ / inactiveᶜᶠᶜ (strain_rates.jl:2-3) */
static int inactive_u(const og *g, int i, int j, const double *m) { return periph(g, i, j, F, C) | ice_free_u(g, i, j, m); }
static int inactive_v(const og *g, int i, int j, const double *m) { return periph(g, i, j, C, F) | ice_free_v(g, i, j, m); }

/* interior_u / interior_v (strain_rates.jl:5-6) */
static double interior_u(const og *g, int i, int j, const double *u, const double *m) { return inactive_u(g, i, j, m) ? 0.0 : AT(g, u, i, j); }
static double interior_v(const og *g, int i, int j, const double *v, const double *m) { return inactive_v(g, i, j, m) ? 0.0 : AT(g, v, i, j); }

/* strain_rate_xx / yy (strain_rates.jl:8-9): ∂xᶜᶜᶜ = δxᶜ / Δxᶜᶜᶜ */
static double strain_xx(const og *g, int i, int j, const double *u, const double *m) {
    return (interior_u(g, i + 1, j, u, m) - interior_u(g, i, j, u, m)) / g->dx;
}
static double strain_yy(const og *g, int i, int j, const double *v, const double *m) {
    return (interior_v(g, i, j + 1, v, m) - interior_v(g, i, j, v, m)) / g->dy;
}
/* wall_shear_factor (strain_rates.jl:11-12) */
static double wall_factor(int walls) { return walls == NK_NO_SLIP ? 2.0 : 0.0; }
/* strain_rate_xy (strain_rates.jl:14-17) */
static double strain_xy(const og *g, int i, int j, const double *u, const double *v, int walls, const double *m) {
    double e = ((interior_u(g, i, j, u, m) - interior_u(g, i, j - 1, u, m)) / g->dy +
                (interior_v(g, i, j, v, m) - interior_v(g, i - 1, j, v, m)) / g->dx) / 2.0;
    return periph(g, i, j, F, F) ? wall_factor(walls) * e : e;
}

/* normal_stress_xx / yy and shear_stress (ssa_operator.jl:60-62) */
static double nsxx(const og *g, int i, int j, const double *u, const double *v, const double *eta) {
    return 2.0 * AT(g, eta, i, j) * (2.0 * strain_xx(g, i, j, u, eta) + strain_yy(g, i, j, v, eta));
}
static double nsyy(const og *g, int i, int j, const double *u, const double *v, const double *eta) {
    return 2.0 * AT(g, eta, i, j) * (2.0 * strain_yy(g, i, j, v, eta) + strain_xx(g, i, j, u, eta));
}
static double shear(const og *g, int i, int j, const double *u, const double *v, const double *etaff, int walls, const double *eta) {
    return 2.0 * AT(g, etaff, i, j) * strain_xy(g, i, j, u, v, walls, eta);
}

/* x_stress_balance / y_stress_balance (ssa_operator.jl:69-79) */
static double x_stress_balance(const og *g, int i, int j, const double *u, const double *v, const double *eta,
                               const double *etaff, const double *bu, int walls) {
    double div = (g->dy * nsxx(g, i, j, u, v, eta) - g->dy * nsxx(g, i - 1, j, u, v, eta)) +
                 (g->dx * shear(g, i, j + 1, u, v, etaff, walls, eta) - g->dx * shear(g, i, j, u, v, etaff, walls, eta));
    double Az = g->dx * g->dy;
    return Az * AT(g, bu, i, j) * AT(g, u, i, j) - div;
}
static double y_stress_balance(const og *g, int i, int j, const double *u, const double *v, const double *eta,
                               const double *etaff, const double *bv, int walls) {
    double div = (g->dx * nsyy(g, i, j, u, v, eta) - g->dx * nsyy(g, i, j - 1, u, v, eta)) +
                 (g->dy * shear(g, i + 1, j, u, v, etaff, walls, eta) - g->dy * shear(g, i, j, u, v, etaff, walls, eta));
    double Az = g->dx * g->dy;
    return Az * AT(g, bv, i, j) * AT(g, v, i, j) - div;
}

/* _ssa_operator! over :xyz (ssa_operator.jl:81-89); the halos of p, η̄H, η̄Hᶠᶠ must be filled. */
static void ssa_operator(const og *g, double *qu, double *qv, const double *pu, const double *pv, const double *eta,
                         const double *etaff, const double *bu, const double *bv, int walls) {
    #pragma omp parallel for schedule(static) if (OMP_BIG(g))
    for (int j = 1; j <= g->Ny; j++)
        for (int i = 1; i <= g->Nx; i++) {
            int wu = inactive_u(g, i, j, eta), wv = inactive_v(g, i, j, eta);
            AT(g, qu, i, j) = wu ? AT(g, pu, i, j) : x_stress_balance(g, i, j, pu, pv, eta, etaff, bu, walls);
            AT(g, qv, i, j) = wv ? AT(g, pv, i, j) : y_stress_balance(g, i, j, pu, pv, eta, etaff, bv, walls);
        }
}

/* effective_strain_rate² (strain_rates.jl:27-32): strain rates without η̄H (walls only); ℑxyᶜᶜ of ε̇xy². */
static double eff_strain2(const og *g, int i, int j, const double *u, const double *v, int walls) {
    double exx = strain_xx(g, i, j, u, NULL), eyy = strain_yy(g, i, j, v, NULL);
    double a = strain_xy(g, i, j, u, v, walls, NULL), b = strain_xy(g, i + 1, j, u, v, walls, NULL);
    double c = strain_xy(g, i, j + 1, u, v, walls, NULL), d = strain_xy(g, i + 1, j + 1, u, v, walls, NULL);
    double exy2 = ((a * a + b * b) / 2.0 + (c * c + d * d) / 2.0) / 2.0;
    return exx * exx + eyy * eyy + exx * eyy + exy2;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Diagonal preconditioner (src/StressBalances/diagonal_preconditioner.jl:20-54)                                  */
/* ------------------------------------------------------------------------------------------------------------ */

static double x_nd(const og *g, int i, int j, const double *eta) { return 4.0 * g->dy * AT(g, eta, i, j) / g->dx; }
static double y_nd(const og *g, int i, int j, const double *eta) { return 4.0 * g->dx * AT(g, eta, i, j) / g->dy; }
static double corner_factor(const og *g, int i, int j, int walls) { return periph(g, i, j, F, F) ? wall_factor(walls) : 1.0; }
static double x_sd(const og *g, int i, int j, const double *etaff, int walls) {
    return corner_factor(g, i, j, walls) * g->dx * AT(g, etaff, i, j) / g->dy;
}
static double y_sd(const og *g, int i, int j, const double *etaff, int walls) {
    return corner_factor(g, i, j, walls) * g->dy * AT(g, etaff, i, j) / g->dx;
}
static double x_diagonal(const og *g, int i, int j, const double *eta, const double *etaff, const double *bu, int walls) {
    double d = g->dx * g->dy * AT(g, bu, i, j) + x_nd(g, i, j, eta) + x_nd(g, i - 1, j, eta) +
               x_sd(g, i, j, etaff, walls) + x_sd(g, i, j + 1, etaff, walls);
    return (inactive_u(g, i, j, eta) | (d == 0.0)) ? 1.0 : d;
}
static double y_diagonal(const og *g, int i, int j, const double *eta, const double *etaff, const double *bv, int walls) {
    double d = g->dx * g->dy * AT(g, bv, i, j) + y_nd(g, i, j, eta) + y_nd(g, i, j - 1, eta) +
               y_sd(g, i, j, etaff, walls) + y_sd(g, i + 1, j, etaff, walls);
    return (inactive_v(g, i, j, eta) | (d == 0.0)) ? 1.0 : d;
}
static void diagonal_precondition(const og *g, double *zu, double *zv, const double *ru, const double *rv,
                                  const double *eta, const double *etaff, const double *bu, const double *bv, int walls) {
    #pragma omp parallel for schedule(static) if (OMP_BIG(g))
    for (int j = 1; j <= ny_of(g, C); j++)
        for (int i = 1; i <= nx_of(g, F); i++) AT(g, zu, i, j) = AT(g, ru, i, j) / x_diagonal(g, i, j, eta, etaff, bu, walls);
    #pragma omp parallel for schedule(static) if (OMP_BIG(g))
    for (int j = 1; j <= ny_of(g, F); j++)
        for (int i = 1; i <= nx_of(g, C); i++) AT(g, zv, i, j) = AT(g, rv, i, j) / y_diagonal(g, i, j, eta, etaff, bv, walls);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Vector operations (src/Solvers/velocity_vector.jl:48-141)                                                      */
/* ------------------------------------------------------------------------------------------------------------ */

/* _linear_combination! over :xyz: z = a x + b y (velocity_vector.jl:54-62). */
static void lincomb(const og *g, double *z, double a, const double *x, double b, const double *y) {
    #pragma omp parallel for schedule(static) if (OMP_BIG(g))
    for (int j = 1; j <= g->Ny; j++)
        for (int i = 1; i <= g->Nx; i++) AT(g, z, i, j) = a * AT(g, x, i, j) + b * AT(g, y, i, j);
}
static void vlincomb(const og *g, double *zu, double *zv, double a, const double *xu, const double *xv, double b,
                     const double *yu, const double *yv) {
    lincomb(g, zu, a, xu, b, yu);
    lincomb(g, zv, a, xv, b, yv);
}

/* Julia's pairwise `sum` over a vector (Base.mapreduce_impl with pairwise block 1024). */
static double pairwise_sum(const double *w, long lo, long hi) {
    if (lo == hi) return w[lo];
    if (hi - lo < 1024) {
        double v = w[lo] + w[lo + 1];
        for (long i = lo + 2; i <= hi; i++) v += w[i];
        return v;
    }
    long mid = lo + ((hi - lo) >> 1);
    return pairwise_sum(w, lo, mid) + pairwise_sum(w, mid + 1, hi);
}
/* The same recursion with the halves of large ranges summed as parallel tasks (same split points, same result). */
static double pairwise_sum_task(const double *w, long lo, long hi) {
    if (hi - lo < (1L << 16)) return pairwise_sum(w, lo, hi);
    long mid = lo + ((hi - lo) >> 1);
    double a = 0.0, b = 0.0;
#pragma omp task shared(a)
    a = pairwise_sum_task(w, lo, mid);
    b = pairwise_sum_task(w, mid + 1, hi);
#pragma omp taskwait
    return a + b;
}
static double pairwise_sum_par(const double *w, long lo, long hi) {
    if (hi - lo < (1L << 16)) return pairwise_sum(w, lo, hi);
    double r = 0.0;
#pragma omp parallel
#pragma omp single
    r = pairwise_sum_task(w, lo, hi);
    return r;
}

/* component_dot! (velocity_vector.jl:129-133): product over :xyz into a zeroed parent workspace, then the sum of
 * the whole parent array. */
static double component_dot(const og *g, double *w, const double *x, const double *y) {
    memset(w, 0, g->plane * sizeof(double));
    #pragma omp parallel for schedule(static) if (OMP_BIG(g))
    for (int j = 1; j <= g->Ny; j++)
        for (int i = 1; i <= g->Nx; i++) AT(g, w, i, j) = AT(g, x, i, j) * AT(g, y, i, j);
    return pairwise_sum_par(w, 0, (long)g->plane - 1);
}
static double vdot(const og *g, double *w, const double *xu, const double *xv, const double *yu, const double *yv) {
    return component_dot(g, w, xu, yu) + component_dot(g, w, xv, yv);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Multigrid (src/StressBalances/multigrid_preconditioner.jl)                                                    */
/* ------------------------------------------------------------------------------------------------------------ */

typedef struct {
    og g;
    int cx, cy;  /* 1 = Halved, 0 = Unchanged, relative to the next finer level */
    double *eta, *etaff, *bu, *bv, *xu, *xv, *b_u, *b_v, *qu, *qv, *t;
} nko_level;

typedef struct {
    int nlev;
    nko_level *lev;
    double *ru, *rv, *t;  /* finest residual workspace and line scratch */
} nko_mg;

/* coarsening (multigrid_preconditioner.jl:86) */
static int coarsening(int N, int H, int minimum) { int m = minimum > H ? minimum : H; return (N % 2 == 0) && (N / 2 >= m); }

/* ℓ1 row sums (multigrid_preconditioner.jl:195-211) */
static double x_abs_row_sum(const og *g, int i, int j, const double *eta, const double *etaff, const double *bu, int walls) {
    double s = g->dx * g->dy * AT(g, bu, i, j) +
               2.0 * (x_nd(g, i, j, eta) + x_nd(g, i - 1, j, eta)) +
               4.0 * (AT(g, eta, i, j) + AT(g, eta, i - 1, j)) +
               2.0 * (x_sd(g, i, j, etaff, walls) + x_sd(g, i, j + 1, etaff, walls)) +
               2.0 * (corner_factor(g, i, j, walls) * AT(g, etaff, i, j) + corner_factor(g, i, j + 1, walls) * AT(g, etaff, i, j + 1));
    return (inactive_u(g, i, j, eta) | (s == 0.0)) ? 1.0 : s;
}
static double y_abs_row_sum(const og *g, int i, int j, const double *eta, const double *etaff, const double *bv, int walls) {
    double s = g->dx * g->dy * AT(g, bv, i, j) +
               2.0 * (y_nd(g, i, j, eta) + y_nd(g, i, j - 1, eta)) +
               4.0 * (AT(g, eta, i, j) + AT(g, eta, i, j - 1)) +
               2.0 * (y_sd(g, i, j, etaff, walls) + y_sd(g, i + 1, j, etaff, walls)) +
               2.0 * (corner_factor(g, i, j, walls) * AT(g, etaff, i, j) + corner_factor(g, i + 1, j, walls) * AT(g, etaff, i + 1, j));
    return (inactive_v(g, i, j, eta) | (s == 0.0)) ? 1.0 : s;
}

/* x_line_coupling / y_line_coupling (multigrid_preconditioner.jl:223-227) */
static double x_line_coupling(const og *g, int i, int j, const double *eta, const double *etaff, int walls) {
    return (inactive_u(g, i, j - 1, eta) | inactive_u(g, i, j, eta)) ? 0.0 : x_sd(g, i, j, etaff, walls);
}
static double y_line_coupling(const og *g, int i, int j, const double *eta, int walls) {
    (void)walls;
    return (inactive_v(g, i, j - 1, eta) | inactive_v(g, i, j, eta)) ? 0.0 : y_nd(g, i, j - 1, eta);
}

/* line_relax! (multigrid_preconditioner.jl:231-253): Thomas algorithm along j. comp = 0 for u, 1 for v. */
static void line_relax(const og *g, double *x, const double *b, double *q, double *t, int i, int comp,
                       const double *eta, const double *etaff, const double *beta, int walls) {
    int Ny = g->Ny;
    double c = 0.0, tj = 0.0, d = 0.0;
    for (int j = 1; j <= Ny; j++) {
        double s = comp == 0 ? x_abs_row_sum(g, i, j, eta, etaff, beta, walls) : y_abs_row_sum(g, i, j, eta, etaff, beta, walls);
        double cn = comp == 0 ? x_line_coupling(g, i, j + 1, eta, etaff, walls) : y_line_coupling(g, i, j + 1, eta, walls);
        cn = (j == Ny) ? 0.0 : cn;
        double m = s - c - cn + c * tj;
        d = (AT(g, b, i, j) - AT(g, q, i, j) + c * d) / m;
        tj = -cn / m;
        AT(g, t, i, j) = tj;
        AT(g, q, i, j) = d;
        c = cn;
    }
    double delta = 0.0;
    for (int j = Ny; j >= 1; j--) {
        delta = AT(g, q, i, j) - AT(g, t, i, j) * delta;
        AT(g, x, i, j) += delta;
    }
}

typedef struct {
    const double *eta, *etaff, *bu, *bv;
} coeffs;

/* smooth! (multigrid_preconditioner.jl:165-180) */
static void smooth(const og *g, double *xu, double *xv, const double *bu_, const double *bv_, double *qu, double *qv,
                   double *t, coeffs c, int walls, int smoother, int sweeps) {
    for (int s = 0; s < sweeps; s++) {
        fill_halo(g, xu, F, C, 1);
        fill_halo(g, xv, C, F, 1);
        ssa_operator(g, qu, qv, xu, xv, c.eta, c.etaff, c.bu, c.bv, walls);
        if (smoother == NK_POINT_JACOBI) {
            /* _smooth! (multigrid_preconditioner.jl:213-219) */
            #pragma omp parallel for schedule(static) if (OMP_BIG(g))
            for (int j = 1; j <= g->Ny; j++)
                for (int i = 1; i <= g->Nx; i++) {
                    AT(g, xu, i, j) += (AT(g, bu_, i, j) - AT(g, qu, i, j)) / x_abs_row_sum(g, i, j, c.eta, c.etaff, c.bu, walls);
                    AT(g, xv, i, j) += (AT(g, bv_, i, j) - AT(g, qv, i, j)) / y_abs_row_sum(g, i, j, c.eta, c.etaff, c.bv, walls);
                }
        } else {
            /* _line_smooth! over :xz (multigrid_preconditioner.jl:255-259) */
            #pragma omp parallel for schedule(static) if (OMP_BIG(g))
            for (int i = 1; i <= g->Nx; i++) {
                line_relax(g, xu, bu_, qu, t, i, 0, c.eta, c.etaff, c.bu, walls);
                line_relax(g, xv, bv_, qv, t, i, 1, c.eta, c.etaff, c.bv, walls);
            }
        }
    }
}

/* restriction_index / restriction_weight (multigrid_preconditioner.jl:264-283). lf: F or C; h: halved. */
static int r_offsets_lo(int lf, int h) { return h ? -1 : 0; }
static int r_offsets_hi(int lf, int h) { return h ? (lf == F ? 1 : 2) : 0; }
static int r_index(int I, int a, int N, int bounded, int lf, int h) {
    if (!h) return I;
    if (!bounded) return wrap(2 * I - 1 + a, N);
    if (lf == F) { int i = 2 * I - 1 + a; return i > 1 ? i : 1; }
    return clampi(2 * I - 1 + a, 1, N);
}
static double r_weight(int I, int a, int Nc, int bounded, int lf, int h) {
    if (!h) return 1.0;
    if (lf == F) {
        double w = (a == 0) ? 1.0 : 1.0 / 2.0;
        return (bounded && (2 * I - 1 + a < 1)) ? 0.0 : w;
    }
    double w = ((a == 0) | (a == 1)) ? 3.0 * 1.0 / 4.0 : 1.0 / 4.0;
    if (!bounded) return w;
    int outside = ((I == 1) & (a == -1)) | ((I == Nc) & (a == 2));
    int folded = ((I == 1) & (a == 0)) | ((I == Nc) & (a == 1));
    w = w + (folded ? 1.0 / 4.0 : 0.0);
    return outside ? 0.0 : w;
}

/* parents (multigrid_preconditioner.jl:286-305) */
static void parents(int i, int Nc, int bounded, int lf, int h, int *I1, int *I2, double *w1, double *w2) {
    if (!h) { *I1 = i; *I2 = i; *w1 = 1.0; *w2 = 0.0; return; }
    int I = (i + 1) / 2, odd = i & 1;
    if (lf == F) {
        int up = bounded ? I + 1 : wrap(I + 1, Nc);
        *I1 = I; *I2 = odd ? I : up; *w1 = odd ? 1.0 : 1.0 / 2.0; *w2 = odd ? 0.0 : 1.0 / 2.0;
    } else {
        int Ip = odd ? I - 1 : I + 1;
        *I1 = I; *I2 = bounded ? clampi(Ip, 1, Nc) : wrap(Ip, Nc); *w1 = 3.0 * 1.0 / 4.0; *w2 = 1.0 / 4.0;
    }
}

/* Pointwise function evaluated by restrict/prolongate: either interior_u/v of a vector or Az·ϕ. */
typedef struct {
    int kind;  /* 0 interior_u, 1 interior_v, 2 area weighted */
    const double *f, *eta;
} phi_fn;
static double phi_eval(const og *g, int i, int j, phi_fn p) {
    if (p.kind == 0) return interior_u(g, i, j, p.f, p.eta);
    if (p.kind == 1) return interior_v(g, i, j, p.f, p.eta);
    return g->dx * g->dy * AT(g, p.f, i, j);
}

/* restrict (multigrid_preconditioner.jl:307-318) */
static double restrict_at(const og *cg, const og *fg, int I, int J, int lx, int ly, int cx, int cy, phi_fn p) {
    double s = 0.0;
    for (int a = r_offsets_lo(lx, cx); a <= r_offsets_hi(lx, cx); a++)
        for (int b = r_offsets_lo(ly, cy); b <= r_offsets_hi(ly, cy); b++) {
            int i = r_index(I, a, fg->Nx, fg->tx, lx, cx);
            int j = r_index(J, b, fg->Ny, fg->ty, ly, cy);
            double w = r_weight(I, a, cg->Nx, fg->tx, lx, cx) * r_weight(J, b, cg->Ny, fg->ty, ly, cy);
            s += w * phi_eval(fg, i, j, p);
        }
    return s;
}

/* prolongate (multigrid_preconditioner.jl:320-327) */
static double prolongate_at(const og *fg, const og *cg, int i, int j, int lx, int ly, int cx, int cy, phi_fn p) {
    int I1, I2, J1, J2;
    double wx1, wx2, wy1, wy2;
    parents(i, cg->Nx, fg->tx, lx, cx, &I1, &I2, &wx1, &wx2);
    parents(j, cg->Ny, fg->ty, ly, cy, &J1, &J2, &wy1, &wy2);
    return wx1 * (wy1 * phi_eval(cg, I1, J1, p) + wy2 * phi_eval(cg, I1, J2, p)) +
           wx2 * (wy1 * phi_eval(cg, I2, J1, p) + wy2 * phi_eval(cg, I2, J2, p));
}

/* _interpolate_to_corners! (src/StressBalances/stress_balance_coefficients.jl:26-29): ℑxyᶠᶠ = ℑy∘ℑx. */
static void interpolate_to_corners(const og *g, double *etaff, const double *eta) {
    #pragma omp parallel for schedule(static) if (OMP_BIG(g))
    for (int j = 1; j <= ny_of(g, F); j++)
        for (int i = 1; i <= nx_of(g, F); i++)
            AT(g, etaff, i, j) = ((AT(g, eta, i - 1, j - 1) + AT(g, eta, i, j - 1)) / 2.0 +
                                  (AT(g, eta, i - 1, j) + AT(g, eta, i, j)) / 2.0) / 2.0;
    fill_halo(g, etaff, F, F, 1);
}

/* _coarsen_coefficients! (multigrid_preconditioner.jl:331-344) and coarsen_coefficients! (:132-141) */
static void coarsen_coefficients(nko_mg *mg, const og *fg, const double *eta, const double *bu, const double *bv) {
    const og *f = fg;
    const double *fe = eta, *fbu = bu, *fbv = bv;
    for (int l = 0; l < mg->nlev; l++) {
        nko_level *L = &mg->lev[l];
        const og *g = &L->g;
        int fx = L->cx ? 2 : 1, fy = L->cy ? 2 : 1;
        #pragma omp parallel for schedule(static) if (OMP_BIG(g))
        for (int J = 1; J <= g->Ny; J++)
            for (int I = 1; I <= g->Nx; I++) {
                int i = fx * (I - 1) + 1, j = fy * (J - 1) + 1;
                double S = 0.0;
                for (int a = 0; a <= fx - 1; a++)
                    for (int b = 0; b <= fy - 1; b++) S += f->dx * f->dy * AT(f, fe, i + a, j + b);
                AT(g, L->eta, I, J) = S / (g->dx * g->dy);
                phi_fn pu = {2, fbu, NULL}, pv = {2, fbv, NULL};
                AT(g, L->bu, I, J) = restrict_at(g, f, I, J, F, C, L->cx, L->cy, pu) / (g->dx * g->dy);
                AT(g, L->bv, I, J) = restrict_at(g, f, I, J, C, F, L->cx, L->cy, pv) / (g->dx * g->dy);
            }
        fill_halo(g, L->eta, C, C, 1);
        interpolate_to_corners(g, L->etaff, L->eta);
        fill_halo(g, L->bu, F, C, 1);
        fill_halo(g, L->bv, C, F, 1);
        f = g; fe = L->eta; fbu = L->bu; fbv = L->bv;
    }
}

/* _restrict! (multigrid_preconditioner.jl:346-352) */
static void restrict_vec(const og *cg, const og *fg, int cx, int cy, double *bu, double *bv, const double *ru,
                         const double *rv, const double *ceta, const double *feta) {
#pragma omp parallel for schedule(static) if (OMP_BIG(cg))
    for (int J = 1; J <= cg->Ny; J++)
        for (int I = 1; I <= cg->Nx; I++) {
            phi_fn pu = {0, ru, feta}, pv = {1, rv, feta};
            AT(cg, bu, I, J) = inactive_u(cg, I, J, ceta) ? 0.0 : restrict_at(cg, fg, I, J, F, C, cx, cy, pu);
            AT(cg, bv, I, J) = inactive_v(cg, I, J, ceta) ? 0.0 : restrict_at(cg, fg, I, J, C, F, cx, cy, pv);
        }
}

/* _prolongate! (multigrid_preconditioner.jl:354-360) */
static void prolongate_vec(const og *fg, const og *cg, int cx, int cy, double *xu, double *xv, const double *cu,
                           const double *cv, const double *feta, const double *ceta) {
#pragma omp parallel for schedule(static) if (OMP_BIG(fg))
    for (int j = 1; j <= fg->Ny; j++)
        for (int i = 1; i <= fg->Nx; i++) {
            phi_fn pu = {0, cu, ceta}, pv = {1, cv, ceta};
            AT(fg, xu, i, j) += inactive_u(fg, i, j, feta) ? 0.0 : prolongate_at(fg, cg, i, j, F, C, cx, cy, pu);
            AT(fg, xv, i, j) += inactive_v(fg, i, j, feta) ? 0.0 : prolongate_at(fg, cg, i, j, C, F, cx, cy, pv);
        }
}

/* fill!(x::VelocityVector, value) over the interior of each component (velocity_vector.jl:71-75) */
static void vfill(const og *g, double *xu, double *xv, double value) {
    for (int j = 1; j <= ny_of(g, C); j++) for (int i = 1; i <= nx_of(g, F); i++) AT(g, xu, i, j) = value;
    for (int j = 1; j <= ny_of(g, F); j++) for (int i = 1; i <= nx_of(g, C); i++) AT(g, xv, i, j) = value;
}

typedef struct {
    int smoother, sweeps, coarsest_sweeps, walls;
} mg_params;

/* v_cycle! (multigrid_preconditioner.jl:144-163); level = index of the first coarse level below grid g. */
static void v_cycle(nko_mg *mg, int level, const og *g, double *xu, double *xv, const double *bu_, const double *bv_,
                    double *qu, double *qv, double *t, coeffs c, mg_params p) {
    if (level >= mg->nlev) {
        vfill(g, xu, xv, 0.0);
        smooth(g, xu, xv, bu_, bv_, qu, qv, t, c, p.walls, p.smoother, p.coarsest_sweeps);
        return;
    }
    nko_level *L = &mg->lev[level];
    vfill(g, xu, xv, 0.0);
    smooth(g, xu, xv, bu_, bv_, qu, qv, t, c, p.walls, p.smoother, p.sweeps);
    fill_halo(g, xu, F, C, 1);
    fill_halo(g, xv, C, F, 1);
    ssa_operator(g, qu, qv, xu, xv, c.eta, c.etaff, c.bu, c.bv, p.walls);
    vlincomb(g, qu, qv, 1.0, bu_, bv_, -1.0, qu, qv);  /* subtract!(q, b, q) */
    restrict_vec(&L->g, g, L->cx, L->cy, L->b_u, L->b_v, qu, qv, L->eta, c.eta);
    coeffs cc = {L->eta, L->etaff, L->bu, L->bv};
    v_cycle(mg, level + 1, &L->g, L->xu, L->xv, L->b_u, L->b_v, L->qu, L->qv, L->t, cc, p);
    prolongate_vec(g, &L->g, L->cx, L->cy, xu, xv, L->xu, L->xv, c.eta, L->eta);
    smooth(g, xu, xv, bu_, bv_, qu, qv, t, c, p.walls, p.smoother, p.sweeps);
}

static nko_mg *mg_create(const og *fine, int halo, int minimum) {
    nko_mg *mg = calloc(1, sizeof(nko_mg));
    og cur = *fine;
    int cap = 64;
    mg->lev = calloc(cap, sizeof(nko_level));
    for (;;) {
        int cx = coarsening(cur.Nx, halo, minimum), cy = coarsening(cur.Ny, halo, minimum);
        if (!cx && !cy) break;
        nko_level *L = &mg->lev[mg->nlev++];
        og_init(&L->g, cur.Nx / (cx ? 2 : 1), cur.Ny / (cy ? 2 : 1), cur.tx, cur.ty, cur.dx * (cx ? 2 : 1),
                cur.dy * (cy ? 2 : 1));
        L->cx = cx; L->cy = cy;
        double **bufs[] = {&L->eta, &L->etaff, &L->bu, &L->bv, &L->xu, &L->xv, &L->b_u, &L->b_v, &L->qu, &L->qv, &L->t};
        for (size_t b = 0; b < sizeof(bufs) / sizeof(bufs[0]); b++) *bufs[b] = og_alloc(&L->g, 1);
        cur = L->g;
    }
    mg->ru = og_alloc(fine, 1);
    mg->rv = og_alloc(fine, 1);
    mg->t = og_alloc(fine, 1);
    return mg;
}

static void mg_destroy(nko_mg *mg) {
    if (!mg) return;
    for (int l = 0; l < mg->nlev; l++) {
        nko_level *L = &mg->lev[l];
        double *bufs[] = {L->eta, L->etaff, L->bu, L->bv, L->xu, L->xv, L->b_u, L->b_v, L->qu, L->qv, L->t};
        for (size_t b = 0; b < sizeof(bufs) / sizeof(bufs[0]); b++) free(bufs[b]);
    }
    free(mg->lev); free(mg->ru); free(mg->rv); free(mg->t); free(mg);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* The model                                                                                                     */
/* ------------------------------------------------------------------------------------------------------------ */

typedef struct nko_model {
    nk_grid grid;
    nk_physics ph;
    nk_solver_options op;
    og g;
    int Ns;
    double *f[NK_FIELD_COUNT];
    uint8_t *class_u, *class_v;
    double *zf, *zc, *dz, *w2;  /* σ faces (1..Ns+1), centers, layer thickness, layer_weight(k, 2) */
    /* CG workspace (src/Solvers/conjugate_gradient.jl:93-116) and Picard previous velocity */
    double *ru, *rv, *zu, *zv, *pu, *pv, *qu, *qv, *w, *xmu, *xmv;
    nko_mg *mg;
    double weno_bounds[2];  /* WENO(order = 5, bounds): (0, 1e5) unless set */
    char err[256];
} nko_model;

static int nz_of(const nko_model *m, int id) {
    if (LOCS[id].sigma) return m->Ns;
    if ((id == NK_FIELD_A || id == NK_FIELD_ABAR) && m->ph.rate_factor_depth == NK_RATE_FACTOR_3D) return m->Ns;
    return 1;
}

/* sigma_faces (src/Grids/sigma_grid.jl:20-22) and the σ-grid metrics used by column_integrals.jl. Julia's
 * Float64^Int is correctly rounded up to its compensated power; long double reproduces that. */
static double jpow_int(double x, int n) {
    long double r = 1.0L;
    for (int k = 0; k < n; k++) r *= (long double)x;
    return (double)r;
}
static double sigma_face(int k0, int Ns, double refinement) {
    double q = (double)k0 / (double)Ns;
    double r = refinement;
    if (r == floor(r) && r >= 0 && r <= 16) return jpow_int(q, (int)r);
    return pow(q, r);
}
/* power_integral (column_integrals.jl:2) */
static double power_integral(double zm, double zp, int p) { return (jpow_int(1.0 - zm, p + 1) - jpow_int(1.0 - zp, p + 1)) / (double)(p + 1); }

EXPORT nko_model *nko_create(const nk_grid *grid, const nk_physics *ph, const nk_solver_options *op) {
    nko_model *m = calloc(1, sizeof(nko_model));
    m->grid = *grid; m->ph = *ph; m->op = *op;
    m->weno_bounds[0] = 0.0; m->weno_bounds[1] = 1e5;
    og_init(&m->g, grid->Nx, grid->Ny, grid->topology_x, grid->topology_y, grid->dx, grid->dy);
    m->Ns = grid->Nsigma;
    for (int id = 0; id < NK_FIELD_COUNT; id++)
        m->f[id] = og_alloc(&m->g, (LOCS[id].sigma || id == NK_FIELD_A || id == NK_FIELD_ABAR) ? m->Ns : 1);
    m->class_u = calloc(m->g.plane, 1);
    m->class_v = calloc(m->g.plane, 1);
    int Ns = m->Ns;
    m->zf = calloc(Ns + 2, sizeof(double)); m->zc = calloc(Ns + 2, sizeof(double));
    m->dz = calloc(Ns + 2, sizeof(double)); m->w2 = calloc(Ns + 2, sizeof(double));
    for (int k = 1; k <= Ns + 1; k++) m->zf[k] = sigma_face(k - 1, Ns, grid->bed_refinement);
    for (int k = 1; k <= Ns; k++) {
        m->zc[k] = (m->zf[k] + m->zf[k + 1]) / 2.0;
        m->dz[k] = m->zf[k + 1] - m->zf[k];
        m->w2[k] = power_integral(m->zf[k], m->zf[k + 1], 2);
    }
    double **ws[] = {&m->ru, &m->rv, &m->zu, &m->zv, &m->pu, &m->pv, &m->qu, &m->qv, &m->w, &m->xmu, &m->xmv};
    for (size_t b = 0; b < sizeof(ws) / sizeof(ws[0]); b++) *ws[b] = og_alloc(&m->g, 1);
    if (op->preconditioner == NK_PRECONDITIONER_MULTIGRID) m->mg = mg_create(&m->g, grid->halo, op->minimum_coarse_size);
    return m;
}

EXPORT void nko_destroy(nko_model *m) {
    if (!m) return;
    for (int id = 0; id < NK_FIELD_COUNT; id++) free(m->f[id]);
    free(m->class_u); free(m->class_v);
    free(m->zf); free(m->zc); free(m->dz); free(m->w2);
    double *ws[] = {m->ru, m->rv, m->zu, m->zv, m->pu, m->pv, m->qu, m->qv, m->w, m->xmu, m->xmv};
    for (size_t b = 0; b < sizeof(ws) / sizeof(ws[0]); b++) free(ws[b]);
    mg_destroy(m->mg);
    free(m);
}

EXPORT void nko_set_physics(nko_model *m, const nk_physics *ph) { m->ph = *ph; }

/* New options on a live model; the multigrid hierarchy is rebuilt when the preconditioner needs one it lacks or
 * its coarsening rule changed (StressBalanceMultigridPreconditioner(grid; ...) builds it per grid). */
EXPORT void nko_set_options(nko_model *m, const nk_solver_options *op) {
    int rebuild = op->preconditioner == NK_PRECONDITIONER_MULTIGRID &&
                  (!m->mg || op->minimum_coarse_size != m->op.minimum_coarse_size);
    m->op = *op;
    if (rebuild) {
        mg_destroy(m->mg);
        m->mg = mg_create(&m->g, m->grid.halo, op->minimum_coarse_size);
    }
}

/* Threads of the parallel row loops (the timed CPU baseline); returns the count in effect. */
EXPORT int nko_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

EXPORT void nko_field_shape(const nko_model *m, int id, int32_t *shape) {
    shape[0] = nx_of(&m->g, LOCS[id].lx);
    shape[1] = ny_of(&m->g, LOCS[id].ly);
    shape[2] = nz_of(m, id);
}

static void fill_field_halo(nko_model *m, int id) { fill_halo(&m->g, m->f[id], LOCS[id].lx, LOCS[id].ly, nz_of(m, id)); }

EXPORT void nko_set_field(nko_model *m, int id, const double *data, int64_t ld, int64_t lds) {
    int32_t s[3];
    nko_field_shape(m, id, s);
    if (ld == 0) ld = s[0];
    if (lds == 0) lds = ld * s[1];
    for (int k = 0; k < s[2]; k++)
        for (int j = 1; j <= s[1]; j++)
            for (int i = 1; i <= s[0]; i++)
                m->f[id][(size_t)k * m->g.plane + IDX(&m->g, i, j)] = data[(i - 1) + (j - 1) * ld + k * lds];
    fill_field_halo(m, id);
}

EXPORT void nko_get_field(nko_model *m, int id, double *data, int64_t ld, int64_t lds) {
    int32_t s[3];
    nko_field_shape(m, id, s);
    if (ld == 0) ld = s[0];
    if (lds == 0) lds = ld * s[1];
    for (int k = 0; k < s[2]; k++)
        for (int j = 1; j <= s[1]; j++)
            for (int i = 1; i <= s[0]; i++)
                data[(i - 1) + (j - 1) * ld + k * lds] = m->f[id][(size_t)k * m->g.plane + IDX(&m->g, i, j)];
}

EXPORT void nko_get_classes(nko_model *m, uint8_t *cu, int64_t ldu, uint8_t *cv, int64_t ldv) {
    const og *g = &m->g;
    int nxu = nx_of(g, F), nyu = ny_of(g, C), nxv = nx_of(g, C), nyv = ny_of(g, F);
    if (ldu == 0) ldu = nxu;
    if (ldv == 0) ldv = nxv;
    for (int j = 1; j <= nyu; j++) for (int i = 1; i <= nxu; i++) cu[(i - 1) + (j - 1) * ldu] = m->class_u[IDX(g, i, j)];
    for (int j = 1; j <= nyv; j++) for (int i = 1; i <= nxv; i++) cv[(i - 1) + (j - 1) * ldv] = m->class_v[IDX(g, i, j)];
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Geometry (src/Groundings/subgrid_grounded_fraction.jl, src/Models/smooth_ice_mask.jl)                         */
/* ------------------------------------------------------------------------------------------------------------ */

/* ice_mask_value (smooth_ice_mask.jl:50-53) */
EXPORT double nko_ice_mask_value(double H, double hmin, double htr) {
    double t = jclamp((H - hmin) / htr, 0.0, 1.0);
    return t * t * (3.0 - 2.0 * t);
}
static double nonzero(double x) { return x == 0.0 ? 1.0 : x; }
/* segment_grounded_fraction (subgrid_grounded_fraction.jl:159-163) */
EXPORT double nko_segment_grounded_fraction(double f1, double f2) {
    double fp = jmax(f1, 0.0) + jmax(f2, 0.0);
    double sf = fabs(f1) + fabs(f2);
    return sf == 0.0 ? 1.0 : fp / nonzero(sf);
}
/* triangle_grounded_fraction (subgrid_grounded_fraction.jl:166-174) */
EXPORT double nko_triangle_grounded_fraction(double f1, double f2, double f3) {
    double a = jmin(jmin(f1, f2), f3);
    double c = jmax(jmax(f1, f2), f3);
    double b = f1 + f2 + f3 - a - c;
    double one_grounded = c * c / nonzero((c - a) * (c - b));
    double one_floating = 1.0 - a * a / nonzero((b - a) * (c - a));
    double phi = b > 0 ? one_floating : one_grounded;
    return a >= 0 ? 1.0 : (c <= 0 ? 0.0 : phi);
}

static double flotation(const og *g, int i, int j, const double *H, const double *b, double r) { return AT(g, H, i, j) + r * AT(g, b, i, j); }

/* grounded_fractionᶜᶜᶜ (subgrid_grounded_fraction.jl:179-192) */
static double grounded_fraction_cc(const og *g, int i, int j, const double *H, const double *b, double r) {
#define FL(a, c) flotation(g, a, c, H, b, r)
    double fc = FL(i, j);
    double fe = (fc + FL(i + 1, j)) / 2.0, fw = (fc + FL(i - 1, j)) / 2.0;
    double fn = (fc + FL(i, j + 1)) / 2.0, fs = (fc + FL(i, j - 1)) / 2.0;
    double fne = (fc + FL(i + 1, j) + FL(i, j + 1) + FL(i + 1, j + 1)) / 4.0;
    double fnw = (fc + FL(i - 1, j) + FL(i, j + 1) + FL(i - 1, j + 1)) / 4.0;
    double fse = (fc + FL(i + 1, j) + FL(i, j - 1) + FL(i + 1, j - 1)) / 4.0;
    double fsw = (fc + FL(i - 1, j) + FL(i, j - 1) + FL(i - 1, j - 1)) / 4.0;
#undef FL
    double phi = nko_triangle_grounded_fraction(fc, fe, fne) + nko_triangle_grounded_fraction(fc, fn, fne) +
                 nko_triangle_grounded_fraction(fc, fn, fnw) + nko_triangle_grounded_fraction(fc, fw, fnw) +
                 nko_triangle_grounded_fraction(fc, fw, fsw) + nko_triangle_grounded_fraction(fc, fs, fsw) +
                 nko_triangle_grounded_fraction(fc, fs, fse) + nko_triangle_grounded_fraction(fc, fe, fse);
    return phi / 8.0;
}

/* update_geometry! (src/Models/ice_sheet_model_methods.jl:103-111) with compute_surface_elevation!,
 * compute_grounded_fractions!, compute_effective_pressure! (subgrid_grounded_fraction.jl:109-232) and
 * compute_ice_mask! (smooth_ice_mask.jl:60-69). */
EXPORT void nko_update_geometry(nko_model *m) {
    const og *g = &m->g;
    const nk_physics *ph = &m->ph;
    double *H = m->f[NK_FIELD_H], *b = m->f[NK_FIELD_BED], *s = m->f[NK_FIELD_S], *phi = m->f[NK_FIELD_PHI];
    double *phiu = m->f[NK_FIELD_PHIU], *phiv = m->f[NK_FIELD_PHIV], *N = m->f[NK_FIELD_N], *mu = m->f[NK_FIELD_MU];
    double ri_rw = ph->ice_density / ph->ocean_density;
    double rw_ri = ph->ocean_density / ph->ice_density;
    double rig = ph->ice_density * ph->gravitational_acceleration;
    double rwg = ph->ocean_density * ph->gravitational_acceleration;
    int sub = ph->grounding != NK_GROUNDING_NONE;
    for (int j = 1; j <= g->Ny; j++)
        for (int i = 1; i <= g->Nx; i++) {
            AT(g, s, i, j) = sub ? jmax(AT(g, b, i, j) + AT(g, H, i, j), (1.0 - ri_rw) * AT(g, H, i, j)) : AT(g, b, i, j) + AT(g, H, i, j);
            AT(g, phi, i, j) = sub ? grounded_fraction_cc(g, i, j, H, b, rw_ri) : 1.0;
            double P0 = rig * AT(g, H, i, j);
            AT(g, N, i, j) = jmax(P0 + rwg * jmin(AT(g, b, i, j), 0.0), 0.0);
            AT(g, mu, i, j) = nko_ice_mask_value(AT(g, H, i, j), ph->mask_minimum_thickness, ph->mask_transition_thickness);
        }
    for (int j = 1; j <= ny_of(g, C); j++)
        for (int i = 1; i <= nx_of(g, F); i++)
            AT(g, phiu, i, j) = sub ? nko_segment_grounded_fraction(flotation(g, i - 1, j, H, b, rw_ri), flotation(g, i, j, H, b, rw_ri)) : 1.0;
    for (int j = 1; j <= ny_of(g, F); j++)
        for (int i = 1; i <= nx_of(g, C); i++)
            AT(g, phiv, i, j) = sub ? nko_segment_grounded_fraction(flotation(g, i, j - 1, H, b, rw_ri), flotation(g, i, j, H, b, rw_ri)) : 1.0;
    int ids[] = {NK_FIELD_S, NK_FIELD_PHI, NK_FIELD_PHIU, NK_FIELD_PHIV, NK_FIELD_N, NK_FIELD_MU};
    for (size_t k = 0; k < 6; k++) fill_field_halo(m, ids[k]);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Driving stress (src/StressBalances/driving_stress.jl:43-125)                                                  */
/* ------------------------------------------------------------------------------------------------------------ */

static double front_force(const og *g, int i, int j, const double *H, const double *s, double ri, double rw, double gr) {
    double Hi = AT(g, H, i, j);
    double D = jclamp(Hi - AT(g, s, i, j), 0.0, Hi);
    return gr * (ri * (Hi * Hi) - rw * (D * D)) / 2.0;
}
static int has_ice(const og *g, int i, int j, const double *mu) { return AT(g, mu, i, j) > 0; }

EXPORT double nko_grounding_line_stress(double H1, double H2, double s1, double s2, double b1, double b2, double ri,
                                        double rw, double gr, int *straddles) {
    double f1 = H1 + rw / ri * b1;
    double f2 = H2 + rw / ri * b2;
    int st = ((f1 > 0) != (f2 > 0)) & (H1 > 0) & (H2 > 0);
    double theta = f1 / (st ? f1 - f2 : 1.0);
    double Hg = H1 + theta * (H2 - H1);
    double sg = (1.0 - ri / rw) * Hg;
    *straddles = st;
    return -ri * gr * ((H1 + Hg) * (sg - s1) + (Hg + H2) * (s2 - sg)) / 2.0;
}

EXPORT void nko_compute_driving_stress(nko_model *m) {
    const og *g = &m->g;
    const nk_physics *ph = &m->ph;
    const double *H = m->f[NK_FIELD_H], *s = m->f[NK_FIELD_S];
    const double *mu = ph->masking ? m->f[NK_FIELD_MU] : NULL;
    const double *b = ph->grounding == NK_GROUNDING_SUBGRID_DRIVING_STRESS ? m->f[NK_FIELD_BED] : NULL;
    double ri = ph->ice_density, rw = ph->ocean_density, gr = ph->gravitational_acceleration;
    double *tu = m->f[NK_FIELD_TAUU], *tv = m->f[NK_FIELD_TAUV];
    double Az = g->dx * g->dy;
    for (int j = 1; j <= ny_of(g, C); j++)
        for (int i = 1; i <= nx_of(g, F); i++) {
            int wall = inactive_u(g, i, j, mu);
            int front = 0;
            double tf = 0.0;
            if (mu) {  /* x_front_stress (driving_stress.jl:64-70) */
                int west = has_ice(g, i - 1, j, mu) & !has_ice(g, i, j, mu);
                int east = (!has_ice(g, i - 1, j, mu)) & has_ice(g, i, j, mu);
                double Fv = (west ? front_force(g, i - 1, j, H, s, ri, rw, gr) : 0.0) - (east ? front_force(g, i, j, H, s, ri, rw, gr) : 0.0);
                front = west | east;
                tf = g->dy * Fv;
            }
            int st = 0;
            double tg = 0.0;
            if (b) {
                tg = g->dy * nko_grounding_line_stress(AT(g, H, i - 1, j), AT(g, H, i, j), AT(g, s, i - 1, j), AT(g, s, i, j),
                                                       AT(g, b, i - 1, j), AT(g, b, i, j), ri, rw, gr, &st);
            }
            double Hf = (AT(g, H, i - 1, j) + AT(g, H, i, j)) / 2.0;
            double ds = (AT(g, s, i, j) - AT(g, s, i - 1, j)) / g->dx;
            double t = -Az * ri * gr * Hf * ds;
            AT(g, tu, i, j) = wall ? 0.0 : (front ? tf : (st ? tg : t));
            m->class_u[IDX(g, i, j)] = (uint8_t)((periph(g, i, j, F, C) ? NK_CLASS_WALL : 0) |
                                                 (ice_free_u(g, i, j, mu) ? NK_CLASS_ICE_FREE : 0) |
                                                 (front ? NK_CLASS_FRONT : 0) | (st ? NK_CLASS_GROUNDING : 0));
        }
    for (int j = 1; j <= ny_of(g, F); j++)
        for (int i = 1; i <= nx_of(g, C); i++) {
            int wall = inactive_v(g, i, j, mu);
            int front = 0;
            double tf = 0.0;
            if (mu) {  /* y_front_stress (driving_stress.jl:72-78) */
                int south = has_ice(g, i, j - 1, mu) & !has_ice(g, i, j, mu);
                int north = (!has_ice(g, i, j - 1, mu)) & has_ice(g, i, j, mu);
                double Fv = (south ? front_force(g, i, j - 1, H, s, ri, rw, gr) : 0.0) - (north ? front_force(g, i, j, H, s, ri, rw, gr) : 0.0);
                front = south | north;
                tf = g->dx * Fv;
            }
            int st = 0;
            double tg = 0.0;
            if (b) {
                tg = g->dx * nko_grounding_line_stress(AT(g, H, i, j - 1), AT(g, H, i, j), AT(g, s, i, j - 1), AT(g, s, i, j),
                                                       AT(g, b, i, j - 1), AT(g, b, i, j), ri, rw, gr, &st);
            }
            double Hf = (AT(g, H, i, j - 1) + AT(g, H, i, j)) / 2.0;
            double ds = (AT(g, s, i, j) - AT(g, s, i, j - 1)) / g->dy;
            double t = -Az * ri * gr * Hf * ds;
            AT(g, tv, i, j) = wall ? 0.0 : (front ? tf : (st ? tg : t));
            m->class_v[IDX(g, i, j)] = (uint8_t)((periph(g, i, j, C, F) ? NK_CLASS_WALL : 0) |
                                                 (ice_free_v(g, i, j, mu) ? NK_CLASS_ICE_FREE : 0) |
                                                 (front ? NK_CLASS_FRONT : 0) | (st ? NK_CLASS_GROUNDING : 0));
        }
    fill_field_halo(m, NK_FIELD_TAUU);
    fill_field_halo(m, NK_FIELD_TAUV);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Coefficients                                                                                                  */
/* ------------------------------------------------------------------------------------------------------------ */

static double rate_factor(const nko_model *m, int i, int j, int k) {
    const og *g = &m->g;
    const double *A = m->f[NK_FIELD_A];
    if (m->ph.rate_factor_depth == NK_RATE_FACTOR_3D) return A[(size_t)(k - 1) * g->plane + IDX(g, i, j)];
    return AT(g, A, i, j);
}

/* _compute_viscosity! (src/Rheologies/glen_flow_law.jl:107-111) over the σ grid :xyz */
static void compute_viscosity(nko_model *m, const double *u, const double *v) {
    const og *g = &m->g;
    double n = m->ph.glen_exponent, e0 = m->ph.strain_rate_regularization;
    double *es = m->f[NK_FIELD_ETASIGMA];
    for (int k = 1; k <= m->Ns; k++)
        #pragma omp parallel for schedule(static) if (OMP_BIG(g))
        for (int j = 1; j <= g->Ny; j++)
            for (int i = 1; i <= g->Nx; i++) {
                double e2 = eff_strain2(g, i, j, u, v, m->ph.walls) + e0 * e0;
                es[(size_t)(k - 1) * g->plane + IDX(g, i, j)] = pow(rate_factor(m, i, j, k), -1.0 / n) * pow(e2, (1.0 - n) / (2.0 * n)) / 2.0;
            }
}

/* _compute_shear_viscosity! (src/Rheologies/glen_flow_law.jl:134-143): the DIVA vertical shear with the lagged ησ */
static void compute_shear_viscosity(nko_model *m, const double *u, const double *v) {
    const og *g = &m->g;
    double n = m->ph.glen_exponent, e0 = m->ph.strain_rate_regularization;
    double *es = m->f[NK_FIELD_ETASIGMA];
    const double *bbu = m->f[NK_FIELD_BETABU], *bbv = m->f[NK_FIELD_BETABV], *ub = m->f[NK_FIELD_UB], *vb = m->f[NK_FIELD_VB];
    for (int k = 1; k <= m->Ns; k++)
        #pragma omp parallel for schedule(static) if (OMP_BIG(g))
        for (int j = 1; j <= g->Ny; j++)
            for (int i = 1; i <= g->Nx; i++) {
                double zeta = m->zc[k];
                double *p = &es[(size_t)(k - 1) * g->plane + IDX(g, i, j)];
                double em = *p;
                em = em == 0.0 ? INFINITY : em;
                double tx = (AT(g, bbu, i, j) * AT(g, ub, i, j) + AT(g, bbu, i + 1, j) * AT(g, ub, i + 1, j)) / 2.0;
                double ty = (AT(g, bbv, i, j) * AT(g, vb, i, j) + AT(g, bbv, i, j + 1) * AT(g, vb, i, j + 1)) / 2.0;
                double exz = tx * (1.0 - zeta) / (2.0 * em);
                double eyz = ty * (1.0 - zeta) / (2.0 * em);
                double e2 = eff_strain2(g, i, j, u, v, m->ph.walls) + exz * exz + eyz * eyz + e0 * e0;
                *p = pow(rate_factor(m, i, j, k), -1.0 / n) * pow(e2, (1.0 - n) / (2.0 * n)) / 2.0;
            }
}

/* _compute_column_integrals! (src/StressBalances/column_integrals.jl:61-72) */
static void compute_column_integrals(nko_model *m) {
    const og *g = &m->g;
    const double *es = m->f[NK_FIELD_ETASIGMA], *H = m->f[NK_FIELD_H];
    const double *mu = m->ph.masking ? m->f[NK_FIELD_MU] : NULL;
    double *eta = m->f[NK_FIELD_ETAH], *F2 = m->f[NK_FIELD_F2];
    #pragma omp parallel for schedule(static) if (OMP_BIG(g))
    for (int j = 1; j <= g->Ny; j++)
        for (int i = 1; i <= g->Nx; i++) {
            double e = 0.0, Fs = 0.0;
            for (int k = 1; k <= m->Ns; k++) {
                double ek = es[(size_t)(k - 1) * g->plane + IDX(g, i, j)];
                e += ek * m->dz[k];
                Fs += 1.0 / ek * m->w2[k];
            }
            AT(g, eta, i, j) = e * AT(g, H, i, j) * (mu ? AT(g, mu, i, j) : 1.0);
            AT(g, F2, i, j) = Fs * AT(g, H, i, j);
        }
    fill_field_halo(m, NK_FIELD_ETAH);
    fill_field_halo(m, NK_FIELD_F2);
}

/* basal_drag (src/Frictions/frictions.jl:83-89) */
EXPORT double nko_basal_drag(double U, double Cf, double Nf, int law, double m_, double alpha2) {
    if (law == NK_FRICTION_WEERTMAN) return Cf * pow(U, m_ - 1.0);
    double tc = alpha2 * Nf;
    return Cf * pow(U, m_ - 1.0) * tc / pow(pow(Cf, 1.0 / m_) * U + pow(tc, 1.0 / m_), m_);
}

/* compute_basal_drag! (src/Frictions/frictions.jl:127-160): β at velocity points from (u, v). */
static void compute_basal_drag(nko_model *m, double *bu, double *bv, const double *u, const double *v) {
    const nk_physics *ph = &m->ph;
    if (ph->friction == NK_FRICTION_NONE) return;
    const og *g = &m->g;
    const double *Cc = m->f[NK_FIELD_C], *N = m->f[NK_FIELD_N], *phiu = m->f[NK_FIELD_PHIU], *phiv = m->f[NK_FIELD_PHIV];
    int useN = ph->friction == NK_FRICTION_REGULARIZED_COULOMB;
    double u0 = ph->velocity_regularization;
    #pragma omp parallel for schedule(static) if (OMP_BIG(g))
    for (int j = 1; j <= ny_of(g, C); j++)
        for (int i = 1; i <= nx_of(g, F); i++) {
            double vm = ((AT(g, v, i - 1, j) + AT(g, v, i, j)) / 2.0 + (AT(g, v, i - 1, j + 1) + AT(g, v, i, j + 1)) / 2.0) / 2.0;
            double U = sqrt(AT(g, u, i, j) * AT(g, u, i, j) + vm * vm + u0 * u0);
            double Cf = (AT(g, Cc, i - 1, j) + AT(g, Cc, i, j)) / 2.0;
            double Nf = useN ? (AT(g, N, i - 1, j) + AT(g, N, i, j)) / 2.0 : 0.0;
            double beta = AT(g, phiu, i, j) * nko_basal_drag(U, Cf, Nf, ph->friction, ph->friction_exponent, ph->coulomb_coefficient);
            AT(g, bu, i, j) = periph(g, i, j, F, C) ? 0.0 : beta;
        }
    #pragma omp parallel for schedule(static) if (OMP_BIG(g))
    for (int j = 1; j <= ny_of(g, F); j++)
        for (int i = 1; i <= nx_of(g, C); i++) {
            double um = ((AT(g, u, i, j - 1) + AT(g, u, i + 1, j - 1)) / 2.0 + (AT(g, u, i, j) + AT(g, u, i + 1, j)) / 2.0) / 2.0;
            double U = sqrt(AT(g, v, i, j) * AT(g, v, i, j) + um * um + u0 * u0);
            double Cf = (AT(g, Cc, i, j - 1) + AT(g, Cc, i, j)) / 2.0;
            double Nf = useN ? (AT(g, N, i, j - 1) + AT(g, N, i, j)) / 2.0 : 0.0;
            double beta = AT(g, phiv, i, j) * nko_basal_drag(U, Cf, Nf, ph->friction, ph->friction_exponent, ph->coulomb_coefficient);
            AT(g, bv, i, j) = periph(g, i, j, C, F) ? 0.0 : beta;
        }
    fill_halo(g, bu, F, C, 1);
    fill_halo(g, bv, C, F, 1);
}

/* DIVA shear factor 1 + βᵇ ℑx(F₂) (src/StressBalances/diva.jl:2-3) */
static double x_shear_factor(const og *g, int i, int j, const double *bb, const double *F2) {
    return 1.0 + AT(g, bb, i, j) * ((AT(g, F2, i - 1, j) + AT(g, F2, i, j)) / 2.0);
}
static double y_shear_factor(const og *g, int i, int j, const double *bb, const double *F2) {
    return 1.0 + AT(g, bb, i, j) * ((AT(g, F2, i, j - 1) + AT(g, F2, i, j)) / 2.0);
}

/* compute_basal_velocities! (diva.jl:11-27) */
static void compute_basal_velocities(nko_model *m) {
    const og *g = &m->g;
    const double *u = m->f[NK_FIELD_U], *v = m->f[NK_FIELD_V], *F2 = m->f[NK_FIELD_F2];
    const double *bbu = m->f[NK_FIELD_BETABU], *bbv = m->f[NK_FIELD_BETABV];
    double *ub = m->f[NK_FIELD_UB], *vb = m->f[NK_FIELD_VB];
    for (int j = 1; j <= ny_of(g, C); j++)
        for (int i = 1; i <= nx_of(g, F); i++) AT(g, ub, i, j) = AT(g, u, i, j) / x_shear_factor(g, i, j, bbu, F2);
    for (int j = 1; j <= ny_of(g, F); j++)
        for (int i = 1; i <= nx_of(g, C); i++) AT(g, vb, i, j) = AT(g, v, i, j) / y_shear_factor(g, i, j, bbv, F2);
    fill_field_halo(m, NK_FIELD_UB);
    fill_field_halo(m, NK_FIELD_VB);
}

/* update_stress_balance_coefficients! (src/StressBalances/stress_balance_coefficients.jl:10-20) and update_diva_coefficients!
 * (diva.jl:61-75), at the velocity (U, V). */
EXPORT void nko_update_coefficients(nko_model *m) {
    const og *g = &m->g;
    double *u = m->f[NK_FIELD_U], *v = m->f[NK_FIELD_V];
    fill_field_halo(m, NK_FIELD_U);
    fill_field_halo(m, NK_FIELD_V);
    if (m->ph.stress_balance == NK_SSA) {
        compute_viscosity(m, u, v);
        compute_column_integrals(m);
        interpolate_to_corners(g, m->f[NK_FIELD_ETAHFF], m->f[NK_FIELD_ETAH]);
        compute_basal_drag(m, m->f[NK_FIELD_BETAU], m->f[NK_FIELD_BETAV], u, v);
    } else {
        compute_basal_velocities(m);
        compute_basal_drag(m, m->f[NK_FIELD_BETABU], m->f[NK_FIELD_BETABV], m->f[NK_FIELD_UB], m->f[NK_FIELD_VB]);
        compute_shear_viscosity(m, u, v);
        compute_column_integrals(m);
        interpolate_to_corners(g, m->f[NK_FIELD_ETAHFF], m->f[NK_FIELD_ETAH]);
        /* compute_effective_drag! (diva.jl:35-51) */
        const double *F2 = m->f[NK_FIELD_F2], *bbu = m->f[NK_FIELD_BETABU], *bbv = m->f[NK_FIELD_BETABV];
        double *bu = m->f[NK_FIELD_BETAU], *bv = m->f[NK_FIELD_BETAV];
        for (int j = 1; j <= ny_of(g, C); j++)
            for (int i = 1; i <= nx_of(g, F); i++) AT(g, bu, i, j) = AT(g, bbu, i, j) / x_shear_factor(g, i, j, bbu, F2);
        for (int j = 1; j <= ny_of(g, F); j++)
            for (int i = 1; i <= nx_of(g, C); i++) AT(g, bv, i, j) = AT(g, bbv, i, j) / y_shear_factor(g, i, j, bbv, F2);
        fill_field_halo(m, NK_FIELD_BETAU);
        fill_field_halo(m, NK_FIELD_BETAV);
    }
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Operator, preconditioner and PCG entry points                                                                 */
/* ------------------------------------------------------------------------------------------------------------ */

EXPORT void nko_apply_operator(nko_model *m) {
    fill_field_halo(m, NK_FIELD_PU);
    fill_field_halo(m, NK_FIELD_PV);
    ssa_operator(&m->g, m->f[NK_FIELD_QU], m->f[NK_FIELD_QV], m->f[NK_FIELD_PU], m->f[NK_FIELD_PV], m->f[NK_FIELD_ETAH],
                 m->f[NK_FIELD_ETAHFF], m->f[NK_FIELD_BETAU], m->f[NK_FIELD_BETAV], m->ph.walls);
}

static coeffs model_coeffs(nko_model *m) {
    coeffs c = {m->f[NK_FIELD_ETAH], m->f[NK_FIELD_ETAHFF], m->f[NK_FIELD_BETAU], m->f[NK_FIELD_BETAV]};
    return c;
}

/* update_preconditioner! (src/Solvers/conjugate_gradient.jl:177, multigrid_preconditioner.jl:128) */
static void update_preconditioner(nko_model *m) {
    if (m->op.preconditioner == NK_PRECONDITIONER_MULTIGRID && m->mg)
        coarsen_coefficients(m->mg, &m->g, m->f[NK_FIELD_ETAH], m->f[NK_FIELD_BETAU], m->f[NK_FIELD_BETAV]);
}

/* precondition! for nothing / StressBalanceDiagonalPreconditioner / StressBalanceMultigridPreconditioner */
static void precondition(nko_model *m, double *zu, double *zv, const double *ru, const double *rv) {
    const og *g = &m->g;
    coeffs c = model_coeffs(m);
    switch (m->op.preconditioner) {
    case NK_PRECONDITIONER_NONE:
        vlincomb(g, zu, zv, 1.0, ru, rv, 0.0, ru, rv);  /* copyto! */
        break;
    case NK_PRECONDITIONER_DIAGONAL:
        diagonal_precondition(g, zu, zv, ru, rv, c.eta, c.etaff, c.bu, c.bv, m->ph.walls);
        break;
    default: {
        mg_params p = {m->op.smoother, m->op.smoothing_sweeps, m->op.coarsest_sweeps, m->ph.walls};
        v_cycle(m->mg, 0, g, zu, zv, ru, rv, m->mg->ru, m->mg->rv, m->mg->t, c, p);
    }
    }
}

EXPORT void nko_precondition(nko_model *m) {
    update_preconditioner(m);
    fill_field_halo(m, NK_FIELD_PU);
    fill_field_halo(m, NK_FIELD_PV);
    precondition(m, m->f[NK_FIELD_QU], m->f[NK_FIELD_QV], m->f[NK_FIELD_PU], m->f[NK_FIELD_PV]);
}

/* conjugate_gradient! and the step (src/Solvers/conjugate_gradient.jl:158-223). Returns the iteration count; the
 * final residual norm goes to *rnorm. */
static int conjugate_gradient(nko_model *m, double *xu, double *xv, const double *bu_, const double *bv_, double *rnorm) {
    const og *g = &m->g;
    coeffs c = model_coeffs(m);
    int walls = m->ph.walls;
    double *ru = m->ru, *rv = m->rv, *zu = m->zu, *zv = m->zv, *pu = m->pu, *pv = m->pv, *qu = m->qu, *qv = m->qv, *w = m->w;
    fill_halo(g, xu, F, C, 1);
    fill_halo(g, xv, C, F, 1);
    ssa_operator(g, qu, qv, xu, xv, c.eta, c.etaff, c.bu, c.bv, walls);
    vlincomb(g, ru, rv, 1.0, bu_, bv_, -1.0, qu, qv);
    update_preconditioner(m);
    precondition(m, zu, zv, ru, rv);
    vlincomb(g, pu, pv, 1.0, zu, zv, 0.0, zu, zv);
    double rho = vdot(g, w, ru, rv, zu, zv);
    const nk_control *ctl = &m->op.linear;
    int fixed = ctl->kind == NK_FIXED_ITERATIONS;
    double r0 = fixed ? 0.0 : sqrt(vdot(g, w, ru, rv, ru, rv));
    double tol = jmax(ctl->reltol * r0, ctl->abstol);
    double res = r0;
    int it = 0;
    for (;;) {
        if (fixed) { if (it >= ctl->iterations) break; }
        else if (!(res > tol && it < ctl->maxiter)) break;
        /* conjugate_gradient_step! */
        fill_halo(g, pu, F, C, 1);
        fill_halo(g, pv, C, F, 1);
        ssa_operator(g, qu, qv, pu, pv, c.eta, c.etaff, c.bu, c.bv, walls);
        double pq = vdot(g, w, pu, pv, qu, qv);
        double alpha = pq == 0.0 ? 0.0 : rho / pq;
        vlincomb(g, xu, xv, 1.0, xu, xv, alpha, pu, pv);
        vlincomb(g, ru, rv, 1.0, ru, rv, -alpha, qu, qv);
        precondition(m, zu, zv, ru, rv);
        double rho1 = vdot(g, w, ru, rv, zu, zv);
        double beta = rho == 0.0 ? 0.0 : rho1 / rho;
        vlincomb(g, pu, pv, 1.0, zu, zv, beta, pu, pv);
        rho = rho1;
        if (!fixed) res = sqrt(vdot(g, w, ru, rv, ru, rv));
        it++;
    }
    *rnorm = fixed ? sqrt(vdot(g, w, ru, rv, ru, rv)) : res;
    return it;
}

EXPORT int nko_solve(nko_model *m, nk_info *info) {
    double rn = 0.0;
    int it = conjugate_gradient(m, m->f[NK_FIELD_U], m->f[NK_FIELD_V], m->f[NK_FIELD_TAUU], m->f[NK_FIELD_TAUV], &rn);
    fill_field_halo(m, NK_FIELD_U);
    fill_field_halo(m, NK_FIELD_V);
    if (info) { info->linear_iterations = it; info->total_linear_iterations += it; info->last_residual_norm = rn; }
    return 0;
}

/* picard_solve! for Convergence and FixedIterations (src/Solvers/picard.jl:71-90) */
EXPORT int nko_picard_solve(nko_model *m, nk_info *info) {
    const og *g = &m->g;
    double *xu = m->f[NK_FIELD_U], *xv = m->f[NK_FIELD_V];
    const nk_control *ctl = &m->op.picard;
    int it = 0;
    double D = INFINITY, X = 0.0;
    if (info) info->total_linear_iterations = 0;
    for (;;) {
        if (ctl->kind == NK_FIXED_ITERATIONS) { if (it >= ctl->iterations) break; }
        else if (!(D > jmax(ctl->reltol * X, ctl->abstol) && it < ctl->maxiter)) break;
        if (ctl->kind != NK_FIXED_ITERATIONS) vlincomb(g, m->xmu, m->xmv, 1.0, xu, xv, 0.0, xu, xv);  /* copyto!(x⁻, x) */
        nko_update_coefficients(m);
        nko_solve(m, info);
        if (ctl->kind != NK_FIXED_ITERATIONS) {
            vlincomb(g, m->xmu, m->xmv, 1.0, m->xmu, m->xmv, -1.0, xu, xv);  /* axpy!(-1, x, x⁻) */
            D = sqrt(vdot(g, m->w, m->xmu, m->xmv, m->xmu, m->xmv));
            X = sqrt(vdot(g, m->w, xu, xv, xu, xv));
        }
        it++;
    }
    if (info) { info->picard_iterations = it; info->last_picard_change = D; info->last_picard_norm = X; }
    return it;
}

/* compute_velocity_profile! (src/StressBalances/diva.jl:89-109) and extend_depth_average!
 * (src/Models/ice_sheet_model_methods.jl:167-180) */
static void post_process(nko_model *m) {
    const og *g = &m->g;
    int Ns = m->Ns;
    double *u = m->f[NK_FIELD_U], *v = m->f[NK_FIELD_V], *ub = m->f[NK_FIELD_UB], *vb = m->f[NK_FIELD_VB];
    double *us = m->f[NK_FIELD_USIGMA], *vs = m->f[NK_FIELD_VSIGMA];
    if (m->ph.stress_balance == NK_SSA) {
        memcpy(ub, u, g->plane * sizeof(double));
        memcpy(vb, v, g->plane * sizeof(double));
        for (int k = 1; k <= Ns; k++) {
            for (int j = 1; j <= ny_of(g, C); j++)
                for (int i = 1; i <= nx_of(g, F); i++) us[(size_t)(k - 1) * g->plane + IDX(g, i, j)] = AT(g, u, i, j);
            for (int j = 1; j <= ny_of(g, F); j++)
                for (int i = 1; i <= nx_of(g, C); i++) vs[(size_t)(k - 1) * g->plane + IDX(g, i, j)] = AT(g, v, i, j);
        }
        return;
    }
    fill_field_halo(m, NK_FIELD_ETASIGMA);
    compute_basal_velocities(m);
    const double *es = m->f[NK_FIELD_ETASIGMA], *H = m->f[NK_FIELD_H];
    const double *bbu = m->f[NK_FIELD_BETABU], *bbv = m->f[NK_FIELD_BETABV];
    /* shear_integral (src/StressBalances/column_integrals.jl:42) with profile_weight (:149-154), p = 1 */
    double *S = calloc(g->plane, sizeof(double));
    for (int k = 1; k <= Ns; k++) {
        for (int j = 1 - g->H; j <= g->Ny + 1 + g->H; j++)
            for (int i = 1 - g->H; i <= g->Nx + 1 + g->H; i++) {
                double I = 0.0;
                for (int kp = 1; kp <= k; kp++) {
                    double zm = m->zf[kp], zp = (kp == k) ? m->zc[k] : m->zf[kp + 1];
                    double w = power_integral(zm, zp, 1);
                    I += 1.0 / es[(size_t)(kp - 1) * g->plane + IDX(g, i, j)] * w;
                }
                AT(g, S, i, j) = I * AT(g, H, i, j);
            }
        for (int j = 1; j <= ny_of(g, C); j++)
            for (int i = 1; i <= nx_of(g, F); i++) {
                double F1 = (AT(g, S, i - 1, j) + AT(g, S, i, j)) / 2.0;
                us[(size_t)(k - 1) * g->plane + IDX(g, i, j)] = AT(g, ub, i, j) * (1.0 + AT(g, bbu, i, j) * F1);
            }
        for (int j = 1; j <= ny_of(g, F); j++)
            for (int i = 1; i <= nx_of(g, C); i++) {
                double F1 = (AT(g, S, i, j - 1) + AT(g, S, i, j)) / 2.0;
                vs[(size_t)(k - 1) * g->plane + IDX(g, i, j)] = AT(g, vb, i, j) * (1.0 + AT(g, bbv, i, j) * F1);
            }
    }
    free(S);
}

/* compute_velocity_profile! / extend_depth_average! alone, after a picard_solve! (test/test_diva.jl:33-34) */
EXPORT void nko_compute_velocity_profile(nko_model *m) {
    fill_field_halo(m, NK_FIELD_U);
    fill_field_halo(m, NK_FIELD_V);
    post_process(m);
}

/* update_velocities! (src/Models/ice_sheet_model_methods.jl:119-147) */
EXPORT int nko_update_velocities(nko_model *m, int geometry, nk_info *info) {
    if (geometry) nko_update_geometry(m);
    nko_compute_driving_stress(m);
    nko_picard_solve(m, info);
    post_process(m);
    return 0;
}

/* Multigrid level shapes */
EXPORT int nko_multigrid_levels(const nko_model *m, int32_t *shapes, int max_levels) {
    if (!m->mg) return 0;
    for (int l = 0; l < m->mg->nlev && l < max_levels; l++) { shapes[2 * l] = m->mg->lev[l].g.Nx; shapes[2 * l + 1] = m->mg->lev[l].g.Ny; }
    return m->mg->nlev;
}

/* Transfers between the finest grid and the first coarse level, after coarsen_coefficients!, for the
 * restriction = prolongationᵀ check of test/test_multigrid_preconditioner.jl:40-51. Arrays are interior-packed
 * at the respective grids' velocity shapes. */
static void copy_in(const og *g, double *f, const double *a, int lx, int ly) {
    int nx = nx_of(g, lx), ny = ny_of(g, ly);
    memset(f, 0, g->plane * sizeof(double));
    for (int j = 1; j <= ny; j++) for (int i = 1; i <= nx; i++) AT(g, f, i, j) = a[(i - 1) + (j - 1) * nx];
    fill_halo(g, f, lx, ly, 1);
}
static void copy_out(const og *g, const double *f, double *a, int lx, int ly) {
    int nx = nx_of(g, lx), ny = ny_of(g, ly);
    for (int j = 1; j <= ny; j++) for (int i = 1; i <= nx; i++) a[(i - 1) + (j - 1) * nx] = AT(g, f, i, j);
}
EXPORT int nko_mg_transfer(nko_model *m, int direction, const double *in_u, const double *in_v, double *out_u, double *out_v) {
    if (!m->mg || m->mg->nlev == 0) return 1;
    update_preconditioner(m);
    nko_level *L = &m->mg->lev[0];
    const og *fg = &m->g, *cg = &L->g;
    double *fu = og_alloc(fg, 1), *fv = og_alloc(fg, 1), *cu = og_alloc(cg, 1), *cv = og_alloc(cg, 1);
    if (direction == 0) {  /* restrict fine → coarse */
        copy_in(fg, fu, in_u, F, C); copy_in(fg, fv, in_v, C, F);
        restrict_vec(cg, fg, L->cx, L->cy, cu, cv, fu, fv, L->eta, m->f[NK_FIELD_ETAH]);
        copy_out(cg, cu, out_u, F, C); copy_out(cg, cv, out_v, C, F);
    } else {  /* prolongate coarse → fine, into zeros */
        copy_in(cg, cu, in_u, F, C); copy_in(cg, cv, in_v, C, F);
        prolongate_vec(fg, cg, L->cx, L->cy, fu, fv, cu, cv, m->f[NK_FIELD_ETAH], L->eta);
        copy_out(fg, fu, out_u, F, C); copy_out(fg, fv, out_v, C, F);
    }
    free(fu); free(fv); free(cu); free(cv);
    return 0;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Rate factor from temperature (config input preparation)                                                       */
/* ------------------------------------------------------------------------------------------------------------ */

/* Defaults of EnthalpyThermodynamics (src/Thermodynamics/enthalpy_thermodynamics.jl:96-105) and
 * ArrheniusRateFactor (src/Rheologies/arrhenius_rate_factor.jl:60-69). */
static const double E_CI = 2009.0, E_L = 3.34e5, E_BETA = 7.9e-8, E_T0 = 273.15, E_TR = 223.15;
static const double A_COLD = 3.61e-13, A_WARM = 1.73e3, Q_COLD = 6.0e4, Q_WARM = 13.9e4, T_TRANS = 263.15,
                    T_WIDTH = 1.0, R_GAS = 8.31441, C_WATER = 181.25, W_MAX = 0.01;

/* arrhenius_rate_factor (arrhenius_rate_factor.jl:102-110) */
EXPORT double nko_arrhenius_rate_factor(double T, double omega) {
    double lc = log(A_COLD) - Q_COLD / (R_GAS * T);
    double lw = log(A_WARM) - Q_WARM / (R_GAS * T);
    double x = jclamp((T - T_TRANS) / T_WIDTH + 1.0 / 2.0, 0.0, 1.0);
    double s = x * x * (3.0 - 2.0 * x);
    double soft = 1.0 + C_WATER * jmin(omega, W_MAX);
    return exp((1.0 - s) * lc + s * lw) * soft;
}
static double pm_temperature(double p) { return E_T0 - E_BETA * p; }                 /* enthalpy_conversions.jl:6 */
static double pm_enthalpy(double p) { return E_CI * (pm_temperature(p) - E_TR); }    /* :123 */
static double e_temperature(double E, double p) { return jmin(pm_temperature(p), E_TR + E / E_CI); }  /* :145 */
static double e_water(double E, double p) { return jmax(0.0, E - pm_enthalpy(p)) / E_L; }            /* :152 */
static double e_enthalpy(double T, double w, double p) { double Tt = jmin(T, pm_temperature(p)); return E_CI * (Tt - E_TR) + w * E_L; }  /* :159-162 */

/* set_enthalpy_from_temperature! → _compute_thermal_diagnostics! → _compute_rate_factor!
 * (src/Models/ice_sheet_model_methods.jl:47-57, src/Thermodynamics/column_energy_solve.jl:105-117) */
EXPORT void nko_rate_factor_from_temperature(nko_model *m) {
    const og *g = &m->g;
    const double *T = m->f[NK_FIELD_TEMPERATURE], *H = m->f[NK_FIELD_H];
    double *A = m->f[NK_FIELD_A];
    double rg = m->ph.ice_density * m->ph.gravitational_acceleration;
    for (int k = 1; k <= m->Ns; k++)
        for (int j = 1; j <= g->Ny; j++)
            for (int i = 1; i <= g->Nx; i++) {
                size_t o = (size_t)(k - 1) * g->plane + IDX(g, i, j);
                double p = rg * AT(g, H, i, j) * (1.0 - m->zc[k]);   /* layer_pressure (enthalpy_conversions.jl:61-64) */
                double E = e_enthalpy(T[o], 0.0, p);
                m->f[NK_FIELD_ENTHALPY][o] = E;
                double Tk = e_temperature(E, p), w = e_water(E, p);
                double Tp = Tk - pm_temperature(p) + E_T0;
                A[o] = nko_arrhenius_rate_factor(Tp, w);
            }
    fill_field_halo(m, NK_FIELD_A);
    fill_field_halo(m, NK_FIELD_ENTHALPY);
}

/* update_thermal_diagnostics! (src/Thermodynamics/column_energy_solve.jl:85-117): _compute_thermal_diagnostics!
 * (:113-117) then _compute_rate_factor! (:106-111) from the enthalpy field. */
EXPORT void nko_thermal_diagnostics(nko_model *m) {
    const og *g = &m->g;
    const double *E = m->f[NK_FIELD_ENTHALPY], *H = m->f[NK_FIELD_H];
    double *T = m->f[NK_FIELD_TEMPERATURE], *W = m->f[NK_FIELD_WATER_FRACTION], *A = m->f[NK_FIELD_A];
    double rg = m->ph.ice_density * m->ph.gravitational_acceleration;
    for (int k = 1; k <= m->Ns; k++)
        for (int j = 1; j <= g->Ny; j++)
            for (int i = 1; i <= g->Nx; i++) {
                size_t o = (size_t)(k - 1) * g->plane + IDX(g, i, j);
                double p = rg * AT(g, H, i, j) * (1.0 - m->zc[k]);   /* layer_pressure (enthalpy_conversions.jl:61-64) */
                T[o] = e_temperature(E[o], p);
                W[o] = e_water(E[o], p);
                A[o] = nko_arrhenius_rate_factor(T[o] - pm_temperature(p) + E_T0, W[o]);
            }
    fill_field_halo(m, NK_FIELD_TEMPERATURE);
    fill_field_halo(m, NK_FIELD_WATER_FRACTION);
    fill_field_halo(m, NK_FIELD_A);
}

/* enthalpy (enthalpy_conversions.jl:49-52) and its inverses at pressure p, for tests */
EXPORT void nko_enthalpy_conversions(double T, double w, double p, double *E, double *T_of_E, double *w_of_E) {
    *E = e_enthalpy(T, w, p);
    *T_of_E = e_temperature(*E, p);
    *w_of_E = e_water(*E, p);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Thickness transport (src/Models/time_stepping.jl:1-72), flux form of Oceananigans' div_Uc (third-party: the      */
/* Oceananigans 0.117 dependency is not in the reference tree; restated from its published finite-volume form)    */
/* ------------------------------------------------------------------------------------------------------------ */

/* face thickness: UpwindBiased(order = 1) takes the upwind cell (the right one unless the velocity is positive),
 * Centered(order = 2) the mean of the two cells (halo cells hold the zero-gradient / periodic fill) */
static double face_H(double left, double right, double vel, int adv) { return adv == 0 ? (vel > 0.0 ? left : right) : (left + right) / 2.0; }

/* WENO(order = 5) face thickness and the bounds-preserving limiter of WENO(order = 5, bounds) — the reference's
 * default thickness advection (src/Models/ice_sheet_model.jl:123; limiter updated before each tendency,
 * src/Models/time_stepping.jl:37; face values rescaled per cell as in src/Models/partial_cell_front.jl:139-151).
 * The scheme is Oceananigans' (0.117, outside the reference tree), restated from its published form: Jiang & Shu
 * (1996) stencils and smoothness indicators with WENO-Z weights (Borges et al. 2008, αₖ = dₖ(1 + τ/(βₖ + ε)),
 * ε = 1e-8), third order and then first-order upwind where a stencil would leave a Bounded direction, and the
 * Zhang & Shu (2010) scaling limiter with ω̂ = 5/18, one θ per cell over both directions. Parity with Oceananigans'
 * own operation order and boundary rule is unpinned; the reference's transport tests and the scheme's order pin it. */
#define WENO_EPS 1e-8
#define WENO_OMEGA (5.0 / 18.0)
static double weno5z(double a, double b, double c, double d, double e) {
    double q0 = (2.0 * a - 7.0 * b + 11.0 * c) / 6.0;
    double q1 = (-b + 5.0 * c + 2.0 * d) / 6.0;
    double q2 = (2.0 * c + 5.0 * d - e) / 6.0;
    double s0 = a - 2.0 * b + c, t0 = a - 4.0 * b + 3.0 * c;
    double s1 = b - 2.0 * c + d, t1 = b - d;
    double s2 = c - 2.0 * d + e, t2 = 3.0 * c - 4.0 * d + e;
    double b0 = 13.0 / 12.0 * (s0 * s0) + 0.25 * (t0 * t0);
    double b1 = 13.0 / 12.0 * (s1 * s1) + 0.25 * (t1 * t1);
    double b2 = 13.0 / 12.0 * (s2 * s2) + 0.25 * (t2 * t2);
    double tau = fabs(b0 - b2);
    double r0 = tau / (b0 + WENO_EPS), r1 = tau / (b1 + WENO_EPS), r2 = tau / (b2 + WENO_EPS);
    double a0 = 0.1 * (1.0 + r0), a1 = 0.6 * (1.0 + r1), a2 = 0.3 * (1.0 + r2);
    return (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
}
static double weno3z(double a, double b, double c) {
    double q0 = (-a + 3.0 * b) / 2.0, q1 = (b + c) / 2.0;
    double b0 = (b - a) * (b - a), b1 = (c - b) * (c - b);
    double tau = fabs(b0 - b1);
    double r0 = tau / (b0 + WENO_EPS), r1 = tau / (b1 + WENO_EPS);
    double a0 = (1.0 / 3.0) * (1.0 + r0), a1 = (2.0 / 3.0) * (1.0 + r1);
    return (a0 * q0 + a1 * q1) / (a0 + a1);
}
EXPORT double nko_weno5z(double a, double b, double c, double d, double e) { return weno5z(a, b, c, d, e); }
EXPORT double nko_weno3z(double a, double b, double c) { return weno3z(a, b, c); }
static int weno_order(int k, int N, int bounded) {
    if (!bounded) return 5;
    if (k >= 4 && k <= N - 2) return 5;
    if (k >= 3 && k <= N - 1) return 3;
    return 1;
}
/* reconstruction at face (i, j) of direction dir from the upwind cell below the face (left) or at it (right); the
 * halo (3 cells) holds the clamped or wrapped values */
static double weno_face(const og *g, const double *H, int dir, int i, int j, int left) {
#define WC(d) (dir == 0 ? AT(g, H, i + (d), j) : AT(g, H, i, j + (d)))
    int order = dir == 0 ? weno_order(i, g->Nx, g->tx) : weno_order(j, g->Ny, g->ty);
    double r;
    if (left) r = order == 5 ? weno5z(WC(-3), WC(-2), WC(-1), WC(0), WC(1)) : (order == 3 ? weno3z(WC(-2), WC(-1), WC(0)) : WC(-1));
    else r = order == 5 ? weno5z(WC(2), WC(1), WC(0), WC(-1), WC(-2)) : (order == 3 ? weno3z(WC(1), WC(0), WC(-1)) : WC(0));
#undef WC
    return r;
}
static double limit_ratio(double bound, double c, double ext) {
    double d = ext - c;
    return d == 0.0 ? 1.0 : fabs((bound - c) / d);
}
EXPORT void nko_set_thickness_bounds(nko_model *m, double lower, double upper) {
    m->weno_bounds[0] = lower;
    m->weno_bounds[1] = upper;
}
/* the face thickness of the WENO schemes: the upwind cell's reconstruction, rescaled by its θ when bounded;
 * upwind_biased_product ((u + |u|) ψᴸ + (u − |u|) ψᴿ) / 2 */
static double weno_flux_H(const og *g, const double *H, const double *theta, int dir, int i, int j, double vel) {
    double hl = weno_face(g, H, dir, i, j, 1), hr = weno_face(g, H, dir, i, j, 0);
    double ul = hl, ur = hr;
    if (theta) {
        int il = dir == 0 ? i - 1 : i, jl = dir == 0 ? j : j - 1;
        double cl = AT(g, H, il, jl), cr = AT(g, H, i, j);
        ul = AT(g, theta, il, jl) * (hl - cl) + cl;
        ur = AT(g, theta, i, j) * (hr - cr) + cr;
    }
    return ((vel + fabs(vel)) * ul + (vel - fabs(vel)) * ur) / 2.0;
}

/* compute_thickness_tendency! / _compute_thickness_tendency! (time_stepping.jl:34-63); adv 0 UpwindBiased(1),
 * 1 Centered(2), 2 WENO(5), 3 WENO(5) with the bounds-preserving limiter */
EXPORT void nko_thickness_tendency(nko_model *m, int adv, int coupled) {
    const og *g = &m->g;
    const double *u = m->f[NK_FIELD_U], *v = m->f[NK_FIELD_V], *H = m->f[NK_FIELD_H];
    const double *a = m->f[NK_FIELD_SMB], *mm = m->f[NK_FIELD_BASAL_MELT], *phi = m->f[NK_FIELD_PHI], *mg = m->f[NK_FIELD_GROUNDED_MELT];
    double *G = m->f[NK_FIELD_GH];
    double *theta = NULL;
    if (adv == 3) {   /* update_bounds_preserving_limiter!: θ per cell, halo filled as H's */
        theta = og_alloc(g, 1);
        for (int j = 1; j <= g->Ny; j++)
            for (int i = 1; i <= g->Nx; i++) {
                double c = AT(g, H, i, j), th = 1.0;
                for (int dir = 0; dir < 2; dir++) {
                    double pm = weno_face(g, H, dir, i, j, 0);
                    double pp = dir == 0 ? weno_face(g, H, 0, i + 1, j, 1) : weno_face(g, H, 1, i, j + 1, 1);
                    double pt = (c - WENO_OMEGA * pm - WENO_OMEGA * pp) / (1.0 - 2.0 * WENO_OMEGA);
                    double M = fmax(pt, fmax(pm, pp)), mn = fmin(pt, fmin(pm, pp));
                    th = fmin(th, fmin(limit_ratio(m->weno_bounds[1], c, M), limit_ratio(m->weno_bounds[0], c, mn)));
                }
                AT(g, theta, i, j) = th;
            }
        fill_halo(g, theta, C, C, 1);
    }
    for (int j = 1; j <= g->Ny; j++)
        for (int i = 1; i <= g->Nx; i++) {
            double u0 = AT(g, u, i, j), u1 = AT(g, u, i + 1, j), v0 = AT(g, v, i, j), v1 = AT(g, v, i, j + 1);
            double Fx0, Fx1, Fy0, Fy1;
            if (adv >= 2) {
                Fx0 = g->dy * weno_flux_H(g, H, theta, 0, i, j, u0);
                Fx1 = g->dy * weno_flux_H(g, H, theta, 0, i + 1, j, u1);
                Fy0 = g->dx * weno_flux_H(g, H, theta, 1, i, j, v0);
                Fy1 = g->dx * weno_flux_H(g, H, theta, 1, i, j + 1, v1);
            } else {
                Fx0 = g->dy * u0 * face_H(AT(g, H, i - 1, j), AT(g, H, i, j), u0, adv);
                Fx1 = g->dy * u1 * face_H(AT(g, H, i, j), AT(g, H, i + 1, j), u1, adv);
                Fy0 = g->dx * v0 * face_H(AT(g, H, i, j - 1), AT(g, H, i, j), v0, adv);
                Fy1 = g->dx * v1 * face_H(AT(g, H, i, j), AT(g, H, i, j + 1), v1, adv);
            }
            double div = (1.0 / (g->dx * g->dy)) * ((Fx1 - Fx0) + (Fy1 - Fy0));
            /* basal_melt (:51-52): (1 − φ) m with a grounding treatment; grounded_melt (:57-58): mᵍ */
            double melt = m->ph.grounding != NK_GROUNDING_NONE ? (1.0 - AT(g, phi, i, j)) * AT(g, mm, i, j) : AT(g, mm, i, j);
            double S = (AT(g, a, i, j) - melt) - (coupled ? AT(g, mg, i, j) : 0.0);
            AT(g, G, i, j) = S - div;
        }
    free(theta);
    fill_field_halo(m, NK_FIELD_GH);
}

/* _thickness_substep! (time_stepping.jl:65-72) */
EXPORT void nko_thickness_substep(nko_model *m, double dtau, int final_stage) {
    const og *g = &m->g;
    const double *Hm = m->f[NK_FIELD_HPREV], *G = m->f[NK_FIELD_GH];
    double *H = m->f[NK_FIELD_H], *Hc = m->f[NK_FIELD_HCLIP];
    for (int j = 1; j <= g->Ny; j++)
        for (int i = 1; i <= g->Nx; i++) {
            double Hs = AT(g, Hm, i, j) + dtau * AT(g, G, i, j);
            AT(g, H, i, j) = jmax(0.0, Hs);
            AT(g, Hc, i, j) = AT(g, Hc, i, j) + (final_stage ? jmax(0.0, -Hs) : 0.0);
        }
    fill_field_halo(m, NK_FIELD_H);
    fill_field_halo(m, NK_FIELD_HCLIP);
}

/* cache_current_fields! (time_stepping.jl:1-5) */
EXPORT void nko_cache_thickness(nko_model *m) {
    memcpy(m->f[NK_FIELD_HPREV], m->f[NK_FIELD_H], sizeof(double) * m->g.plane);
}

/* Oceananigans cell_advection_timescale: min over cells of 1/(|u|/Δx + |v|/Δy) */
EXPORT double nko_cell_advection_timescale(const nko_model *m) {
    const og *g = &m->g;
    const double *u = m->f[NK_FIELD_U], *v = m->f[NK_FIELD_V];
    double t = INFINITY;
    for (int j = 1; j <= g->Ny; j++)
        for (int i = 1; i <= g->Nx; i++) t = fmin(t, 1.0 / (fabs(AT(g, u, i, j)) / g->dx + fabs(AT(g, v, i, j)) / g->dy));
    return t;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Enthalpy step (src/Thermodynamics/column_energy_solve.jl, layer_transport.jl, strain_heating.jl)               */
/* ------------------------------------------------------------------------------------------------------------ */

/* EnthalpyParameters defaults not above (enthalpy_thermodynamics.jl:151-166) */
static const double E_K = 2.1, E_RATIO = 1e-3, E_DE = 1.0, E_HWB = 1e-3, E_HMIN = 1.0, E_WD = 0.01, E_RMAX = 1.58443823077064e-9;

static double unit_ramp(double x, double w) { return jclamp(x / w, 0.0, 1.0); }                 /* enthalpy_conversions.jl:69 */
static double col_thickness(double H) { return jmax(H, E_HMIN); }                              /* :166 */
/* layer_diffusivity (:180-188) at enthalpy E and pressure p */
static double layer_diffusivity(double E, double p) {
    double w = unit_ramp(E - pm_enthalpy(p) + E_DE, E_DE);
    double Kc = E_K / E_CI;
    return Kc * (1.0 - w + w * E_RATIO);
}
static double upwind_value(double u, double cm, double cp) { return u > 0.0 ? cm : cp; }        /* layer_transport.jl:5 */

/* compute_strain_heating! for SSA and DIVA (strain_heating.jl:8-56) */
EXPORT void nko_compute_strain_heating(nko_model *m) {
    const og *g = &m->g;
    const double *u = m->f[NK_FIELD_U], *v = m->f[NK_FIELD_V], *es = m->f[NK_FIELD_ETASIGMA];
    const double *ub = m->f[NK_FIELD_UB], *vb = m->f[NK_FIELD_VB], *bbu = m->f[NK_FIELD_BETABU], *bbv = m->f[NK_FIELD_BETABV];
    double *Phi = m->f[NK_FIELD_STRAIN_HEATING];
    int diva = m->ph.stress_balance == NK_DIVA;
    for (int k = 1; k <= m->Ns; k++)
        for (int j = 1; j <= g->Ny; j++)
            for (int i = 1; i <= g->Nx; i++) {
                size_t o = (size_t)(k - 1) * g->plane + IDX(g, i, j);
                double eta = es[o];
                double e2 = eff_strain2(g, i, j, u, v, m->ph.walls);
                double val = 4.0 * eta * e2;   /* _compute_membrane_heating! (:41-44) */
                if (diva) {                    /* _compute_shear_heating! (:48-56): ℑx, ℑy of (β u)² */
                    double s0 = AT(g, bbu, i, j) * AT(g, ub, i, j), s1 = AT(g, bbu, i + 1, j) * AT(g, ub, i + 1, j);
                    double t0 = AT(g, bbv, i, j) * AT(g, vb, i, j), t1 = AT(g, bbv, i, j + 1) * AT(g, vb, i, j + 1);
                    double tb2 = (s0 * s0 + s1 * s1) / 2.0 + (t0 * t0 + t1 * t1) / 2.0;
                    double shear = tb2 * m->w2[k] / m->dz[k] / (eta == 0.0 ? INFINITY : eta);
                    val = val + shear;
                }
                Phi[o] = val;
            }
    fill_field_halo(m, NK_FIELD_STRAIN_HEATING);
}

/* advance_enthalpy! (column_energy_solve.jl:48-84) with update_thermal_diagnostics! and Hⁿ = H */
/* effective_pressure with TillEffectivePressure (till_effective_pressure.jl:66-73) in _compute_effective_pressure!
 * (subgrid_grounded_fraction.jl:176-183) */
EXPORT void nko_till_effective_pressure(nko_model *m, const nk_till *t) {
    const og *g = &m->g;
    const double *H = m->f[NK_FIELD_H], *b = m->f[NK_FIELD_BED], *Hw = m->f[NK_FIELD_HW];
    double *N = m->f[NK_FIELD_N];
    double rig = m->ph.ice_density * m->ph.gravitational_acceleration, rwg = m->ph.ocean_density * m->ph.gravitational_acceleration;
    for (int j = 1; j <= g->Ny; j++)
        for (int i = 1; i <= g->Nx; i++) {
            double P0 = rig * AT(g, H, i, j);
            double No = jmax(P0 + rwg * jmin(AT(g, b, i, j), 0.0), 0.0);
            double N0 = t->reference_effective_pressure, d = t->overburden_fraction;
            double s = jclamp(AT(g, Hw, i, j) / t->maximum_water_thickness, 0.0, 1.0);
            double Pt = P0 > 0.0 ? P0 : N0 / d;
            double Nt = N0 * pow(d * Pt / N0, s) * pow(10.0, t->reference_void_ratio / t->compressibility * (1.0 - s));
            AT(g, N, i, j) = jmin(Nt, No);
        }
    fill_field_halo(m, NK_FIELD_N);
}

EXPORT void nko_energy_step(nko_model *m, double dt, int coupled, const nk_till *till) {
    const og *g = &m->g;
    const int Ns = m->Ns;
    const double rho = m->ph.ice_density, rg = m->ph.ice_density * m->ph.gravitational_acceleration, Az = g->dx * g->dy;
    const int grounding = m->ph.grounding != NK_GROUNDING_NONE, diva = m->ph.stress_balance == NK_DIVA;
    const double *H = m->f[NK_FIELD_H], *Hn = m->f[NK_FIELD_HN], *phi = m->f[NK_FIELD_PHI], *mm = m->f[NK_FIELD_BASAL_MELT];
    const double *G = m->f[NK_FIELD_GEOTHERMAL], *Ts = m->f[NK_FIELD_TSURF], *To = m->f[NK_FIELD_TOCEAN], *Phi = m->f[NK_FIELD_STRAIN_HEATING];
    const double *us = m->f[NK_FIELD_USIGMA], *vs = m->f[NK_FIELD_VSIGMA], *ub = m->f[NK_FIELD_UB], *vb = m->f[NK_FIELD_VB];
    const double *bu = diva ? m->f[NK_FIELD_BETABU] : m->f[NK_FIELD_BETAU], *bv = diva ? m->f[NK_FIELD_BETABV] : m->f[NK_FIELD_BETAV];
    double *E = m->f[NK_FIELD_ENTHALPY], *Hw = m->f[NK_FIELD_HW], *mg = m->f[NK_FIELD_GROUNDED_MELT];
    fill_field_halo(m, NK_FIELD_ENTHALPY);
    double *Enew = malloc(sizeof(double) * g->plane * (size_t)Ns);
    memcpy(Enew, E, sizeof(double) * g->plane * (size_t)Ns);
    double *W = malloc(sizeof(double) * (size_t)(Ns + 2)), *a = malloc(sizeof(double) * (size_t)(Ns + 1)), *b = malloc(sizeof(double) * (size_t)(Ns + 1));
    double *c = malloc(sizeof(double) * (size_t)(Ns + 1)), *f = malloc(sizeof(double) * (size_t)(Ns + 1)), *t = calloc((size_t)(Ns + 1), sizeof(double));
    double *x = malloc(sizeof(double) * (size_t)(Ns + 1));
#define EK(i, j, k) E[(size_t)((k) - 1) * g->plane + IDX(g, i, j)]
#define LP(i, j, k) (rg * AT(g, H, i, j) * (1.0 - m->zc[k]))   /* layer_pressure (enthalpy_conversions.jl:61-64) */
    for (int j = 1; j <= g->Ny; j++)
        for (int i = 1; i <= g->Nx; i++) {
            double Hc = AT(g, H, i, j), Ht = col_thickness(Hc), Htn = col_thickness(AT(g, Hn, i, j));
            int mu = Hc > E_HMIN;                                                   /* column_has_ice (:168) */
            double phg = grounding ? AT(g, phi, i, j) : 1.0;                        /* grounded_part (:119-120) */
            /* _compute_layer_volume_fluxes! (layer_transport.jl:34-44) */
            double dHdt = (Ht - Htn) / dt;
            double bm = grounding ? (1.0 - AT(g, phi, i, j)) * AT(g, mm, i, j) : AT(g, mm, i, j);
            W[1] = -bm - (coupled ? AT(g, mg, i, j) : 0.0);
            for (int k = 1; k <= Ns; k++) {
                const double *uk = us + (size_t)(k - 1) * g->plane, *vk = vs + (size_t)(k - 1) * g->plane;
                double xv[2], yv[2];
                for (int s = 0; s < 2; s++) {   /* x_volume_flux, y_volume_flux (:8-16) with wall faces carrying none */
                    double u = periph(g, i + s, j, F, C) ? 0.0 : AT(g, uk, i + s, j);
                    xv[s] = g->dy * u * upwind_value(u, AT(g, Hn, i + s - 1, j), AT(g, Hn, i + s, j));
                    double v = periph(g, i, j + s, C, F) ? 0.0 : AT(g, vk, i, j + s);
                    yv[s] = g->dx * v * upwind_value(v, AT(g, Hn, i, j + s - 1), AT(g, Hn, i, j + s));
                }
                double vfd = (xv[1] - xv[0] + yv[1] - yv[0]) / Az;             /* volume_flux_divergence (:20-23) */
                W[k + 1] = W[k] - m->dz[k] * (dHdt + vfd);
            }
            /* _compute_basal_state! (column_energy_solve.jl:130-147) */
            double pb = rg * Hc;                                                    /* basal_pressure (:176) */
            double D = layer_diffusivity(EK(i, j, 1), LP(i, j, 1)) / (Ht * m->zc[1]);
            double F_ = (AT(g, bu, i, j) * (AT(g, ub, i, j) * AT(g, ub, i, j)) + AT(g, bu, i + 1, j) * (AT(g, ub, i + 1, j) * AT(g, ub, i + 1, j)) +
                         AT(g, bv, i, j) * (AT(g, vb, i, j) * AT(g, vb, i, j)) + AT(g, bv, i, j + 1) * (AT(g, vb, i, j + 1) * AT(g, vb, i, j + 1))) / 2.0;
            double Epmb = pm_enthalpy(pb);
            double Eb = EK(i, j, 1) + (AT(g, G, i, j) + F_) / D;
            double r = jmax(unit_ramp(Eb - Epmb, E_DE), unit_ramp(AT(g, Hw, i, j), E_HWB));
            double sb = unit_ramp(EK(i, j, 1) - pm_enthalpy(LP(i, j, 1)) + E_DE, E_DE);
            double qg = phg * AT(g, G, i, j) + F_;
            /* _compute_column_system! (:185-226) */
            double Es = e_enthalpy(AT(g, Ts, i, j), 0.0, 0.0);                      /* surface_enthalpy (:164) */
            double chi = phg * r * (1.0 - sb);
            double Eo = e_enthalpy(AT(g, To, i, j), 0.0, pb);
            double Fb = rho * W[1];
            double Lam = D * (chi + 1.0 - phg) + jmax(-Fb, 0.0);
            double S = (1.0 - r) * qg + D * (chi * Epmb + (1.0 - phg) * Eo) + jmax(Fb, 0.0) * Epmb;
            double Fs = rho * W[Ns + 1];
            double Ks = layer_diffusivity(EK(i, j, Ns), LP(i, j, Ns)) / (Ht * (1.0 - m->zc[Ns]));   /* surface_conductance */
            double Ds = Ks + jmax(Fs, 0.0), Ss = (Ks + jmax(-Fs, 0.0)) * Es;
            for (int k = 1; k <= Ns; k++) {
                /* face_couplings at the faces k (below layer k) and k + 1 (above), interior faces only */
                double low[2] = {0.0, 0.0}, upp[2] = {0.0, 0.0};
                for (int s = 0; s < 2; s++) {
                    int kf = k + s;
                    if (kf < 2 || kf > Ns) continue;
                    double zf = m->zf[kf], zm = m->zc[kf - 1], zp = m->zc[kf];
                    double th = (zf - zm) / (zp - zm);
                    double Fk = rho * W[kf];
                    double Km = layer_diffusivity(EK(i, j, kf - 1), LP(i, j, kf - 1)), Kp = layer_diffusivity(EK(i, j, kf), LP(i, j, kf));
                    double cond = 1.0 / (Ht * ((zf - zm) / Km + (zp - zf) / Kp));    /* face_conductance (:150-158) */
                    double Dk = jmax(jmax(cond, jmax(Fk, 0.0) * th), jmax(-Fk, 0.0) * (1.0 - th));
                    low[s] = -Fk * (1.0 - th) - Dk;
                    upp[s] = Fk * th - Dk;
                }
                double dm = k == 1 ? Lam : -upp[0];
                double dp = k == Ns ? Ds : -low[1];
                double forcing = (k == 1 ? S : 0.0) + (k == Ns ? Ss : 0.0);
                /* enthalpy_flux_divergence (layer_transport.jl:25-28) with Hⁿ and the old enthalpy */
                const double *uk = us + (size_t)(k - 1) * g->plane, *vk = vs + (size_t)(k - 1) * g->plane;
                double xe[2], ye[2];
                for (int s = 0; s < 2; s++) {
                    double uraw = AT(g, uk, i + s, j), u = periph(g, i + s, j, F, C) ? 0.0 : uraw;
                    xe[s] = g->dy * u * upwind_value(u, AT(g, Hn, i + s - 1, j), AT(g, Hn, i + s, j)) * upwind_value(uraw, EK(i + s - 1, j, k), EK(i + s, j, k));
                    double vraw = AT(g, vk, i, j + s), v = periph(g, i, j + s, C, F) ? 0.0 : vraw;
                    ye[s] = g->dx * v * upwind_value(v, AT(g, Hn, i, j + s - 1), AT(g, Hn, i, j + s)) * upwind_value(vraw, EK(i, j + s - 1, k), EK(i, j + s, k));
                }
                double efd = (xe[1] - xe[0] + ye[1] - ye[0]) / Az;
                double transport = rho * m->dz[k] * efd;
                a[k] = mu ? low[1] : 0.0;
                c[k] = mu ? upp[1] : 0.0;
                b[k] = mu ? rho * Ht * m->dz[k] / dt + dm + dp : 1.0;
                f[k] = mu ? rho * Htn * m->dz[k] * EK(i, j, k) / dt - transport + Ht * m->dz[k] * Phi[(size_t)(k - 1) * g->plane + IDX(g, i, j)] + forcing : Es;
            }
            /* BatchedTridiagonalSolver (Oceananigans): a[k] couples x[k] into row k + 1, c[k] x[k + 1] into row k */
            for (int k = 1; k <= Ns; k++) x[k] = EK(i, j, k);
            double beta = b[1];
            x[1] = f[1] / beta;
            for (int k = 2; k <= Ns; k++) {
                t[k] = c[k - 1] / beta;
                beta = b[k] - a[k - 1] * t[k];
                if (!(fabs(beta) > 10.0 * 2.220446049250313e-16)) break;
                x[k] = (f[k] - a[k - 1] * x[k - 1]) / beta;
            }
            for (int k = Ns - 1; k >= 1; k--) x[k] = x[k] - t[k + 1] * x[k + 1];
            /* _update_basal_water! (:238-267), no till */
            double q = D * (Epmb - x[1]);
            double melt = r * (qg - phg * (1.0 - sb) * q) / (rho * E_L);
            double drained = 0.0;
            for (int k = 1; k <= Ns; k++) {
                double w = e_water(x[k], LP(i, j, k));
                double xr = w / E_WD;   /* drainage_rate (:230-234) */
                double shape = jclamp(xr - 1.0, 0.0, 1.0) / 10.0 + 9.0 * jclamp(xr - 2.0, 0.0, 1.0) / 10.0;
                double dw = jmin(dt * (E_RMAX * shape), jmax(w - E_WD, 0.0));
                dw = mu ? dw : 0.0;
                x[k] = x[k] - E_L * dw;
                drained = drained + Ht * m->dz[k] * dw;
            }
            double Hw0 = AT(g, Hw, i, j);
            double Hstar = Hw0 + dt * (mu ? melt : 0.0) + drained;
            double Hplus = jmax(0.0, Hstar), missing = Hplus - Hstar;
            AT(g, mg, i, j) = (Hplus - Hw0) / dt;
            /* drain_till (till_effective_pressure.jl:77-78) */
            AT(g, Hw, i, j) = till ? jclamp(Hplus - dt * till->drainage_rate, 0.0, till->maximum_water_thickness) : Hplus;
            x[1] = x[1] - E_L * missing / (Ht * m->dz[1]);
            for (int k = 1; k <= Ns; k++) Enew[(size_t)(k - 1) * g->plane + IDX(g, i, j)] = x[k];
        }
#undef EK
#undef LP
    memcpy(E, Enew, sizeof(double) * g->plane * (size_t)Ns);
    free(Enew); free(W); free(a); free(b); free(c); free(f); free(t); free(x);
    fill_field_halo(m, NK_FIELD_ENTHALPY);
    fill_field_halo(m, NK_FIELD_HW);
    fill_field_halo(m, NK_FIELD_GROUNDED_MELT);
    /* update_thermal_diagnostics!: T and ω, A when the rate factor is on the σ grid */
    double *T = m->f[NK_FIELD_TEMPERATURE], *Wf = m->f[NK_FIELD_WATER_FRACTION], *A = m->f[NK_FIELD_A];
    for (int k = 1; k <= Ns; k++)
        for (int j = 1; j <= g->Ny; j++)
            for (int i = 1; i <= g->Nx; i++) {
                size_t o = (size_t)(k - 1) * g->plane + IDX(g, i, j);
                double p = rg * AT(g, H, i, j) * (1.0 - m->zc[k]);
                T[o] = e_temperature(E[o], p);
                Wf[o] = e_water(E[o], p);
                if (m->ph.rate_factor_depth == NK_RATE_FACTOR_3D) A[o] = nko_arrhenius_rate_factor(T[o] - pm_temperature(p) + E_T0, Wf[o]);
            }
    fill_field_halo(m, NK_FIELD_TEMPERATURE);
    fill_field_halo(m, NK_FIELD_WATER_FRACTION);
    if (m->ph.rate_factor_depth == NK_RATE_FACTOR_3D) fill_field_halo(m, NK_FIELD_A);
    memcpy(m->f[NK_FIELD_HN], m->f[NK_FIELD_H], sizeof(double) * g->plane);   /* parent(Hⁿ) .= parent(H) */
}

/* σ-grid metrics for tests */
EXPORT void nko_sigma_metrics(const nko_model *m, double *zf, double *zc, double *dz, double *w2) {
    for (int k = 1; k <= m->Ns + 1; k++) zf[k - 1] = m->zf[k];
    for (int k = 1; k <= m->Ns; k++) { zc[k - 1] = m->zc[k]; dz[k - 1] = m->dz[k]; w2[k - 1] = m->w2[k]; }
}
