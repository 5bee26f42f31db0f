// nk_solver.hip — host side of libnunatak_hip.so: the handle, its device fields, and the drivers of the
// reference's velocity-solve functions (update_velocities!, picard_solve!, solve!, precondition!, ssa_operator!)
// over the gfx950 kernels of nk_device.hpp / nk_physics.hpp. C ABI in include/nunatak_hip.h.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nunatak_hip.h"
#include "nk_device.hpp"
#include "nk_physics.hpp"
#include "nk_adjoint.hpp"
#include "nk_galerkin.hpp"
#include "nk_transport.hpp"
#include "nk_thermal.hpp"
#include "nk_internal.hpp"

using namespace nk;

namespace {

struct LocInfo {
    int lx, ly, sigma;  // 1 = Face
};
const LocInfo LOC[NK_FIELD_COUNT] = {
    {1, 0, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 0},
    {1, 0, 0}, {0, 1, 0}, {0, 0, 0}, {1, 1, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 1}, {1, 0, 0}, {0, 1, 0},
    {1, 0, 0}, {0, 1, 0}, {1, 0, 1}, {0, 1, 1}, {0, 0, 2}, {0, 0, 0}, {0, 0, 1}, {1, 0, 0}, {0, 1, 0}, {1, 0, 0},
    {0, 1, 0}, {1, 0, 0}, {0, 1, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 2}, {1, 0, 1}, {0, 1, 1}, {0, 0, 1},
    {0, 0, 1}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0},
    {0, 0, 0}, {0, 0, 0}, {0, 0, 1}, {0, 0, 0}, {0, 0, 0},
};
static_assert(sizeof(LOC) / sizeof(LOC[0]) == NK_FIELD_COUNT, "one location per field id");

template <typename T> struct Vecs {
    T *x0u = nullptr, *x0v = nullptr, *x1u = nullptr, *x1v = nullptr, *bu = nullptr, *bv = nullptr, *qu = nullptr,
      *qv = nullptr, *t = nullptr;
};

struct Level {
    KG g{};
    int cx = 0, cy = 0;  // halved relative to the next finer level
    double *eta = nullptr, *etaff = nullptr, *bu = nullptr, *bv = nullptr;
    float *eta32 = nullptr, *etaff32 = nullptr, *bu32 = nullptr, *bv32 = nullptr;
    uint8_t *act = nullptr;
    Vecs<double> v64;
    Vecs<float> v32;
    double *gs64 = nullptr;  // Galerkin stencil + 1/ℓ1 planes (GS_ALLOC) at the V-cycle's precision
    float *gs32 = nullptr;
    void *gcst = nullptr;    // compact copy of the Galerkin stencil (k_g_compact) for levels the LDS tail may hold
    void *gpk = nullptr;     // symmetric packed copy of the Galerkin stencil (k_g_pack) for the launched sweeps
};

// Levels small enough for the Galerkin LDS tail keep a compact copy of their stencil (shape-only rule, so the
// device-memory estimate can mirror it).
constexpr size_t GTAIL_BUDGET = 150 * 1024;
// Points the Galerkin LDS tail takes on by default beyond the two coarsest levels: 16² and below. From 32² the one
// workgroup spends longer on that level than four launched kernels do (1024² mixed solve, same-box A/B: 1390.6 ms with
// the tail from 32², 1374.2 ms from 16², 1515.8 ms from 8²).
constexpr size_t GTAIL_POINTS = 640;
bool wants_gcst(const KG &g, int l, size_t es) {
    return l >= 1 && (size_t)(g.Nx + 1) * (size_t)(g.Ny + 1) * (8 * es + 1) <= GTAIL_BUDGET;
}

template <typename T> struct Work {  // PCG workspace of one precision on the finest grid
    T *ru, *rv, *zu, *zv, *z2u, *z2v, *pu, *pv, *qu, *qv, *du, *dv;
};

double jmax_h(double x, double y) { double d = x - y; return (std::isnan(x) || std::isnan(y)) ? d : (std::signbit(d) ? y : x); }

}  // namespace

struct nk_solver {
    nk_grid grid{};
    nk_physics ph{};
    nk_solver_options op{};
    nk_solver_options op0{};   // options at nk_create: they fix the hierarchy, the precision and the tails
    KG g{};
    long plane = 0;  // elements per 2D plane
    int Ns = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double *f[NK_FIELD_COUNT] = {};
    uint8_t *cls_u = nullptr, *cls_v = nullptr;
    double *sig = nullptr;
    Work<double> w64{};
    Work<float> w32{};
    double *xmu = nullptr, *xmv = nullptr;
    // adjoint work planes: q̄, ȳ, λ⁺, r (u and v), η̄H̄ (centres), ε̇xy adjoint (corners)
    double *ad_qu = nullptr, *ad_qv = nullptr, *ad_yu = nullptr, *ad_yv = nullptr, *ad_lu = nullptr, *ad_lv = nullptr,
           *ad_ru = nullptr, *ad_rv = nullptr, *ad_eta = nullptr, *ad_exy = nullptr;
    // DIVA adjoint: βᵇ̄ (total), uᵇ̄, F̄₂ of this step's state; the sweep state (st_*) and the step's outputs for the
    // lagged state (out_*), swapped after each fixed-point iteration; ησ̄ planes (Nσ layers)
    double *ad_bbu = nullptr, *ad_bbv = nullptr, *ad_ubb = nullptr, *ad_vbb = nullptr, *ad_F2n = nullptr;
    double *st_bbu = nullptr, *st_bbv = nullptr, *st_F2 = nullptr, *out_bbu = nullptr, *out_bbv = nullptr, *out_F2 = nullptr;
    double *st_es = nullptr, *out_es = nullptr;
    // shadows the cost gives the DIVA state (through the layer velocities) and the resulting x̄ of the adjoint
    double *cs_bbu = nullptr, *cs_bbv = nullptr, *cs_F2 = nullptr, *cs_es = nullptr, *ad_xbu = nullptr, *ad_xbv = nullptr;
    // unrolled adjoint of FixedIterations Picard: the initial guess of the last such Picard solve (recorded by picard()),
    // and the tape of its iterates x_0 .. x_n (2 planes each; DIVA adds the lagged state βᵇ, F₂, ησ of each step) plus
    // λ_k, allocated by the first nk_adjoint
    double *x0u = nullptr, *x0v = nullptr;
    double *s0_bbu = nullptr, *s0_bbv = nullptr, *s0_F2 = nullptr, *s0_es = nullptr;   // DIVA: the lagged state it started from
    bool x0_valid = false;
    double *tape = nullptr;
    long tape_planes = 0;
    // mixed precision: while set, the finest level's last FP32 sweep also writes z64 and sums r64·z64
    const double *fz_r64u = nullptr, *fz_r64v = nullptr;
    double *fz_z64u = nullptr, *fz_z64v = nullptr;
    float *rl1u32 = nullptr, *rl1v32 = nullptr;   // mixed multigrid: 1/ℓ1 of the finest rows in FP32 (k_l1_recip)
    float *b32u = nullptr, *b32v = nullptr;       // iterative refinement: the scaled FP32 defect, right-hand side of the inner PCG
    double refine_inner = 1e-4;                   // ... and the inner PCG's relative tolerance (NK_REFINE_INNER_RELTOL: tuning)
    bool fuse_pre = true;   // first two sweeps from zero in one kernel (NK_NO_PRE_FUSE=1 at creation: separate kernels)
    bool fuse_pro = true;   // packed Galerkin levels: prolongation inside the first post-sweep (NK_NO_PRO_FUSE=1: k_prolongate)
    long pro_fuse_points = 256L * 256;   // ... on levels of at most this many points (NK_PRO_FUSE_POINTS: tuning)
    long latency_points = 256L * 256;    // packed Galerkin levels of at most this many points launch the latency-bound
                                         // forms k_gp_sweep_tl, k_gp_residual_tl, k_gp_sweep_pre_tl (NK_LATENCY_POINTS: tuning)
    uint16_t *nm16 = nullptr;   // neighbour masks of the finest points (k_nmask), rebuilt with the finest activity flags
    bool use_z32 = true;        // mixed: the direction update reads the FP32 z (NK_NO_Z32=1: FP64 z via z64)
    bool apply_tile = true;     // FP64 finest operator applies and sweeps, and the packed Galerkin sweeps and residuals,
                                // stage the iterate through LDS (NK_NO_ATILE=1: k_apply_m, k_sweep_m, k_sweep_m_dot,
                                // k_gp_sweep, k_gp_sweep_pre, k_gp_residual)
    bool use_gpack = true;      // launched Galerkin sweeps read the packed stencil (NK_NO_GPACK=1 at creation: full)
    bool fuse_rr = true;        // packed Galerkin levels of up to rr_points points: residual and restriction in one
                                // launch (k_gp_residual_restrict; NK_NO_RR_FUSE=1 at creation: a residual kernel + k_restrict)
    long rr_points = 256L * 256;   // (NK_RR_POINTS: tuning)
    bool rap_split = true;      // triple products into levels of at most rap_split_points points split each support
                                // point's couplings between two lanes by column component (k_g_rap_split;
                                // NK_NO_RAP_SPLIT=1 at creation: one lane per support point, k_g_rap)
    bool rap_compact = true;    // the products also write the tail levels' compact stencil copies (NK_NO_RAP_COMPACT=1
                                // at creation: k_g_compact copies them after the products)
    long rap_split_points = 256L * 256;   // (NK_RAP_SPLIT_POINTS: tuning; into 512² from 1024² the split measured
                                          // 491 against 444 µs, into 256² and below 335 against 344 µs down to 30 against 53)
    bool fin_pattern = true;    // the first triple product reads only the finest stencil's 9 couplings per row
                                // (NK_NO_FIN_PATTERN=1 at creation: all 31, the other 22 zeros skipped as they are read)
    bool ginv_regs = true;      // the coarsest Galerkin inverse eliminates in registers (NK_NO_GINV_REGS=1 at creation:
                                // in LDS, element by element)
    bool use_nmask = true;      // finest-level operator applies and sweeps read nm16 (NK_NO_NMASK=1 at creation: the
                                // flag-byte kernels k_apply, k_sweep, k_sweep_dot, k_sweep_z64)
    // Multigrid PCG on grids of at most fold_blocks workgroups (NK_FOLD_BLOCKS; NK_NO_FOLD=1 at creation: never): the α
    // and β finishers run inside k_update_xr and k_update_p / k_update_p32 (every block sums the slots itself) instead
    // of as k_finish launches — two launches fewer per CG iteration. Measured (mixed solve, 3 Picard iterations): 4 %
    // faster at 512² (1,024 slots); at 1024² (4,096 slots) within 1 %, with k_update_xr at 24.7 instead of 19.9 µs (every
    // block re-reads 32 KB of slots), so the largest grids keep the launches.
    bool fold = false;
    long fold_blocks = 1024;
    int fold_pending = 0, fold_nb = 0;   // slots of the r·z reduction the preconditioner left for the direction update
    // enthalpy step workspace, allocated by the first nk_energy_step: face volume fluxes (Nσ + 1 planes), the column
    // system a, b, c, f and the solver's t (Nσ planes each), the basal state rᵇ, sᵇ, Dᵇ, qᵍ (4 planes)
    double *en_W = nullptr, *en_a = nullptr, *en_b = nullptr, *en_c = nullptr, *en_f = nullptr, *en_t = nullptr, *en_bas = nullptr;
    // bounds-preserving WENO thickness advection: the per-cell limiter θ (allocated by the first WENO tendency) and the
    // bounds (the reference's default WENO(order = 5, bounds = (0, 1e5)))
    double *weno_theta = nullptr;
    double thick_lower = 0.0, thick_upper = 1e5;
    std::vector<Level> lev;  // lev[0] = the finest grid
    double *partials = nullptr;
    double *partials_xr = nullptr;   // k_update_xr's r·r slots: apart from the p·q slots it may read (folded α)
    Scal *S = nullptr;
    Scal *S_host = nullptr;
    std::vector<void *> allocs;
    size_t device_bytes = 0;
    std::string err;
    bool flags_valid = false;
    bool verbose = false;  // NK_VERBOSE=1: one progress line per Picard iteration on stderr
    int tail_start[2] = {-1, -1};           // first multigrid level run by the LDS tail kernel (FP64, FP32); −1: none
    size_t tail_bytes[2] = {0, 0};
    hipGraphExec_t chunk_graph = nullptr;  // captured PCG chunk (check_interval iterations)
    int chunk_graph_key = -1;
    const void *chunk_graph_x = nullptr;  // solution buffer the captured chunk updates
    int *gerr = nullptr;                   // Galerkin products: couplings that fell outside the stencil pattern
    int gtail_start = -1;                  // first level of the Galerkin LDS tail (-1: none)
    void *cinv = nullptr;                  // FP64 dense inverse of the coarsest Galerkin operator (both precisions), or null
    // opt-in event timing of one finest-level kernel (nk_solver_options.profile_kernel)
    std::vector<hipEvent_t> prof_events;
    size_t prof_used = 0;
    double prof_ms = 0.0;
    long prof_launches = 0, prof_points = 0;
    int prof_precision = 0;
};

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            s->err = std::string(#expr) + ": " + hipGetErrorString(e_);                                \
            return NK_DEVICE_ERROR;                                                                    \
        }                                                                                              \
    } while (0)

#define CHK(expr)                               \
    do {                                        \
        nk_status st_ = (expr);                 \
        if (st_ != NK_OK) return st_;           \
    } while (0)

namespace {

// Serialises graph capture against device allocation and release across the handles of one process.
std::mutex g_device_mu;

// Message of the last failed handle-free call on this thread (nk_last_error(NULL)).
thread_local std::string g_free_err = "null handle";

KG make_kg(int Nx, int Ny, int tx, int ty, double dx, double dy) {
    KG g;
    g.Nx = Nx; g.Ny = Ny; g.tx = tx; g.ty = ty; g.dx = dx; g.dy = dy;
    g.rdx = 1.0 / dx; g.rdy = 1.0 / dy;   // correctly rounded: the device forms a / dx from them (qd with Dv)
    g.pitch = ((Nx + 1) + 63) / 64 * 64;
    return g;
}
long plane_of(const KG &g) { return (long)g.pitch * (long)(g.Ny + 1); }

dim3 blocks_for(int ex, int ey) { return dim3((ex + BX - 1) / BX, (ey + BY - 1) / BY); }
dim3 blocks_for_rows(int ex, int ey, int rows) { return dim3((ex + BX - 1) / BX, (ey + BY * rows - 1) / (BY * rows)); }
const dim3 THREADS(BX, BY);

// Launch a stencil kernel instantiated for the grid's topology (Bounded = 1 / Periodic = 0 along x and y), so the
// index maps compile to straight-line code.
#define NK_LAUNCH_TOPO_EV(G, KERNEL, T, GRID, BLOCK, SHM, STREAM, E0, E1, ...)                                 \
    do {                                                                                                      \
        if ((G).tx) {                                                                                         \
            if ((G).ty) hipExtLaunchKernelGGL(KERNEL<T, 1, 1>, GRID, BLOCK, SHM, STREAM, E0, E1, 0, __VA_ARGS__); \
            else hipExtLaunchKernelGGL(KERNEL<T, 1, 0>, GRID, BLOCK, SHM, STREAM, E0, E1, 0, __VA_ARGS__);        \
        } else {                                                                                              \
            if ((G).ty) hipExtLaunchKernelGGL(KERNEL<T, 0, 1>, GRID, BLOCK, SHM, STREAM, E0, E1, 0, __VA_ARGS__); \
            else hipExtLaunchKernelGGL(KERNEL<T, 0, 0>, GRID, BLOCK, SHM, STREAM, E0, E1, 0, __VA_ARGS__);        \
        }                                                                                                     \
    } while (0)

#define NK_LAUNCH_TOPO(G, KERNEL, T, GRID, BLOCK, SHM, STREAM, ...)                                  \
    do {                                                                                             \
        if ((G).tx) {                                                                                \
            if ((G).ty) KERNEL<T, 1, 1><<<GRID, BLOCK, SHM, STREAM>>>(__VA_ARGS__);                   \
            else KERNEL<T, 1, 0><<<GRID, BLOCK, SHM, STREAM>>>(__VA_ARGS__);                          \
        } else {                                                                                     \
            if ((G).ty) KERNEL<T, 0, 1><<<GRID, BLOCK, SHM, STREAM>>>(__VA_ARGS__);                   \
            else KERNEL<T, 0, 0><<<GRID, BLOCK, SHM, STREAM>>>(__VA_ARGS__);                          \
        }                                                                                            \
    } while (0)

#define NK_LAUNCH_TOPO2(G, KERNEL, GRID, BLOCK, SHM, STREAM, ...)                                    \
    do {                                                                                             \
        if ((G).tx) {                                                                                \
            if ((G).ty) KERNEL<1, 1><<<GRID, BLOCK, SHM, STREAM>>>(__VA_ARGS__);                      \
            else KERNEL<1, 0><<<GRID, BLOCK, SHM, STREAM>>>(__VA_ARGS__);                             \
        } else {                                                                                     \
            if ((G).ty) KERNEL<0, 1><<<GRID, BLOCK, SHM, STREAM>>>(__VA_ARGS__);                      \
            else KERNEL<0, 0><<<GRID, BLOCK, SHM, STREAM>>>(__VA_ARGS__);                             \
        }                                                                                            \
    } while (0)

nk_status alloc_bytes(nk_solver *s, void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        s->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        return NK_OUT_OF_MEMORY;
    }
    s->allocs.push_back(*p);
    s->device_bytes += bytes;
    e = hipMemsetAsync(*p, 0, bytes, s->stream);
    if (e != hipSuccess) { s->err = "hipMemset failed"; return NK_DEVICE_ERROR; }
    return NK_OK;
}
template <typename T> nk_status alloc_plane(nk_solver *s, T **p, const KG &g, int nz = 1) {
    return alloc_bytes(s, (void **)p, sizeof(T) * (size_t)plane_of(g) * (size_t)nz);
}

int coarsening(int N, int H, int minimum) { int m = minimum > H ? minimum : H; return (N % 2 == 0) && (N / 2 >= m); }

// Julia Float64^Int is correctly rounded up to its compensated power; long double reproduces that.
double jpow_int(double x, int n) { long double r = 1.0L; for (int k = 0; k < n; k++) r *= (long double)x; return (double)r; }
double sigma_face(int k0, int Ns, double r) {
    double q = (double)k0 / (double)Ns;
    if (r == std::floor(r) && r >= 0 && r <= 16) return jpow_int(q, (int)r);
    return std::pow(q, r);
}
double power_integral(double zm, double zp, int p) { return (jpow_int(1.0 - zm, p + 1) - jpow_int(1.0 - zp, p + 1)) / (double)(p + 1); }

PhysK physk(const nk_solver *s) {
    const nk_physics &ph = s->ph;
    PhysK p;
    p.ri = ph.ice_density; p.rw = ph.ocean_density; p.gr = ph.gravitational_acceleration;
    p.ri_rw = ph.ice_density / ph.ocean_density;
    p.rw_ri = ph.ocean_density / ph.ice_density;
    p.rig = ph.ice_density * ph.gravitational_acceleration;
    p.rwg = ph.ocean_density * ph.gravitational_acceleration;
    p.grounding = ph.grounding; p.masking = ph.masking;
    p.hmin = ph.mask_minimum_thickness; p.htr = ph.mask_transition_thickness;
    p.n = ph.glen_exponent; p.e0 = ph.strain_rate_regularization;
    p.inv_n = -1.0 / ph.glen_exponent;
    p.exp_e = (1.0 - ph.glen_exponent) / (2.0 * ph.glen_exponent);
    p.A3d = ph.rate_factor_depth == NK_RATE_FACTOR_3D; p.friction = ph.friction;
    p.m = ph.friction_exponent; p.alpha2 = ph.coulomb_coefficient; p.u0 = ph.velocity_regularization;
    p.walls = ph.walls; p.Ns = s->Ns;
    p.wf = ph.walls == NK_NO_SLIP ? 2.0 : 0.0;
    return p;
}

Sig sig_tables(const nk_solver *s) {
    int Ns = s->Ns;
    Sig t;
    t.zf = s->sig;
    t.zc = s->sig + (Ns + 1);
    t.dz = t.zc + Ns;
    t.w2 = t.dz + Ns;
    t.pw = t.w2 + Ns;
    return t;
}

Red red(nk_solver *s, int fin) { Red r; r.partials = s->partials; r.S = s->S; r.fin = fin; return r; }

template <typename T> Coef<T> coef(const Level &L);
template <> Coef<double> coef<double>(const Level &L) { return Coef<double>{L.eta, L.etaff, L.bu, L.bv, L.act}; }
template <> Coef<float> coef<float>(const Level &L) { return Coef<float>{L.eta32, L.etaff32, L.bu32, L.bv32, L.act}; }
template <typename T> Vecs<T> &vecs(Level &L);
template <> Vecs<double> &vecs<double>(Level &L) { return L.v64; }
template <> Vecs<float> &vecs<float>(Level &L) { return L.v32; }
template <typename T> T *gs(Level &L);
template <> double *gs<double>(Level &L) { return L.gs64; }
template <> float *gs<float>(Level &L) { return L.gs32; }
template <typename T> Work<T> &work(nk_solver *s);
template <> Work<double> &work<double>(nk_solver *s) { return s->w64; }
template <> Work<float> &work<float>(nk_solver *s) { return s->w32; }

// Event bracket around one launch of the profiled kernel class on the finest grid.
// The launch inside a ProfScope goes through hipExtLaunchKernelGGL with the scope's event pair, which the runtime
// records at the kernel's start and end (null events when the scope is inactive: a plain launch).
struct ProfScope {
    nk_solver *s;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ProfScope(nk_solver *s_, int kind, int level, int precision_bytes) : s(s_) {
        if (s->op.profile_kernel != kind || kind == NK_PROFILE_NONE) return;
        if (level != (kind == NK_PROFILE_GALERKIN_SWEEP ? 1 : 0)) return;
        if (s->prof_used + 2 > s->prof_events.size()) {
            for (int k = 0; k < 256; k++) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return;
                s->prof_events.push_back(e);
            }
        }
        e0 = s->prof_events[s->prof_used];
        e1 = s->prof_events[s->prof_used + 1];
        s->prof_used += 2;
        s->prof_launches += 1;
        s->prof_points += (long)s->lev[level].g.Nx * s->lev[level].g.Ny;
        s->prof_precision = precision_bytes;
    }
};

// Profiling with chunks of PCG iterations: the launches of iterations after the one that stopped the solve return at
// once (stop flag), so their event pairs are dropped before the sum. A chunk records the same number of pairs per
// iteration; `ran` iterations counted by the device ran their operator, and all but the last one (when it stopped
// the solve) ran their V-cycle.
void prof_trim_chunk(nk_solver *s, int chunk, int ran, bool stopped) {
    const size_t pairs = s->prof_used / 2;
    if (s->op.profile_kernel == NK_PROFILE_NONE || pairs == 0 || pairs % (size_t)chunk != 0) return;
    const size_t per = pairs / (size_t)chunk;
    const int real = s->op.profile_kernel == NK_PROFILE_OPERATOR ? ran : ran - (stopped ? 1 : 0);
    const size_t keep = per * (size_t)(real < 0 ? 0 : (real > chunk ? chunk : real));
    const long dropped = (long)(pairs - keep);
    if (dropped <= 0) return;
    const long per_launch_points = s->prof_launches > 0 ? s->prof_points / s->prof_launches : 0;
    s->prof_launches -= dropped;
    s->prof_points -= dropped * per_launch_points;
    s->prof_used = 2 * keep;
}

// Sum the recorded event pairs (after the stream has been synchronised).
void prof_harvest(nk_solver *s) {
    for (size_t k = 0; k + 1 < s->prof_used; k += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, s->prof_events[k], s->prof_events[k + 1]) == hipSuccess) s->prof_ms += ms;
    }
    s->prof_used = 0;
}

nk_status launch_check(nk_solver *s) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { s->err = std::string("kernel launch: ") + hipGetErrorString(e); return NK_DEVICE_ERROR; }
    return NK_OK;
}

// Finish a reduction of nv values over the blocks of grid B (after the kernel that wrote the partials).
nk_status finish_red(nk_solver *s, dim3 B, int nv, int fin, const Scal *gate, const double *partials = nullptr) {
    k_finish<<<1, THREADS, 0, s->stream>>>(partials ? partials : s->partials, (int)(B.x * B.y), nv, s->S, fin, gate);
    return launch_check(s);
}

// The PCG's β reduction at the end of the preconditioner: a k_finish launch, or — when the PCG folds it into the
// direction update (fold_pending) — only the slot count for that kernel (r·z is the reduction's first value).
nk_status finish_beta(nk_solver *s, dim3 B, int nv, const Scal *gate) {
    if (s->fold_pending) {
        s->fold_nb = (int)(B.x * B.y);
        s->fold_pending = 0;
        return NK_OK;
    }
    return finish_red(s, B, nv, FIN_BETA, gate);
}

// ---- coefficient bookkeeping ------------------------------------------------------------------------------------

// η̄Hᶠᶠ (optional) and activity flags of level l from its η̄H
nk_status corners_flags(nk_solver *s, int l, bool write_corners) {
    Level &L = s->lev[l];
    k_corners_flags<<<blocks_for(L.g.Nx + 1, L.g.Ny + 1), THREADS, 0, s->stream>>>(L.g, L.eta, write_corners ? L.etaff : nullptr, L.act);
    CHK(launch_check(s));
    if (l == 0 && s->nm16) {
        NK_LAUNCH_TOPO2(L.g, k_nmask, blocks_for(L.g.Nx, L.g.Ny), THREADS, 0, s->stream, L.g, L.act, s->nm16);
        CHK(launch_check(s));
    }
    return NK_OK;
}

// q = A p, or r = b − A p with b, at the finest level: through the neighbour masks (k_apply_m) unless NK_NO_NMASK.
template <typename T>
void launch_apply0(nk_solver *s, const Coef<T> &c, T wf, const T *pu, const T *pv, T *qu, T *qv, const T *bu, const T *bv, Red r,
                   int nred, const Scal *gate, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const KG &g = s->lev[0].g;
    const dim3 B = blocks_for(g.Nx, g.Ny);
    if (s->nm16 && s->use_nmask && s->apply_tile && sizeof(T) == 8)
        NK_LAUNCH_TOPO_EV(g, k_apply_mt, T, B, THREADS, 0, s->stream, e0, e1, g, pu, pv, c, (const uint16_t *)s->nm16, wf, qu, qv, bu, bv,
                          r, nred, gate);
    else if (s->nm16 && s->use_nmask)
        NK_LAUNCH_TOPO_EV(g, k_apply_m, T, B, THREADS, 0, s->stream, e0, e1, g, pu, pv, c, (const uint16_t *)s->nm16, wf, qu, qv, bu, bv,
                          r, nred, gate);
    else
        NK_LAUNCH_TOPO_EV(g, k_apply, T, B, THREADS, 0, s->stream, e0, e1, g, pu, pv, c, wf, qu, qv, bu, bv, r, nred, gate);
}

// update_preconditioner!: coarse coefficients of every level (multigrid) and FP32 copies (mixed precision)
// Galerkin coarse operators: the finest operator as a stencil, then A_{l+1} = Pᵀ A_l P level by level.
template <typename T> nk_status galerkin_build(nk_solver *s) {
    Level &L0 = s->lev[0];
    const T wf = (T)(s->ph.walls == NK_NO_SLIP ? 2.0 : 0.0);
    NK_LAUNCH_TOPO(L0.g, k_g_assemble, T, blocks_for(L0.g.Nx, L0.g.Ny), THREADS, 0, s->stream, L0.g, coef<T>(L0), wf, gs<T>(L0));
    CHK(launch_check(s));
    HIPCHK(hipMemsetAsync(s->gerr, 0, sizeof(int), s->stream));
    for (size_t l = 1; l < s->lev.size(); l++) {
        Level &F = s->lev[l - 1], &C = s->lev[l];
        const bool split = s->rap_split && (long)C.g.Nx * C.g.Ny <= s->rap_split_points;
        const int npb = split ? GRAP_POINTS / 2 : GRAP_POINTS, fin = (l == 1 && s->fin_pattern) ? 1 : 0;
        const dim3 Br((C.g.Nx * C.g.Ny + npb - 1) / npb);
        T *cc = s->rap_compact ? (T *)C.gcst : nullptr;   // the tail's compact copy, written by the product itself
        if (split)
            NK_LAUNCH_TOPO(C.g, k_g_rap_split, T, Br, THREADS, 0, s->stream, C.g, F.g, C.cx, C.cy, C.act, F.act, gs<T>(F),
                           gs<T>(C), s->gerr, fin, cc);
        else
            NK_LAUNCH_TOPO(C.g, k_g_rap, T, Br, THREADS, 0, s->stream, C.g, F.g, C.cx, C.cy, C.act, F.act, gs<T>(F),
                           gs<T>(C), s->gerr, fin, cc);
        CHK(launch_check(s));
    }
    if (s->cinv) {
        Level &C = s->lev.back();
        const int n = 2 * C.g.Nx * C.g.Ny;
        const size_t shm = n <= GDENSE_LDS ? sizeof(double) * (size_t)n * (size_t)n : 0;
        NK_LAUNCH_TOPO(C.g, k_g_inverse, T, 1, TAIL_THREADS, shm, s->stream, C.g, gs<T>(C), (double *)s->cinv, s->ginv_regs ? 1 : 0);
        CHK(launch_check(s));
    }
    for (size_t l = 1; l < s->lev.size(); l++) {
        Level &L = s->lev[l];
        if (!L.gpk) continue;
        k_g_pack<T><<<blocks_for(L.g.Nx, L.g.Ny), THREADS, 0, s->stream>>>(L.g, gs<T>(L), (T *)L.gpk);
        CHK(launch_check(s));
    }
    for (size_t l = 1; l < s->lev.size() && !s->rap_compact; l++) {
        Level &L = s->lev[l];
        if (!L.gcst) continue;
        k_g_compact<T><<<blocks_for(L.g.Nx + 1, L.g.Ny + 1), THREADS, 0, s->stream>>>(L.g, gs<T>(L), (T *)L.gcst);
        CHK(launch_check(s));
    }
    int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, s->gerr, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (bad) {
        s->err = "Galerkin product: " + std::to_string(bad) + " couplings outside the coarse stencil pattern";
        return NK_DEVICE_ERROR;
    }
    return NK_OK;
}

nk_status update_preconditioner(nk_solver *s) {
    if (s->op.preconditioner == NK_PRECONDITIONER_MULTIGRID) {
        for (size_t l = 1; l < s->lev.size(); l++) {
            Level &F = s->lev[l - 1], &C = s->lev[l];
            k_coarsen<<<blocks_for(C.g.Nx, C.g.Ny), THREADS, 0, s->stream>>>(C.g, F.g, C.cx, C.cy, F.eta, F.bu, F.bv, C.eta, C.bu, C.bv);
            CHK(launch_check(s));
            CHK(corners_flags(s, (int)l, true));
        }
    }
    if (s->op.precision != NK_PRECISION_FP64) {
        for (size_t l = 0; l < s->lev.size(); l++) {
            Level &L = s->lev[l];
            k_cast_coeffs<<<blocks_for(L.g.Nx + 1, L.g.Ny + 1), THREADS, 0, s->stream>>>(L.g, L.eta, L.etaff, L.bu, L.bv, L.eta32,
                                                                                          L.etaff32, L.bu32, L.bv32);
            CHK(launch_check(s));
        }
        if (s->rl1u32 && s->op.preconditioner == NK_PRECONDITIONER_MULTIGRID) {
            Level &L0 = s->lev[0];
            const float wf = (float)(s->ph.walls == NK_NO_SLIP ? 2.0 : 0.0);
            NK_LAUNCH_TOPO(L0.g, k_l1_recip, float, blocks_for(L0.g.Nx, L0.g.Ny), THREADS, 0, s->stream, L0.g, coef<float>(L0), wf,
                           s->rl1u32, s->rl1v32);
            CHK(launch_check(s));
        }
    }
    if (s->op.preconditioner == NK_PRECONDITIONER_MULTIGRID && s->op.coarse_operator == NK_COARSE_GALERKIN && s->lev.size() > 1) {
        if (s->op.precision != NK_PRECISION_FP64) CHK(galerkin_build<float>(s));
        else CHK(galerkin_build<double>(s));
    }
    return NK_OK;
}

// ---- multigrid V-cycle ------------------------------------------------------------------------------------------

template <typename T> struct Pair { T *u, *v; };

// The finest-level sweeps through the neighbour masks (k_sweep_m, k_sweep_m_dot; k_sweep_m_z64 in the mixed form's FP32
// V-cycle) unless NK_NO_NMASK. Returns false (nothing launched) on coarser levels.
template <typename T>
bool launch_sweep_m(nk_solver *s, int l, const KG &g, dim3 B2, const ProfScope &ps, Pair<T> cur, Pair<T> oth, const Coef<T> &c, T wf,
                    const T *bu, const T *bv, const Scal *gate, bool dot) {
    if (l != 0 || !s->nm16 || !s->use_nmask) return false;
    const bool tile = s->apply_tile && sizeof(T) == 8;   // measured: FP64 gains, the FP32 sweep loses (11.1 vs 9.4 µs)
    if (dot && tile)
        NK_LAUNCH_TOPO_EV(g, k_sweep_mt_dot, T, B2, THREADS, 0, s->stream, ps.e0, ps.e1, g, (const T *)cur.u, (const T *)cur.v, c,
                          (const uint16_t *)s->nm16, wf, bu, bv, oth.u, oth.v, red(s, FIN_BETA), gate);
    else if (dot)
        NK_LAUNCH_TOPO_EV(g, k_sweep_m_dot, T, B2, THREADS, 0, s->stream, ps.e0, ps.e1, g, (const T *)cur.u, (const T *)cur.v, c,
                          (const uint16_t *)s->nm16, wf, bu, bv, oth.u, oth.v, red(s, FIN_BETA), gate);
    else if (tile)
        NK_LAUNCH_TOPO_EV(g, k_sweep_mt, T, B2, THREADS, 0, s->stream, ps.e0, ps.e1, g, (const T *)cur.u, (const T *)cur.v, c,
                          (const uint16_t *)s->nm16, wf, bu, bv, oth.u, oth.v, gate);
    else
        NK_LAUNCH_TOPO_EV(g, k_sweep_m, T, B2, THREADS, 0, s->stream, ps.e0, ps.e1, g, (const T *)cur.u, (const T *)cur.v, c,
                          (const uint16_t *)s->nm16, wf, bu, bv, oth.u, oth.v, gate);
    return true;
}
template <typename T>
bool launch_sweep_m_z64(nk_solver *s, int l, const KG &g, dim3 B2, Pair<T> cur, Pair<T> oth, const Coef<T> &c, T wf, const T *bu,
                        const T *bv, const Scal *gate) {
    if constexpr (sizeof(T) == 4) {
        if (l != 0 || !s->nm16 || !s->use_nmask) return false;
        NK_LAUNCH_TOPO(g, k_sweep_m_z64, T, B2, THREADS, 0, s->stream, g, (const T *)cur.u, (const T *)cur.v, c,
                       (const uint16_t *)s->nm16, wf, bu, bv, oth.u, oth.v, red(s, FIN_BETA), gate, (const double *)s->fz_r64u,
                       (const double *)s->fz_r64v, s->fz_z64u, s->fz_z64v);
        return true;
    }
    return false;
}

// smooth! from a zero initial guess (fill!(x, 0) then `sweeps` sweeps); result in x0 or x1 (returned)
template <typename T>
nk_status smooth_from_zero(nk_solver *s, int l, Pair<T> x0, Pair<T> x1, const T *bu, const T *bv, int sweeps, bool final_dot,
                           const Scal *gate, Pair<T> *out) {
    Level &L = s->lev[l];
    Coef<T> c = coef<T>(L);
    const T wf = (T)(s->ph.walls == NK_NO_SLIP ? 2.0 : 0.0);
    const dim3 B = blocks_for(L.g.Nx, L.g.Ny);
    Vecs<T> &V = vecs<T>(L);
    if (s->op.smoother == NK_LINE_JACOBI || sweeps == 0) {
        k_fill2<T><<<blocks_for(L.g.Nx + 1, L.g.Ny + 1), THREADS, 0, s->stream>>>(L.g, x0.u, x0.v, T(0), gate);
        CHK(launch_check(s));
        T *qu = (l == 0) ? work<T>(s).qu : V.qu, *qv = (l == 0) ? work<T>(s).qv : V.qv;
        T *t = V.t;
        for (int k = 0; k < sweeps; k++) {
            NK_LAUNCH_TOPO(L.g, k_apply, T, B, THREADS, 0, s->stream, L.g, x0.u, x0.v, c, wf, qu, qv, nullptr, nullptr, red(s, FIN_NONE), 0, gate);
            CHK(launch_check(s));
            NK_LAUNCH_TOPO(L.g, k_line_relax, T, dim3((L.g.Nx + NT - 1) / NT), THREADS, 0, s->stream, L.g, x0.u, x0.v, bu, bv, qu, qv, t, c, wf, gate);
            CHK(launch_check(s));
        }
        if (final_dot) {
            k_dot<T><<<B, THREADS, 0, s->stream>>>(L.g, bu, bv, x0.u, x0.v, nullptr, nullptr, red(s, FIN_BETA), gate);
            CHK(launch_check(s));
            CHK(finish_beta(s, B, 2, gate));
        }
        *out = x0;
        return NK_OK;
    }
    Pair<T> cur = x0, oth = x1;
    const dim3 B2 = blocks_for_rows(L.g.Nx, L.g.Ny, SROWS);
    int k0 = 0;
    if (sizeof(T) == 4 && l == 0 && s->rl1u32 && s->fuse_pre && sweeps >= 2 && !final_dot) {   // first two FP32 sweeps in one pass
        NK_LAUNCH_TOPO(L.g, k_sweep_pre_t, T, B2, THREADS, 0, s->stream, L.g, c, wf, bu, bv, (const T *)s->rl1u32, (const T *)s->rl1v32,
                       oth.u, oth.v, gate);
        CHK(launch_check(s));
        std::swap(cur, oth);
        k0 = 2;
    }
    for (int k = k0; k < sweeps; k++) {
        const bool last = (k == sweeps - 1) && final_dot;
        if (k == 0) {
            NK_LAUNCH_TOPO(L.g, k_sweep_zero, T, B2, THREADS, 0, s->stream, L.g, c, wf, bu, bv, cur.u, cur.v, red(s, FIN_BETA),
                           last ? 1 : 0, gate);
        } else {
            ProfScope ps(s, NK_PROFILE_SWEEP, l, (int)sizeof(T));
            if (launch_sweep_m(s, l, L.g, B2, ps, cur, oth, c, wf, bu, bv, gate, last)) {
            } else if (last)
                NK_LAUNCH_TOPO_EV(L.g, k_sweep_dot, T, B2, THREADS, 0, s->stream, ps.e0, ps.e1, L.g, (const T *)cur.u, (const T *)cur.v, c,
                                  wf, bu, bv, oth.u, oth.v, red(s, FIN_BETA), gate);
            else
                NK_LAUNCH_TOPO_EV(L.g, k_sweep, T, B2, THREADS, 0, s->stream, ps.e0, ps.e1, L.g, (const T *)cur.u, (const T *)cur.v, c,
                                  wf, bu, bv, oth.u, oth.v, gate);
        }
        CHK(launch_check(s));
        if (last) CHK(finish_beta(s, B2, 1, gate));
        if (k > 0) std::swap(cur, oth);
    }
    *out = cur;
    return NK_OK;
}

// smooth! from the current x (post-smoothing)
template <typename T>
nk_status smooth_from(nk_solver *s, int l, Pair<T> cur, Pair<T> oth, const T *bu, const T *bv, int sweeps, bool final_dot,
                      const Scal *gate, Pair<T> *out) {
    Level &L = s->lev[l];
    Coef<T> c = coef<T>(L);
    const T wf = (T)(s->ph.walls == NK_NO_SLIP ? 2.0 : 0.0);
    const dim3 B = blocks_for(L.g.Nx, L.g.Ny);
    Vecs<T> &V = vecs<T>(L);
    if (s->op.smoother == NK_LINE_JACOBI) {
        T *qu = (l == 0) ? work<T>(s).qu : V.qu, *qv = (l == 0) ? work<T>(s).qv : V.qv;
        for (int k = 0; k < sweeps; k++) {
            NK_LAUNCH_TOPO(L.g, k_apply, T, B, THREADS, 0, s->stream, L.g, cur.u, cur.v, c, wf, qu, qv, nullptr, nullptr, red(s, FIN_NONE), 0, gate);
            CHK(launch_check(s));
            NK_LAUNCH_TOPO(L.g, k_line_relax, T, dim3((L.g.Nx + NT - 1) / NT), THREADS, 0, s->stream, L.g, cur.u, cur.v, bu, bv, qu, qv, V.t, c, wf, gate);
            CHK(launch_check(s));
        }
    } else {
        const dim3 B2 = blocks_for_rows(L.g.Nx, L.g.Ny, SROWS);
        for (int k = 0; k < sweeps; k++) {
            const bool last = (k == sweeps - 1) && final_dot;
            {
                const bool fz = last && l == 0 && sizeof(T) == 4 && s->fz_r64u;   // mixed: r64·z64 (and z64) here
                if (fz) {
                    if (!launch_sweep_m_z64(s, l, L.g, B2, cur, oth, c, wf, bu, bv, gate))
                        NK_LAUNCH_TOPO(L.g, k_sweep_z64, T, B2, THREADS, 0, s->stream, L.g, (const T *)cur.u, (const T *)cur.v, c, wf, bu,
                                       bv, oth.u, oth.v, red(s, FIN_BETA), gate, (const double *)s->fz_r64u, (const double *)s->fz_r64v,
                                       s->fz_z64u, s->fz_z64v);
                } else {
                    ProfScope ps(s, NK_PROFILE_SWEEP, l, (int)sizeof(T));
                    if (launch_sweep_m(s, l, L.g, B2, ps, cur, oth, c, wf, bu, bv, gate, last)) {
                    } else if (last)
                        NK_LAUNCH_TOPO_EV(L.g, k_sweep_dot, T, B2, THREADS, 0, s->stream, ps.e0, ps.e1, L.g, (const T *)cur.u,
                                          (const T *)cur.v, c, wf, bu, bv, oth.u, oth.v, red(s, FIN_BETA), gate);
                    else
                        NK_LAUNCH_TOPO_EV(L.g, k_sweep, T, B2, THREADS, 0, s->stream, ps.e0, ps.e1, L.g, (const T *)cur.u,
                                          (const T *)cur.v, c, wf, bu, bv, oth.u, oth.v, gate);
                }
            }
            CHK(launch_check(s));
            if (last) CHK(finish_beta(s, B2, 1, gate));
            std::swap(cur, oth);
        }
        if (final_dot && sweeps > 0) { *out = cur; return NK_OK; }
    }
    if (final_dot) {
        k_dot<T><<<B, THREADS, 0, s->stream>>>(L.g, bu, bv, cur.u, cur.v, nullptr, nullptr, red(s, FIN_BETA), gate);
        CHK(launch_check(s));
        CHK(finish_beta(s, B, 2, gate));
    }
    *out = cur;
    return NK_OK;
}

// v_cycle! (src/StressBalances/multigrid_preconditioner.jl:144-163). x ≈ A⁻¹ b from zero at level l; the result
// lands in x0 or x1 (returned in *out). final_dot: Σ b·x at the end → FIN_BETA (finest level of the PCG).
// The coarse tail (levels tail_start..coarsest) in one LDS-resident workgroup: x at level l from b at level l.
template <typename T>
nk_status coarse_tail(nk_solver *s, int l, Pair<T> x, const T *bu, const T *bv, const Scal *gate) {
    const int pi = sizeof(T) == 8 ? 0 : 1;
    TailDesc<T> d;
    std::memset(&d, 0, sizeof(d));
    d.n = (int)s->lev.size() - l;
    d.sweeps = s->op.smoothing_sweeps;
    d.coarsest = s->op.coarsest_sweeps;
    d.wf = (T)(s->ph.walls == NK_NO_SLIP ? 2.0 : 0.0);
    d.gbu = bu; d.gbv = bv; d.gxu = x.u; d.gxv = x.v;
    const KG L0g = s->lev[l].g;
    int off_t = 0;
    for (int k = 0; k < d.n; k++) {
        Level &L = s->lev[l + k];
        TailLevel<T> &t = d.lv[k];
        t.lg = L.g;
        t.lg.pitch = L.g.Nx + 1;
        t.gpitch = L.g.pitch;
        t.hx = L.cx; t.hy = L.cy;
        Coef<T> c = coef<T>(L);
        t.eta = c.eta; t.etaff = c.etaff; t.bu = c.bu; t.bv = c.bv; t.act = L.act;
        const int plane = (L.g.Nx + 1) * (L.g.Ny + 1);
        int *offs[] = {&t.o_eta, &t.o_etaff, &t.o_bu, &t.o_bv, &t.o_xu, &t.o_xv, &t.o_bu_, &t.o_bv_, &t.o_qu, &t.o_qv};
        for (int *o : offs) { *o = off_t; off_t += plane; }
    }
    int off_b = off_t * (int)sizeof(T);
    for (int k = 0; k < d.n; k++) {
        Level &L = s->lev[l + k];
        d.lv[k].o_act = off_b;
        off_b += (L.g.Nx + 1) * (L.g.Ny + 1);
    }
    NK_LAUNCH_TOPO(L0g, k_coarse_cycle, T, 1, TAIL_THREADS, (size_t)off_b, s->stream, d, gate);
    (void)pi;
    return launch_check(s);
}

template <typename T>
nk_status v_cycle(nk_solver *s, int l, Pair<T> x0, Pair<T> x1, const T *bu, const T *bv, bool final_dot, const Scal *gate,
                  Pair<T> *out);

// The V-cycle below a level with a Galerkin operator (l ≥ 1): ℓ1-Jacobi sweeps and residuals with the stencil,
// the same restriction and prolongation kernels as the rediscretized cycle.
// Levels l..coarsest of a Galerkin V-cycle in one LDS-resident workgroup: x at level l from b at level l.
template <typename T>
nk_status galerkin_tail(nk_solver *s, int l, Pair<T> x, const T *bu, const T *bv, const Scal *gate) {
    GTailDesc<T> d;
    std::memset(&d, 0, sizeof(d));
    d.n = (int)s->lev.size() - l;
    d.sweeps = s->op.smoothing_sweeps;
    d.coarsest = s->op.coarsest_sweeps;
    d.gbu = bu; d.gbv = bv; d.gxu = x.u; d.gxv = x.v;
    d.cinv = (const double *)s->cinv;
    int off_t = 0;
    for (int k = 0; k < d.n; k++) {
        Level &L = s->lev[l + k];
        GTailLevel<T> &t = d.lv[k];
        t.lg = L.g;
        t.lg.pitch = L.g.Nx + 1;
        t.gpitch = L.g.pitch;
        t.hx = L.cx; t.hy = L.cy;
        t.cst = (const T *)L.gcst;
        t.act = L.act;
        t.o_st = -1;
        const int plane = (L.g.Nx + 1) * (L.g.Ny + 1);
        int *offs[] = {&t.o_x0u, &t.o_x0v, &t.o_x1u, &t.o_x1v, &t.o_bu, &t.o_bv, &t.o_ru, &t.o_rv};
        for (int *o : offs) { *o = off_t; off_t += plane; }
    }
    if (d.cinv) {   // FP64 partial products of the dense solve and its FP64 right-hand side, 8-byte aligned
        const KG &gc = s->lev.back().g;
        const int n = 2 * gc.Nx * gc.Ny;
        d.dgroups = TAIL_THREADS / n > 0 ? TAIL_THREADS / n : 1;
        constexpr int per8 = 8 / (int)sizeof(T);
        const int at = (off_t + per8 - 1) / per8 * per8;
        d.o_dp = at * (int)sizeof(T);
        off_t = at + (d.dgroups + 1) * n * per8;   // + the coarsest right-hand side in FP64, in unknown order
    }
    {   // copy the stencils of the smallest levels into LDS while they fit (the coarsest needs none with a dense inverse)
        int flags = 0;
        for (int k = 0; k < d.n; k++) flags += (s->lev[l + k].g.Nx + 1) * (s->lev[l + k].g.Ny + 1);
        const int cap = 160 * 1024;
        constexpr int align = 16 / (int)sizeof(T);   // 16-byte vector copies and row loads
        for (int k = d.n - (d.cinv ? 2 : 1); k >= 0; k--) {
            const Level &L = s->lev[l + k];
            const int need = GS_ALLOC * (L.g.Nx + 1) * (L.g.Ny + 1);
            const int at = (off_t + align - 1) / align * align;
            if ((at + need) * (int)sizeof(T) + flags > cap) break;
            d.lv[k].o_st = at;
            off_t = at + need;
        }
    }
    int off_b = off_t * (int)sizeof(T);
    for (int k = 0; k < d.n; k++) {
        Level &L = s->lev[l + k];
        d.lv[k].o_act = off_b;
        off_b += (L.g.Nx + 1) * (L.g.Ny + 1);
    }
    NK_LAUNCH_TOPO(s->lev[l].g, k_g_tail, T, 1, TAIL_THREADS, (size_t)off_b, s->stream, d, gate);
    return launch_check(s);
}

template <typename T>
nk_status v_cycle_galerkin(nk_solver *s, int l, Pair<T> x0, Pair<T> x1, const T *bu, const T *bv, const Scal *gate, Pair<T> *out) {
    const int nlev = (int)s->lev.size();
    if (s->gtail_start >= 1 && l >= s->gtail_start) {
        CHK(galerkin_tail<T>(s, l, x0, bu, bv, gate));
        *out = x0;
        return NK_OK;
    }
    Level &L = s->lev[l];
    Vecs<T> &V = vecs<T>(L);
    const T *st = gs<T>(L);
    const T *pk = (s->use_gpack && L.gpk) ? (const T *)L.gpk : nullptr;
    const dim3 B = blocks_for(L.g.Nx, L.g.Ny);
    const bool lat = (long)L.g.Nx * L.g.Ny <= s->latency_points;   // a latency-bound level
    auto smooth = [&](Pair<T> &cur, Pair<T> &oth, int sweeps, bool from_zero) -> nk_status {
        if (from_zero && sweeps == 0) {
            k_fill2<T><<<blocks_for(L.g.Nx + 1, L.g.Ny + 1), THREADS, 0, s->stream>>>(L.g, cur.u, cur.v, T(0), gate);
            return launch_check(s);
        }
        int k0 = 0;
        if (from_zero && sweeps >= 2 && s->fuse_pre) {   // the first two sweeps in one pass
            if (pk && s->apply_tile && lat) NK_LAUNCH_TOPO(L.g, k_gp_sweep_pre_tl, T, B, THREADS, 0, s->stream, L.g, st, pk, bu, bv, oth.u, oth.v, gate);
            else if (pk && s->apply_tile) NK_LAUNCH_TOPO(L.g, k_gp_sweep_pre_t, T, B, THREADS, 0, s->stream, L.g, st, pk, bu, bv, oth.u, oth.v, gate);
            else if (pk) NK_LAUNCH_TOPO(L.g, k_gp_sweep_pre, T, B, THREADS, 0, s->stream, L.g, st, pk, bu, bv, oth.u, oth.v, gate);
            else NK_LAUNCH_TOPO(L.g, k_g_sweep_pre, T, B, THREADS, 0, s->stream, L.g, st, bu, bv, oth.u, oth.v, gate);
            CHK(launch_check(s));
            std::swap(cur, oth);
            k0 = 2;
        }
        for (int k = k0; k < sweeps; k++) {
            if (from_zero && k == 0) {
                NK_LAUNCH_TOPO(L.g, k_g_sweep, T, B, THREADS, 0, s->stream, L.g, st, cur.u, cur.v, bu, bv, cur.u, cur.v, 1, gate);
            } else {
                ProfScope ps(s, NK_PROFILE_GALERKIN_SWEEP, l, (int)sizeof(T));
                if (pk && s->apply_tile && lat)
                    NK_LAUNCH_TOPO_EV(L.g, k_gp_sweep_tl, T, B, THREADS, 0, s->stream, ps.e0, ps.e1, L.g, st, pk, cur.u, cur.v, bu, bv,
                                      oth.u, oth.v, gate);
                else if (pk && s->apply_tile)
                    NK_LAUNCH_TOPO_EV(L.g, k_gp_sweep_t, T, B, THREADS, 0, s->stream, ps.e0, ps.e1, L.g, st, pk, cur.u, cur.v, bu, bv,
                                      oth.u, oth.v, gate);
                else if (pk)
                    NK_LAUNCH_TOPO_EV(L.g, k_gp_sweep, T, B, THREADS, 0, s->stream, ps.e0, ps.e1, L.g, st, pk, cur.u, cur.v, bu, bv, oth.u,
                                      oth.v, gate);
                else
                    NK_LAUNCH_TOPO_EV(L.g, k_g_sweep, T, B, THREADS, 0, s->stream, ps.e0, ps.e1, L.g, st, cur.u, cur.v, bu, bv, oth.u, oth.v,
                                      0, gate);
                std::swap(cur, oth);
            }
            CHK(launch_check(s));
        }
        return NK_OK;
    };
    Pair<T> cur = x0, oth = x1;
    if (l == nlev - 1) {
        if (s->cinv) {
            k_g_dense_solve<T><<<1, TAIL_THREADS, 0, s->stream>>>(L.g, (const double *)s->cinv, bu, bv, cur.u, cur.v, gate);
            CHK(launch_check(s));
        } else {
            CHK(smooth(cur, oth, s->op.coarsest_sweeps, true));
        }
        *out = cur;
        return NK_OK;
    }
    CHK(smooth(cur, oth, s->op.smoothing_sweeps, true));
    Level &C = s->lev[l + 1];
    Vecs<T> &CV = vecs<T>(C);
    if (pk && s->apply_tile && s->fuse_rr && (long)L.g.Nx * L.g.Ny <= s->rr_points) {   // residual and restriction in one launch
        const dim3 Brr((C.g.Nx + RRX - 1) / RRX, (C.g.Ny + RRY - 1) / RRY);
        NK_LAUNCH_TOPO(L.g, k_gp_residual_restrict, T, Brr, dim3(RR_THREADS), 0, s->stream, L.g, pk, cur.u, cur.v, bu, bv, L.act, C.g,
                       C.cx, C.cy, C.act, CV.bu, CV.bv, gate);
        CHK(launch_check(s));
    } else {
        if (pk && s->apply_tile && lat)
            NK_LAUNCH_TOPO(L.g, k_gp_residual_tl, T, B, THREADS, 0, s->stream, L.g, pk, cur.u, cur.v, bu, bv, V.qu, V.qv, gate);
        else if (pk && s->apply_tile)
            NK_LAUNCH_TOPO(L.g, k_gp_residual_t, T, B, THREADS, 0, s->stream, L.g, pk, cur.u, cur.v, bu, bv, V.qu, V.qv, gate);
        else if (pk) NK_LAUNCH_TOPO(L.g, k_gp_residual, T, B, THREADS, 0, s->stream, L.g, pk, cur.u, cur.v, bu, bv, V.qu, V.qv, gate);
        else NK_LAUNCH_TOPO(L.g, k_g_residual, T, B, THREADS, 0, s->stream, L.g, st, cur.u, cur.v, bu, bv, V.qu, V.qv, gate);
        CHK(launch_check(s));
        NK_LAUNCH_TOPO(C.g, k_restrict, T, blocks_for(C.g.Nx, C.g.Ny), THREADS, 0, s->stream, C.g, L.g, C.cx, C.cy, C.act, L.act, V.qu,
                       V.qv, CV.bu, CV.bv, gate);
        CHK(launch_check(s));
    }
    Pair<T> cres;
    CHK(v_cycle<T>(s, l + 1, Pair<T>{CV.x0u, CV.x0v}, Pair<T>{CV.x1u, CV.x1v}, CV.bu, CV.bv, false, gate, &cres));
    // prolongation inside the first post-sweep on the launch-latency-bound levels (≤ 256 × 256 points by default)
    if (pk && s->apply_tile && s->fuse_pro && s->op.smoothing_sweeps >= 1 && (long)L.g.Nx * L.g.Ny <= s->pro_fuse_points) {
        const ProSrc<T> pr{C.g, C.cx, C.cy, L.act, C.act, cres.u, cres.v};
        NK_LAUNCH_TOPO(L.g, k_gp_sweep_pt, T, B, THREADS, 0, s->stream, L.g, st, pk, cur.u, cur.v, bu, bv, oth.u, oth.v, pr, gate);
        CHK(launch_check(s));
        std::swap(cur, oth);
        CHK(smooth(cur, oth, s->op.smoothing_sweeps - 1, false));
    } else {
        NK_LAUNCH_TOPO(L.g, k_prolongate, T, B, THREADS, 0, s->stream, L.g, C.g, C.cx, C.cy, L.act, C.act, cres.u, cres.v, cur.u, cur.v, gate);
        CHK(launch_check(s));
        CHK(smooth(cur, oth, s->op.smoothing_sweeps, false));
    }
    *out = cur;
    return NK_OK;
}

template <typename T>
nk_status v_cycle(nk_solver *s, int l, Pair<T> x0, Pair<T> x1, const T *bu, const T *bv, bool final_dot, const Scal *gate,
                  Pair<T> *out) {
    const int nlev = (int)s->lev.size();
    Level &L = s->lev[l];
    Vecs<T> &V = vecs<T>(L);
    const int ts = s->tail_start[sizeof(T) == 8 ? 0 : 1];
    if (ts >= 0 && l >= ts) {
        CHK(coarse_tail<T>(s, l, x0, bu, bv, gate));
        if (final_dot) {
            const dim3 B = blocks_for(L.g.Nx, L.g.Ny);
            k_dot<T><<<B, THREADS, 0, s->stream>>>(L.g, bu, bv, x0.u, x0.v, nullptr, nullptr, red(s, FIN_BETA), gate);
            CHK(launch_check(s));
            CHK(finish_beta(s, B, 2, gate));
        }
        *out = x0;
        return NK_OK;
    }
    if (l >= 1 && s->op.coarse_operator == NK_COARSE_GALERKIN) return v_cycle_galerkin<T>(s, l, x0, x1, bu, bv, gate, out);
    if (l == nlev - 1) return smooth_from_zero<T>(s, l, x0, x1, bu, bv, s->op.coarsest_sweeps, final_dot, gate, out);
    Pair<T> cur;
    CHK(smooth_from_zero<T>(s, l, x0, x1, bu, bv, s->op.smoothing_sweeps, false, gate, &cur));
    Pair<T> oth = (cur.u == x0.u) ? x1 : x0;
    Coef<T> c = coef<T>(L);
    const T wf = (T)(s->ph.walls == NK_NO_SLIP ? 2.0 : 0.0);
    T *qu = (l == 0) ? work<T>(s).qu : V.qu, *qv = (l == 0) ? work<T>(s).qv : V.qv;
    if (l == 0) launch_apply0<T>(s, c, wf, cur.u, cur.v, qu, qv, bu, bv, red(s, FIN_NONE), 0, gate);
    else NK_LAUNCH_TOPO(L.g, k_apply, T, blocks_for(L.g.Nx, L.g.Ny), THREADS, 0, s->stream, L.g, cur.u, cur.v, c, wf, qu, qv, bu, bv, red(s, FIN_NONE), 0, gate);
    CHK(launch_check(s));
    Level &C = s->lev[l + 1];
    Vecs<T> &CV = vecs<T>(C);
    NK_LAUNCH_TOPO(C.g, k_restrict, T, blocks_for(C.g.Nx, C.g.Ny), THREADS, 0, s->stream, C.g, L.g, C.cx, C.cy, C.act, L.act, qu, qv, CV.bu, CV.bv, gate);
    CHK(launch_check(s));
    Pair<T> cres;
    CHK(v_cycle<T>(s, l + 1, Pair<T>{CV.x0u, CV.x0v}, Pair<T>{CV.x1u, CV.x1v}, CV.bu, CV.bv, false, gate, &cres));
    NK_LAUNCH_TOPO(L.g, k_prolongate, T, blocks_for(L.g.Nx, L.g.Ny), THREADS, 0, s->stream, L.g, C.g, C.cx, C.cy, L.act, C.act, cres.u, cres.v, cur.u,
                                                                          cur.v, gate);
    CHK(launch_check(s));
    return smooth_from<T>(s, l, cur, oth, bu, bv, s->op.smoothing_sweeps, final_dot, gate, out);
}

// ---- PCG --------------------------------------------------------------------------------------------------------

nk_status read_scalars(nk_solver *s) {
    HIPCHK(hipMemcpyAsync(s->S_host, s->S, sizeof(Scal), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    prof_harvest(s);
    return NK_OK;
}

nk_status write_scalars(nk_solver *s, const nk_control &ctl) {
    Scal h;
    std::memset(&h, 0, sizeof(h));
    h.reltol = ctl.reltol;
    h.abstol = ctl.abstol;
    h.fixed = ctl.kind == NK_FIXED_ITERATIONS;
    h.maxiter = h.fixed ? ctl.iterations : ctl.maxiter;
    h.zs = h.zi = 1.0;
    *s->S_host = h;
    HIPCHK(hipMemcpyAsync(s->S, s->S_host, sizeof(Scal), hipMemcpyHostToDevice, s->stream));
    return NK_OK;
}

// Mixed-precision preconditioner: z = V₃₂(r) with r cast to FP32 and z promoted back to FP64, then the PCG dot(s)
// r·z (and r·r at the start) in FP64 → finisher `fin`. Inside the PCG loop (FIN_BETA) the FP32 copy of r comes from
// k_update_xr, and with point Jacobi the V-cycle's last finest-level sweep writes z64 and sums r·z itself.
// With zout (the PCG's direction update reads the FP32 z itself) the last sweep does not store z64 and *zout names the
// FP32 result; *zout stays null where the result is only in the FP64 z.
nk_status precondition_fp32(nk_solver *s, int fin, const Scal *gate, Pair<float> *zout = nullptr) {
    const KG &g = s->g;
    const dim3 B = blocks_for(g.Nx, g.Ny);
    Work<double> &W = s->w64;
    Work<float> &V = s->w32;
    if (fin == FIN_BETA && s->op.smoother == NK_POINT_JACOBI && s->op.smoothing_sweeps > 0 && s->lev.size() >= 2 &&
        s->tail_start[1] != 0) {
        s->fz_r64u = W.ru; s->fz_r64v = W.rv;
        s->fz_z64u = zout ? nullptr : W.zu;
        s->fz_z64v = zout ? nullptr : W.zv;
        Pair<float> z32;
        const nk_status st = v_cycle<float>(s, 0, Pair<float>{V.zu, V.zv}, Pair<float>{V.z2u, V.z2v}, V.ru, V.rv, true, gate, &z32);
        s->fz_r64u = s->fz_r64v = nullptr; s->fz_z64u = s->fz_z64v = nullptr;
        if (zout) *zout = z32;
        return st;
    }
    if (fin != FIN_BETA) {   // outside the loop: the scale from ‖r‖ first (inside, k_update_xr applied the α-time scale)
        k_dot<double><<<B, THREADS, 0, s->stream>>>(g, W.ru, W.rv, W.ru, W.rv, nullptr, nullptr, red(s, FIN_SCALE), gate);
        CHK(launch_check(s));
        CHK(finish_red(s, B, 2, FIN_SCALE, gate));
        k_demote_r<<<B, THREADS, 0, s->stream>>>(g, W.ru, W.rv, V.ru, V.rv, gate, (const Scal *)s->S);
        CHK(launch_check(s));
    }
    Pair<float> z32;
    CHK(v_cycle<float>(s, 0, Pair<float>{V.zu, V.zv}, Pair<float>{V.z2u, V.z2v}, V.ru, V.rv, false, gate, &z32));
    k_promote_dot<<<B, THREADS, 0, s->stream>>>(g, z32.u, z32.v, W.zu, W.zv, W.ru, W.rv, red(s, fin), gate, (const Scal *)s->S);
    CHK(launch_check(s));
    return fin == FIN_BETA ? finish_beta(s, B, 2, gate) : finish_red(s, B, 2, fin, gate);
}

// conjugate_gradient! (src/Solvers/conjugate_gradient.jl:158-223) at precision T on the finest grid. When
// r_ready, the residual r of the initial guess is already in the workspace (the mixed inner solve from zero).
template <typename T>
nk_status pcg(nk_solver *s, T *xu, T *xv, const T *bu, const T *bv, const nk_control &ctl, bool r_ready, int *iters, double *res) {
    Level &L0 = s->lev[0];
    Work<T> &W = work<T>(s);
    Coef<T> c = coef<T>(L0);
    const T wf = (T)(s->ph.walls == NK_NO_SLIP ? 2.0 : 0.0);
    const KG &g = L0.g;
    const dim3 B = blocks_for(g.Nx, g.Ny);
    const int pre = s->op.preconditioner;
    CHK(write_scalars(s, ctl));
    if (!r_ready) {
        launch_apply0<T>(s, c, wf, xu, xv, W.ru, W.rv, bu, bv, red(s, FIN_NONE), 0, nullptr);
        CHK(launch_check(s));
    }
    Pair<T> z{W.zu, W.zv};
    const bool mixed = sizeof(T) == 8 && s->op.precision != NK_PRECISION_FP64 && pre == NK_PRECONDITIONER_MULTIGRID;
    if (mixed) {
        CHK(precondition_fp32(s, FIN_INIT, nullptr));
        k_update_p<T><<<B, THREADS, 0, s->stream>>>(g, W.zu, W.zv, W.pu, W.pv, 1, nullptr, nullptr, 0);
        CHK(launch_check(s));
    } else if (pre == NK_PRECONDITIONER_MULTIGRID) {
        CHK(v_cycle<T>(s, 0, Pair<T>{W.zu, W.zv}, Pair<T>{W.z2u, W.z2v}, W.ru, W.rv, false, nullptr, &z));
        k_dot<T><<<B, THREADS, 0, s->stream>>>(g, W.ru, W.rv, z.u, z.v, W.ru, W.rv, red(s, FIN_INIT), nullptr);
        CHK(launch_check(s));
        CHK(finish_red(s, B, 2, FIN_INIT, nullptr));
        k_update_p<T><<<B, THREADS, 0, s->stream>>>(g, z.u, z.v, W.pu, W.pv, 1, nullptr, nullptr, 0);
        CHK(launch_check(s));
    } else {
        NK_LAUNCH_TOPO(g, k_update_diag, T, B, THREADS, 0, s->stream, g, c, wf, pre, xu, xv, W.pu, W.pv, W.qu, W.qv, W.ru, W.rv, W.zu, W.zv, W.pu,
                                                        W.pv, 1, red(s, FIN_INIT), nullptr, nullptr, 0);
        CHK(launch_check(s));
        CHK(finish_red(s, B, 2, FIN_INIT, nullptr));
    }
    const int chunk = s->op.check_interval > 0 ? s->op.check_interval : 8;
    const bool fold = s->fold && pre == NK_PRECONDITIONER_MULTIGRID;   // α and β folded
    const bool fold_a = s->fold;                                       // α folded (also into k_update_diag)
    auto launch_chunk = [&]() -> nk_status {
        for (int k = 0; k < chunk; k++) {
            {
                ProfScope ps(s, NK_PROFILE_OPERATOR, 0, (int)sizeof(T));
                launch_apply0<T>(s, c, wf, (const T *)W.pu, (const T *)W.pv, W.qu, W.qv, (const T *)nullptr, (const T *)nullptr,
                                 red(s, FIN_ALPHA), 1, (const Scal *)s->S, ps.e0, ps.e1);
            }
            CHK(launch_check(s));
            if (!fold_a) CHK(finish_red(s, B, 1, FIN_ALPHA, s->S));
            if (pre == NK_PRECONDITIONER_MULTIGRID) {
                {
                    ProfScope ps(s, NK_PROFILE_UPDATE, 0, (int)sizeof(T));
                    Red rxr = red(s, FIN_RES);
                    rxr.partials = s->partials_xr;
                    hipExtLaunchKernelGGL(k_update_xr<T>, B, THREADS, 0, s->stream, ps.e0, ps.e1, 0, g, xu, xv, (const T *)W.pu,
                                          (const T *)W.pv, (const T *)W.qu, (const T *)W.qv, W.ru, W.rv, mixed ? s->w32.ru : nullptr,
                                          mixed ? s->w32.rv : nullptr, rxr, s->S, fold ? (const double *)s->partials : nullptr,
                                          (int)(B.x * B.y));
                }
                CHK(launch_check(s));
                CHK(finish_red(s, B, 1, FIN_RES, s->S, s->partials_xr));
                Pair<T> zz{W.zu, W.zv};
                Pair<float> z32{nullptr, nullptr};
                s->fold_pending = fold ? 1 : 0;
                if (mixed) CHK(precondition_fp32(s, FIN_BETA, s->S, s->use_z32 ? &z32 : nullptr));
                else CHK(v_cycle<T>(s, 0, Pair<T>{W.zu, W.zv}, Pair<T>{W.z2u, W.z2v}, W.ru, W.rv, true, s->S, &zz));
                if (s->fold_pending) {
                    s->fold_pending = 0;
                    s->err = "internal: the preconditioner left no r·z reduction for the folded β";
                    return NK_DEVICE_ERROR;
                }
                const double *fb = fold ? (const double *)s->partials : nullptr;
                if (z32.u) k_update_p32<T><<<B, THREADS, 0, s->stream>>>(g, z32.u, z32.v, W.pu, W.pv, s->S, fb, s->fold_nb);
                else k_update_p<T><<<B, THREADS, 0, s->stream>>>(g, zz.u, zz.v, W.pu, W.pv, 0, s->S, fb, s->fold_nb);
                CHK(launch_check(s));
            } else {
                {
                    ProfScope ps(s, NK_PROFILE_UPDATE, 0, (int)sizeof(T));
                    Red rd = red(s, FIN_BETA_RES);
                    rd.partials = s->partials_xr;
                    NK_LAUNCH_TOPO_EV(g, k_update_diag, T, B, THREADS, 0, s->stream, ps.e0, ps.e1, g, c, wf, pre, xu, xv, (const T *)W.pu,
                                      (const T *)W.pv, (const T *)W.qu, (const T *)W.qv, W.ru, W.rv, W.zu, W.zv, (T *)nullptr, (T *)nullptr, 0,
                                      rd, s->S, fold_a ? (const double *)s->partials : nullptr, (int)(B.x * B.y));
                }
                CHK(launch_check(s));
                CHK(finish_red(s, B, 2, FIN_BETA_RES, s->S, s->partials_xr));
                k_update_p<T><<<B, THREADS, 0, s->stream>>>(g, W.zu, W.zv, W.pu, W.pv, 0, s->S, nullptr, 0);
                CHK(launch_check(s));
            }
        }
        return NK_OK;
    };
    // The chunk's launches have fixed arguments for a given solve path and vector buffers, so they are captured once
    // and replayed; the kernels read the device stop flag, so a replay after convergence does no work.
    const bool use_graph = s->op.graphs && s->op.profile_kernel == NK_PROFILE_NONE;
    const int key = (int)sizeof(T) * 1000 + (mixed ? 500 : 0) + pre * 100 + ((const void *)xu == (const void *)s->f[NK_FIELD_U] ? 1 : 0) + chunk * 10000;
    if (use_graph && (s->chunk_graph_key != key || s->chunk_graph_x != (const void *)xu)) {
        if (s->chunk_graph) { (void)hipGraphExecDestroy(s->chunk_graph); s->chunk_graph = nullptr; }
        hipGraph_t graph;
        // Handles may be driven from several host threads (the server's slots): one capture at a time, never
        // beside an allocation or a free, and relaxed mode so other threads' stream work goes on meanwhile.
        std::lock_guard<std::mutex> lk(g_device_mu);
        HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed));
        nk_status cst = launch_chunk();
        hipError_t e = hipStreamEndCapture(s->stream, &graph);
        if (cst != NK_OK) return cst;
        if (e != hipSuccess) { s->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e); return NK_DEVICE_ERROR; }
        e = hipGraphInstantiate(&s->chunk_graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) { s->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e); return NK_DEVICE_ERROR; }
        s->chunk_graph_key = key;
        s->chunk_graph_x = xu;
    }
    int it0 = -1;
    for (;;) {
        HIPCHK(hipMemcpyAsync(s->S_host, s->S, sizeof(Scal), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        if (it0 >= 0) prof_trim_chunk(s, chunk, s->S_host->it - it0, s->S_host->done != 0);
        prof_harvest(s);
        if (s->S_host->done) break;
        it0 = s->S_host->it;
        if (use_graph) HIPCHK(hipGraphLaunch(s->chunk_graph, s->stream));
        else CHK(launch_chunk());
    }
    *iters = s->S_host->it;
    *res = s->S_host->res;
    if (s->S_host->nonfinite) { s->err = "the linear solve's residual norm is not finite (NaN or Inf in the inputs?)"; return NK_NONFINITE; }
    return NK_OK;
}

// NK_PRECISION_MIXED_REFINED: FP64 iterative refinement around an FP32 PCG. The defect r = b − A x and its norm are
// FP64 (the reference's operator and stopping rule ‖r‖ ≤ max(reltol ‖r₀‖, abstol) on the true residual); each step
// solves A₃₂ e = r·zs from zero with the FP32 PCG (FP32 operator through the neighbour masks, FP32 vectors and V-cycle,
// FP64 dots and α/β) to a relative refine_inner, zs = 2^−⌊log₂‖r‖⌋ keeping the FP32 right-hand side O(1), then
// x += e/zs in FP64. The linear iteration count is the inner PCG's total.
nk_status solve_refined(nk_solver *s, nk_info *info) {
    CHK(update_preconditioner(s));
    const KG &g = s->g;
    const dim3 B = blocks_for(g.Nx, g.Ny);
    Level &L0 = s->lev[0];
    Work<double> &W = s->w64;
    const nk_control &ctl = s->op.linear;
    double *xu = s->f[NK_FIELD_U], *xv = s->f[NK_FIELD_V];
    const double *bu = s->f[NK_FIELD_TAUU], *bv = s->f[NK_FIELD_TAUV];
    const double wf = s->ph.walls == NK_NO_SLIP ? 2.0 : 0.0;
    auto defect = [&](double *res) -> nk_status {
        launch_apply0<double>(s, coef<double>(L0), wf, xu, xv, W.ru, W.rv, bu, bv, red(s, FIN_NONE), 0, nullptr);
        CHK(launch_check(s));
        k_dot<double><<<B, THREADS, 0, s->stream>>>(g, W.ru, W.rv, W.ru, W.rv, nullptr, nullptr, red(s, FIN_NONE), nullptr);
        CHK(launch_check(s));
        CHK(finish_red(s, B, 1, FIN_NONE, nullptr));
        CHK(read_scalars(s));
        *res = std::sqrt(s->S_host->out[0]);
        if (!std::isfinite(*res)) { s->err = "the refinement defect is not finite (NaN or Inf in the inputs?)"; return NK_NONFINITE; }
        return NK_OK;
    };
    double res = 0.0;
    CHK(defect(&res));
    const bool fixed = ctl.kind == NK_FIXED_ITERATIONS;
    const int maxit = fixed ? ctl.iterations : ctl.maxiter;
    const double tol = jmax_h(ctl.reltol * res, ctl.abstol);
    int total = 0, steps = 0;
    while (total < maxit && (fixed || res > tol) && res > 0.0) {
        const int e = std::ilogb(res);
        const int k = e < -1000 ? 1000 : (e > 1000 ? -1000 : -e);
        const double zs = std::ldexp(1.0, k), zi = std::ldexp(1.0, -k);
        k_demote<<<B, THREADS, 0, s->stream>>>(g, W.ru, W.rv, s->b32u, s->b32v, s->w32.du, s->w32.dv, zs);
        CHK(launch_check(s));
        nk_control inner;
        std::memset(&inner, 0, sizeof(inner));
        if (fixed) {   // FixedIterations: the remaining count in one step
            inner.kind = NK_FIXED_ITERATIONS;
            inner.iterations = maxit - total;
        } else {
            inner.kind = NK_CONVERGENCE;
            inner.reltol = s->refine_inner;
            inner.maxiter = maxit - total;
        }
        int it = 0;
        double r32 = 0.0;
        CHK(pcg<float>(s, s->w32.du, s->w32.dv, s->b32u, s->b32v, inner, false, &it, &r32));
        k_promote_add<<<B, THREADS, 0, s->stream>>>(g, s->w32.du, s->w32.dv, xu, xv, zi);
        CHK(launch_check(s));
        total += it;
        steps++;
        CHK(defect(&res));
        if (it == 0 || fixed) break;
    }
    if (info) {
        info->linear_iterations = total;
        info->total_linear_iterations += total;
        info->last_residual_norm = res;
        info->refinement_steps = steps;
    }
    s->S_host->done_by_tol = !fixed && res <= tol;
    return NK_OK;
}

nk_status solve_linear(nk_solver *s, nk_info *info) {
    if (s->op.precision == NK_PRECISION_MIXED_REFINED && s->op.preconditioner == NK_PRECONDITIONER_MULTIGRID) return solve_refined(s, info);
    CHK(update_preconditioner(s));
    int it = 0;
    double res = 0.0;
    CHK(pcg<double>(s, s->f[NK_FIELD_U], s->f[NK_FIELD_V], s->f[NK_FIELD_TAUU], s->f[NK_FIELD_TAUV], s->op.linear, false, &it, &res));
    if (info) {
        info->linear_iterations = it;
        info->total_linear_iterations += it;
        info->last_residual_norm = res;
    }
    return NK_OK;
}

// ---- coefficient update ------------------------------------------------------------------------------------------

nk_status update_coefficients(nk_solver *s) {
    const KG &g = s->g;
    const PhysK p = physk(s);
    const Sig sg = sig_tables(s);
    const dim3 Bc = blocks_for(g.Nx, g.Ny), Be = blocks_for(g.Nx + 1, g.Ny + 1);
    double **f = s->f;
    const double *mu = s->ph.masking ? f[NK_FIELD_MU] : nullptr;
    if (s->ph.stress_balance == NK_SSA) {
        k_viscosity<<<Bc, THREADS, 0, s->stream>>>(g, p, sg, 0, f[NK_FIELD_U], f[NK_FIELD_V], f[NK_FIELD_A], s->plane,
                                                    f[NK_FIELD_ETASIGMA], f[NK_FIELD_H], mu, nullptr, nullptr, nullptr, nullptr,
                                                    f[NK_FIELD_ETAH], f[NK_FIELD_F2]);
        CHK(launch_check(s));
        CHK(corners_flags(s, 0, true));
        if (s->ph.friction != NK_FRICTION_NONE) {
            k_drag<<<Be, THREADS, 0, s->stream>>>(g, p, f[NK_FIELD_U], f[NK_FIELD_V], f[NK_FIELD_C], f[NK_FIELD_N], f[NK_FIELD_PHIU],
                                                   f[NK_FIELD_PHIV], f[NK_FIELD_BETAU], f[NK_FIELD_BETAV]);
            CHK(launch_check(s));
        }
    } else {
        k_basal_velocity<<<Be, THREADS, 0, s->stream>>>(g, f[NK_FIELD_U], f[NK_FIELD_V], f[NK_FIELD_BETABU], f[NK_FIELD_BETABV],
                                                         f[NK_FIELD_F2], f[NK_FIELD_UB], f[NK_FIELD_VB]);
        CHK(launch_check(s));
        if (s->ph.friction != NK_FRICTION_NONE) {
            k_drag<<<Be, THREADS, 0, s->stream>>>(g, p, f[NK_FIELD_UB], f[NK_FIELD_VB], f[NK_FIELD_C], f[NK_FIELD_N], f[NK_FIELD_PHIU],
                                                   f[NK_FIELD_PHIV], f[NK_FIELD_BETABU], f[NK_FIELD_BETABV]);
            CHK(launch_check(s));
        }
        k_viscosity<<<Bc, THREADS, 0, s->stream>>>(g, p, sg, 1, f[NK_FIELD_U], f[NK_FIELD_V], f[NK_FIELD_A], s->plane,
                                                    f[NK_FIELD_ETASIGMA], f[NK_FIELD_H], mu, f[NK_FIELD_BETABU], f[NK_FIELD_BETABV],
                                                    f[NK_FIELD_UB], f[NK_FIELD_VB], f[NK_FIELD_ETAH], f[NK_FIELD_F2]);
        CHK(launch_check(s));
        CHK(corners_flags(s, 0, true));
        k_effective_drag<<<Be, THREADS, 0, s->stream>>>(g, f[NK_FIELD_BETABU], f[NK_FIELD_BETABV], f[NK_FIELD_F2], f[NK_FIELD_BETAU],
                                                         f[NK_FIELD_BETAV]);
        CHK(launch_check(s));
    }
    s->flags_valid = true;
    return NK_OK;
}

nk_status driving_stress(nk_solver *s) {
    const KG &g = s->g;
    double **f = s->f;
    k_driving_stress<<<blocks_for(g.Nx + 1, g.Ny + 1), THREADS, 0, s->stream>>>(
        g, physk(s), f[NK_FIELD_H], f[NK_FIELD_S], s->ph.masking ? f[NK_FIELD_MU] : nullptr,
        s->ph.grounding == NK_GROUNDING_SUBGRID_DRIVING_STRESS ? f[NK_FIELD_BED] : nullptr, f[NK_FIELD_TAUU], f[NK_FIELD_TAUV],
        s->cls_u, s->cls_v);
    return launch_check(s);
}

nk_status post_process(nk_solver *s) {
    const KG &g = s->g;
    double **f = s->f;
    const dim3 Be = blocks_for(g.Nx + 1, g.Ny + 1);
    if (s->ph.stress_balance == NK_SSA) {
        k_extend<<<Be, THREADS, 0, s->stream>>>(g, s->Ns, s->plane, f[NK_FIELD_U], f[NK_FIELD_V], f[NK_FIELD_UB], f[NK_FIELD_VB],
                                                 f[NK_FIELD_USIGMA], f[NK_FIELD_VSIGMA]);
        return launch_check(s);
    }
    k_basal_velocity<<<Be, THREADS, 0, s->stream>>>(g, f[NK_FIELD_U], f[NK_FIELD_V], f[NK_FIELD_BETABU], f[NK_FIELD_BETABV],
                                                     f[NK_FIELD_F2], f[NK_FIELD_UB], f[NK_FIELD_VB]);
    CHK(launch_check(s));
    k_velocity_profile<<<Be, THREADS, 0, s->stream>>>(g, sig_tables(s), s->Ns, s->plane, f[NK_FIELD_ETASIGMA], f[NK_FIELD_H],
                                                       f[NK_FIELD_UB], f[NK_FIELD_VB], f[NK_FIELD_BETABU], f[NK_FIELD_BETABV],
                                                       f[NK_FIELD_USIGMA], f[NK_FIELD_VSIGMA]);
    return launch_check(s);
}

// Copies DIVA's lagged state (βᵇ at u and v faces, F₂, ησ on Nσ layers) from one set of planes to another.
nk_status copy_diva_state(nk_solver *s, const double *bbu, const double *bbv, const double *F2, const double *es, double *bbu2,
                          double *bbv2, double *F22, double *es2) {
    const KG &g = s->g;
    const dim3 Be = blocks_for(g.Nx + 1, g.Ny + 1);
    const double *src[] = {bbu, bbv, F2};
    double *dst[] = {bbu2, bbv2, F22};
    for (int k = 0; k < 3; k++) {
        k_copy_plane<<<Be, THREADS, 0, s->stream>>>(g, src[k], dst[k]);
        CHK(launch_check(s));
    }
    for (int k = 0; k < s->Ns; k++) {
        k_copy_plane<<<Be, THREADS, 0, s->stream>>>(g, es + (long)k * s->plane, es2 + (long)k * s->plane);
        CHK(launch_check(s));
    }
    return NK_OK;
}

// picard_solve! (src/Solvers/picard.jl:71-90)
nk_status picard(nk_solver *s, nk_info *info) {
    const nk_control &ctl = s->op.picard;
    const bool fixed = ctl.kind == NK_FIXED_ITERATIONS;
    const KG &g = s->g;
    const dim3 B = blocks_for(g.Nx, g.Ny);
    double D = INFINITY, X = 0.0;
    int it = 0;
    if (info) info->total_linear_iterations = 0;
    if (fixed) {
        // the initial guess (and DIVA's lagged state), for the unrolled adjoint: nk_adjoint replays the steps from it
        if (!s->x0u) {
            CHK(alloc_plane(s, &s->x0u, g));
            CHK(alloc_plane(s, &s->x0v, g));
        }
        k_copy2<double><<<B, THREADS, 0, s->stream>>>(g, s->f[NK_FIELD_U], s->f[NK_FIELD_V], s->x0u, s->x0v);
        CHK(launch_check(s));
        if (s->ph.stress_balance == NK_DIVA) {
            if (!s->s0_bbu) {
                CHK(alloc_plane(s, &s->s0_bbu, g));
                CHK(alloc_plane(s, &s->s0_bbv, g));
                CHK(alloc_plane(s, &s->s0_F2, g));
                CHK(alloc_plane(s, &s->s0_es, g, s->Ns));
            }
            CHK(copy_diva_state(s, s->f[NK_FIELD_BETABU], s->f[NK_FIELD_BETABV], s->f[NK_FIELD_F2], s->f[NK_FIELD_ETASIGMA],
                                s->s0_bbu, s->s0_bbv, s->s0_F2, s->s0_es));
        }
        s->x0_valid = true;
    } else {
        s->x0_valid = false;
    }
    for (;;) {
        if (fixed) { if (it >= ctl.iterations) break; }
        else if (!(D > jmax_h(ctl.reltol * X, ctl.abstol) && it < ctl.maxiter)) break;
        if (!fixed) {
            k_copy2<double><<<B, THREADS, 0, s->stream>>>(g, s->f[NK_FIELD_U], s->f[NK_FIELD_V], s->xmu, s->xmv);
            CHK(launch_check(s));
        }
        CHK(update_coefficients(s));
        CHK(solve_linear(s, info));
        if (!fixed) {
            k_picard_change<<<B, THREADS, 0, s->stream>>>(g, s->xmu, s->xmv, s->f[NK_FIELD_U], s->f[NK_FIELD_V], red(s, FIN_PICARD));
            CHK(launch_check(s));
            CHK(finish_red(s, B, 2, FIN_PICARD, nullptr));
            CHK(read_scalars(s));
            D = s->S_host->picard_D;
            X = s->S_host->picard_X;
            if (!std::isfinite(D) || !std::isfinite(X)) { s->err = "the Picard change or velocity norm is not finite"; return NK_NONFINITE; }
        }
        it++;
        if (s->verbose) {
            std::fprintf(stderr, "[nunatak_hip] picard %d: linear %d (total %lld), change %.3e, norm %.3e\n", it,
                         info ? info->linear_iterations : -1, info ? (long long)info->total_linear_iterations : -1LL, D, X);
            std::fflush(stderr);
        }
    }
    if (info) {
        info->picard_iterations = it;
        info->last_picard_change = D;
        info->last_picard_norm = X;
    }
    if (s->op.report_not_converged && !fixed && D > jmax_h(ctl.reltol * X, ctl.abstol)) {
        s->err = "Picard iteration stopped at maxiter before meeting its tolerance";
        return NK_NOT_CONVERGED;
    }
    return NK_OK;
}

nk_status check_field(const nk_solver *s, int id) {
    return (id < 0 || id >= NK_FIELD_COUNT) ? NK_INVALID_ARGUMENT : NK_OK;
}

void field_shape(const nk_solver *s, int id, int32_t shape[3]) {
    const LocInfo &L = LOC[id];
    shape[0] = s->grid.Nx + ((L.lx && s->grid.topology_x == NK_BOUNDED) ? 1 : 0);
    shape[1] = s->grid.Ny + ((L.ly && s->grid.topology_y == NK_BOUNDED) ? 1 : 0);
    shape[2] = (L.sigma == 1 || (L.sigma == 2 && s->ph.rate_factor_depth == NK_RATE_FACTOR_3D)) ? s->Ns : 1;
}

struct Timer {
    nk_solver *s;
    nk_info *info;
    Timer(nk_solver *s_, nk_info *i_) : s(s_), info(i_) {
        s->prof_ms = 0.0;
        s->prof_launches = s->prof_points = 0;
        s->prof_used = 0;
        (void)hipEventRecord(s->ev0, s->stream);
    }
    nk_status done(nk_status st) {
        (void)hipEventRecord(s->ev1, s->stream);
        (void)hipEventSynchronize(s->ev1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, s->ev0, s->ev1);
        prof_harvest(s);
        if (info) {
            info->elapsed_ms = ms;
            info->profiled_ms = s->prof_ms;
            info->profiled_launches = s->prof_launches;
            info->profiled_points = s->prof_points;
            info->profiled_precision = s->prof_precision;
        }
        hipError_t e = hipGetLastError();
        if (st == NK_OK && e != hipSuccess) { s->err = hipGetErrorString(e); return NK_DEVICE_ERROR; }
        return st;
    }
};

void drop_chunk_graph(nk_solver *s) {
    if (s->chunk_graph) { (void)hipStreamSynchronize(s->stream); (void)hipGraphExecDestroy(s->chunk_graph); }
    s->chunk_graph = nullptr;
    s->chunk_graph_key = -1;
    s->chunk_graph_x = nullptr;
}

nk_status validate(const nk_grid *g, const nk_physics *ph, const nk_solver_options *op, std::string &err) {
    if (!g || !ph || !op) { err = "null argument"; return NK_INVALID_ARGUMENT; }
    if (g->Nx < 1 || g->Ny < 1 || g->Nsigma < 1 || g->Nsigma > 256) { err = "grid sizes must be positive (Nσ ≤ 256)"; return NK_INVALID_ARGUMENT; }
    if (!(g->dx > 0) || !(g->dy > 0)) { err = "grid spacing must be positive"; return NK_INVALID_ARGUMENT; }
    {   // the stencil kernels address a plane with 32-bit byte offsets
        const double plane_bytes = 8.0 * (double)(g->Ny + 1) * (double)(((g->Nx + 1) + 63) / 64 * 64);
        if (plane_bytes >= 4294967296.0) { err = "grid too large: one FP64 plane must stay below 4 GiB"; return NK_INVALID_ARGUMENT; }
    }
    if (ph->masking && !(ph->mask_transition_thickness > 0)) {
        err = "transition_thickness must be positive";  // SmoothIceMask (src/Models/smooth_ice_mask.jl:42)
        return NK_INVALID_ARGUMENT;
    }
    if (ph->walls != NK_FREE_SLIP && ph->walls != NK_NO_SLIP) {
        err = "the coefficients must include the side walls, FreeSlipWalls() or NoSlipWalls()";  // ssa_operator.jl:56
        return NK_INVALID_ARGUMENT;
    }
    if (op->preconditioner < 0 || op->preconditioner > 2 || op->precision < 0 || op->precision > 2) {
        err = "unknown preconditioner or precision";
        return NK_INVALID_ARGUMENT;
    }
    if (op->precision != NK_PRECISION_FP64 && op->preconditioner != NK_PRECONDITIONER_MULTIGRID) {
        err = "mixed precision runs the multigrid V-cycle in FP32: it needs the multigrid preconditioner";
        return NK_INVALID_ARGUMENT;
    }
    if (op->coarse_operator != NK_COARSE_REDISCRETIZED && op->coarse_operator != NK_COARSE_GALERKIN) {
        err = "unknown coarse operator";
        return NK_INVALID_ARGUMENT;
    }
    if (op->coarse_operator == NK_COARSE_GALERKIN &&
        (op->preconditioner != NK_PRECONDITIONER_MULTIGRID || op->smoother != NK_POINT_JACOBI)) {
        err = "Galerkin coarse operators need the multigrid preconditioner with point-Jacobi smoothing";
        return NK_INVALID_ARGUMENT;
    }
    return NK_OK;
}

}  // namespace

// =====================================================================================================================
// C ABI
// =====================================================================================================================

extern "C" {

int32_t nk_abi_version(void) { return NK_ABI_VERSION; }

void nk_default_grid(nk_grid *g) {
    std::memset(g, 0, sizeof(*g));
    g->Nx = g->Ny = 4;
    g->topology_x = g->topology_y = NK_BOUNDED;
    g->halo = 4;
    g->Nsigma = 10;
    g->dx = g->dy = 1.0;
    g->bed_refinement = 2.0;
}

void nk_default_physics(nk_physics *p) {
    std::memset(p, 0, sizeof(*p));
    p->ice_density = 918.0; p->ocean_density = 1028.0; p->gravitational_acceleration = 9.81;
    p->grounding = NK_GROUNDING_NONE; p->masking = 1;
    p->mask_minimum_thickness = 1.0; p->mask_transition_thickness = 1.0;
    p->glen_exponent = 3.0; p->strain_rate_regularization = 1e-13;
    p->rate_factor_depth = NK_RATE_FACTOR_2D;
    p->friction = NK_FRICTION_NONE; p->friction_exponent = 1.0 / 3.0; p->coulomb_coefficient = 0.0;
    p->velocity_regularization = 1e-8;
    p->walls = NK_FREE_SLIP; p->stress_balance = NK_SSA;
}

void nk_default_till(nk_till *t) {
    if (!t) return;
    t->reference_effective_pressure = 1000.0;
    t->overburden_fraction = 0.02;
    t->reference_void_ratio = 0.69;
    t->compressibility = 0.12;
    t->maximum_water_thickness = 2.0;
    t->drainage_rate = 1e-3 / (365.25 * 86400.0);
}

void nk_default_options(nk_solver_options *o) {
    std::memset(o, 0, sizeof(*o));
    o->linear.kind = NK_CONVERGENCE; o->linear.reltol = 1e-8; o->linear.maxiter = 10000;
    o->picard.kind = NK_CONVERGENCE; o->picard.reltol = 1e-6; o->picard.maxiter = 100;
    o->preconditioner = NK_PRECONDITIONER_DIAGONAL;
    o->smoother = NK_POINT_JACOBI; o->smoothing_sweeps = 2; o->coarsest_sweeps = 20; o->minimum_coarse_size = 2;
    o->precision = NK_PRECISION_FP64; o->coarse_operator = NK_COARSE_REDISCRETIZED; o->check_interval = 8;
    o->profile_kernel = NK_PROFILE_NONE;
    o->graphs = 1;
}

nk_status nk_create(const nk_grid *grid, const nk_physics *physics, const nk_solver_options *options, nk_solver **out) {
    if (!out) return NK_INVALID_ARGUMENT;
    *out = nullptr;
    std::string err;
    nk_status st = validate(grid, physics, options, err);
    if (st != NK_OK) {
        std::fprintf(stderr, "nk_create: %s\n", err.c_str());
        return st;
    }
    std::lock_guard<std::mutex> lk(g_device_mu);
    nk_solver *s = new nk_solver();
    s->grid = *grid; s->ph = *physics; s->op = *options; s->op0 = *options;
    {
        const char *v = std::getenv("NK_VERBOSE");
        s->verbose = v && v[0] && v[0] != '0';
    }
    s->Ns = grid->Nsigma;
    s->g = make_kg(grid->Nx, grid->Ny, grid->topology_x, grid->topology_y, grid->dx, grid->dy);
    s->plane = plane_of(s->g);
    auto fail = [&](nk_status e) { *out = s; return e; };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) { s->err = "hipStreamCreate failed"; return fail(NK_DEVICE_ERROR); }
    (void)hipEventCreate(&s->ev0);
    (void)hipEventCreate(&s->ev1);
    for (int id = 0; id < NK_FIELD_COUNT; id++) {
        int nz = LOC[id].sigma ? s->Ns : 1;
        st = alloc_plane(s, &s->f[id], s->g, nz);
        if (st != NK_OK) return fail(st);
    }
    if ((st = alloc_plane(s, &s->cls_u, s->g)) != NK_OK) return fail(st);
    if ((st = alloc_plane(s, &s->cls_v, s->g)) != NK_OK) return fail(st);
    // σ tables: zf (Ns+1), zc, dz, w2 (Ns each), pw (Ns × Ns), as the reference's SigmaGrid computes them
    {
        int Ns = s->Ns;
        std::vector<double> t((Ns + 1) + 3 * Ns + Ns * Ns, 0.0);
        double *zf = t.data(), *zc = zf + Ns + 1, *dz = zc + Ns, *w2 = dz + Ns, *pw = w2 + Ns;
        for (int k = 0; k <= Ns; k++) zf[k] = sigma_face(k, Ns, grid->bed_refinement);
        for (int k = 0; k < Ns; k++) {
            zc[k] = (zf[k] + zf[k + 1]) / 2.0;
            dz[k] = zf[k + 1] - zf[k];
            w2[k] = power_integral(zf[k], zf[k + 1], 2);
        }
        for (int k = 0; k < Ns; k++)
            for (int kp = 0; kp <= k; kp++) pw[k * Ns + kp] = power_integral(zf[kp], kp == k ? zc[k] : zf[kp + 1], 1);
        if ((st = alloc_bytes(s, (void **)&s->sig, t.size() * sizeof(double))) != NK_OK) return fail(st);
        if (hipMemcpyAsync(s->sig, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, s->stream) != hipSuccess ||
            hipStreamSynchronize(s->stream) != hipSuccess) { s->err = "σ tables"; return fail(NK_DEVICE_ERROR); }
    }
    // levels: the finest uses the model's fields; coarse levels own theirs
    {
        Level L0;
        L0.g = s->g;
        L0.eta = s->f[NK_FIELD_ETAH]; L0.etaff = s->f[NK_FIELD_ETAHFF]; L0.bu = s->f[NK_FIELD_BETAU]; L0.bv = s->f[NK_FIELD_BETAV];
        s->lev.push_back(L0);
        if (options->preconditioner == NK_PRECONDITIONER_MULTIGRID) {
            KG cur = s->g;
            for (;;) {
                // a Galerkin hierarchy stops where the dense inverse takes over
                if (options->coarse_operator == NK_COARSE_GALERKIN && s->lev.size() >= 2 && 2 * cur.Nx * cur.Ny <= GDENSE_STOP) break;
                int cxh = coarsening(cur.Nx, grid->halo, options->minimum_coarse_size);
                int cyh = coarsening(cur.Ny, grid->halo, options->minimum_coarse_size);
                if (!cxh && !cyh) break;
                Level L;
                L.g = make_kg(cur.Nx / (cxh ? 2 : 1), cur.Ny / (cyh ? 2 : 1), cur.tx, cur.ty, cur.dx * (cxh ? 2 : 1), cur.dy * (cyh ? 2 : 1));
                L.cx = cxh; L.cy = cyh;
                s->lev.push_back(L);
                cur = L.g;
            }
        }
        for (size_t l = 0; l < s->lev.size(); l++) {
            Level &L = s->lev[l];
            if (l > 0) {
                double **d[] = {&L.eta, &L.etaff, &L.bu, &L.bv};
                for (auto p : d) if ((st = alloc_plane(s, p, L.g)) != NK_OK) return fail(st);
            }
            if ((st = alloc_plane(s, &L.act, L.g)) != NK_OK) return fail(st);
            double **v64[] = {&L.v64.x0u, &L.v64.x0v, &L.v64.x1u, &L.v64.x1v, &L.v64.bu, &L.v64.bv, &L.v64.qu, &L.v64.qv, &L.v64.t};
            for (auto p : v64) if ((st = alloc_plane(s, p, L.g)) != NK_OK) return fail(st);
            if (options->precision != NK_PRECISION_FP64) {
                float **c32[] = {&L.eta32, &L.etaff32, &L.bu32, &L.bv32};
                for (auto p : c32) if ((st = alloc_plane(s, p, L.g)) != NK_OK) return fail(st);
                float **v32[] = {&L.v32.x0u, &L.v32.x0v, &L.v32.x1u, &L.v32.x1v, &L.v32.bu, &L.v32.bv, &L.v32.qu, &L.v32.qv, &L.v32.t};
                for (auto p : v32) if ((st = alloc_plane(s, p, L.g)) != NK_OK) return fail(st);
            }
            if (options->coarse_operator == NK_COARSE_GALERKIN) {
                if (options->precision != NK_PRECISION_FP64) st = alloc_plane(s, &L.gs32, L.g, GS_ALLOC);
                else st = alloc_plane(s, &L.gs64, L.g, GS_ALLOC);
                if (st != NK_OK) return fail(st);
                const size_t es = options->precision != NK_PRECISION_FP64 ? 4 : 8;
                if (wants_gpack(L.g, (int)l) && (st = alloc_bytes(s, &L.gpk, GP_PLANES * es * (size_t)gp_plane(L.g))) != NK_OK)
                    return fail(st);
                if (wants_gcst(L.g, (int)l, es) &&
                    (st = alloc_bytes(s, &L.gcst, (size_t)(L.g.Nx + 1) * (size_t)(L.g.Ny + 1) * GS_ALLOC * es)) != NK_OK)
                    return fail(st);
            }
        }
    }
    {
        double **w[] = {&s->w64.ru, &s->w64.rv, &s->w64.zu, &s->w64.zv, &s->w64.z2u, &s->w64.z2v, &s->w64.pu, &s->w64.pv,
                        &s->w64.qu, &s->w64.qv, &s->w64.du, &s->w64.dv, &s->xmu, &s->xmv, &s->ad_qu, &s->ad_qv,
                        &s->ad_yu, &s->ad_yv, &s->ad_lu, &s->ad_lv, &s->ad_ru, &s->ad_rv, &s->ad_eta, &s->ad_exy,
                        &s->ad_bbu, &s->ad_bbv, &s->ad_ubb, &s->ad_vbb, &s->ad_F2n, &s->st_bbu, &s->st_bbv, &s->st_F2,
                        &s->out_bbu, &s->out_bbv, &s->out_F2, &s->cs_bbu, &s->cs_bbv, &s->cs_F2, &s->ad_xbu, &s->ad_xbv};
        for (auto p : w) if ((st = alloc_plane(s, p, s->g)) != NK_OK) return fail(st);
        double **wσ[] = {&s->st_es, &s->out_es, &s->cs_es};
        for (auto p : wσ) if ((st = alloc_plane(s, p, s->g, s->Ns)) != NK_OK) return fail(st);
        if ((st = alloc_plane(s, &s->nm16, s->g)) != NK_OK) return fail(st);
        const char *nn = std::getenv("NK_NO_NMASK");
        s->use_nmask = !(nn && nn[0] == '1');
        const char *na = std::getenv("NK_NO_ATILE");
        s->apply_tile = !(na && na[0] == '1');
        const char *nz = std::getenv("NK_NO_Z32");
        s->use_z32 = !(nz && nz[0] == '1');
        const char *np = std::getenv("NK_NO_GPACK");
        s->use_gpack = !(np && np[0] == '1');
        const char *nrr = std::getenv("NK_NO_RR_FUSE");
        s->fuse_rr = !(nrr && nrr[0] == '1');
        const char *nrrp = std::getenv("NK_RR_POINTS");
        if (nrrp) s->rr_points = std::strtol(nrrp, nullptr, 10);
        const char *nrs = std::getenv("NK_NO_RAP_SPLIT");
        s->rap_split = !(nrs && nrs[0] == '1');
        const char *nrc = std::getenv("NK_NO_RAP_COMPACT");
        s->rap_compact = !(nrc && nrc[0] == '1');
        const char *nrsp = std::getenv("NK_RAP_SPLIT_POINTS");
        if (nrsp) s->rap_split_points = std::strtol(nrsp, nullptr, 10);
        const char *nfp = std::getenv("NK_NO_FIN_PATTERN");
        s->fin_pattern = !(nfp && nfp[0] == '1');
        const char *ngr = std::getenv("NK_NO_GINV_REGS");
        s->ginv_regs = !(ngr && ngr[0] == '1');
        if (options->precision != NK_PRECISION_FP64) {
            float **w32[] = {&s->w32.ru, &s->w32.rv, &s->w32.zu, &s->w32.zv, &s->w32.z2u, &s->w32.z2v, &s->w32.pu, &s->w32.pv,
                             &s->w32.qu, &s->w32.qv, &s->w32.du, &s->w32.dv, &s->rl1u32, &s->rl1v32};
            for (auto p : w32) if ((st = alloc_plane(s, p, s->g)) != NK_OK) return fail(st);
            if (options->precision == NK_PRECISION_MIXED_REFINED) {
                if ((st = alloc_plane(s, &s->b32u, s->g)) != NK_OK || (st = alloc_plane(s, &s->b32v, s->g)) != NK_OK) return fail(st);
                const char *ir = std::getenv("NK_REFINE_INNER_RELTOL");
                if (ir) s->refine_inner = std::strtod(ir, nullptr);
            }
        }
    }
    // coarse tail: the longest run of coarsest levels whose LDS image fits the budget (point Jacobi only)
    if (options->preconditioner == NK_PRECONDITIONER_MULTIGRID && options->smoother == NK_POINT_JACOBI && s->lev.size() > 1) {
        const size_t budget = 150 * 1024;
        for (int pi = 0; pi < 2; pi++) {
            const size_t es = pi == 0 ? 8 : 4;
            size_t bytes = 0;
            int start = -1;
            for (int l = (int)s->lev.size() - 1; l >= 1; l--) {
                const Level &L = s->lev[l];
                const size_t plane = (size_t)(L.g.Nx + 1) * (size_t)(L.g.Ny + 1);
                const size_t b = plane * (10 * es + 1);
                if (bytes + b > budget || (int)s->lev.size() - l > TAIL_MAX) break;
                bytes += b;
                start = l;
            }
            if (start >= 0 && (int)s->lev.size() - start >= 2) { s->tail_start[pi] = start; s->tail_bytes[pi] = bytes + 64; }
        }
        const void *tails[] = {(const void *)k_coarse_cycle<double, 0, 0>, (const void *)k_coarse_cycle<double, 0, 1>,
                               (const void *)k_coarse_cycle<double, 1, 0>, (const void *)k_coarse_cycle<double, 1, 1>,
                               (const void *)k_coarse_cycle<float, 0, 0>, (const void *)k_coarse_cycle<float, 0, 1>,
                               (const void *)k_coarse_cycle<float, 1, 0>, (const void *)k_coarse_cycle<float, 1, 1>};
        for (const void *f : tails)
            if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess) s->tail_start[0] = s->tail_start[1] = -1;
        const char *nf = std::getenv("NK_NO_PRE_FUSE");
        s->fuse_pre = !(nf && nf[0] == '1');
        const char *npro = std::getenv("NK_NO_PRO_FUSE");
        s->fuse_pro = !(npro && npro[0] == '1');
        const char *nproc = std::getenv("NK_PRO_FUSE_POINTS");
        if (nproc) s->pro_fuse_points = std::strtol(nproc, nullptr, 10);
        const char *nlat = std::getenv("NK_LATENCY_POINTS");
        if (nlat) s->latency_points = std::strtol(nlat, nullptr, 10);
        const char *nt = std::getenv("NK_NO_TAIL");
        if (nt && nt[0] == '1') s->tail_start[0] = s->tail_start[1] = -1;
        // the LDS tail runs the rediscretized operator: levels with Galerkin operators have a tail of their own
        if (options->coarse_operator == NK_COARSE_GALERKIN) {
            s->tail_start[options->precision != NK_PRECISION_FP64 ? 1 : 0] = -1;
            const size_t es = options->precision != NK_PRECISION_FP64 ? 4 : 8;
            const char *gb = std::getenv("NK_GTAIL_BYTES");   // tuning: LDS bytes of vectors the tail may take
            const size_t gbudget = gb ? (size_t)std::strtoul(gb, nullptr, 10) : GTAIL_BUDGET;
            size_t bytes = 0, pts = 0;
            int start = -1;
            const int nl = (int)s->lev.size();
            for (int l = nl - 1; l >= 1; l--) {
                const Level &L = s->lev[l];
                const size_t p = (size_t)(L.g.Nx + 1) * (size_t)(L.g.Ny + 1), b = p * (8 * es + 1);
                const bool wanted = gb ? bytes + b <= gbudget : (nl - l <= 2 || pts + p <= GTAIL_POINTS);
                if (!wanted || bytes + b > GTAIL_BUDGET || !L.gcst || nl - l > GTAIL_MAX) break;
                bytes += b;
                pts += p;
                start = l;
            }
            const void *gt[] = {(const void *)k_g_tail<double, 0, 0>, (const void *)k_g_tail<double, 0, 1>,
                                (const void *)k_g_tail<double, 1, 0>, (const void *)k_g_tail<double, 1, 1>,
                                (const void *)k_g_tail<float, 0, 0>, (const void *)k_g_tail<float, 0, 1>,
                                (const void *)k_g_tail<float, 1, 0>, (const void *)k_g_tail<float, 1, 1>};
            bool ok = true;
            for (const void *f : gt) ok &= hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) == hipSuccess;
            // the dense inverse eliminates systems of up to GDENSE_LDS unknowns in LDS
            const void *gi[] = {(const void *)k_g_inverse<double, 0, 0>, (const void *)k_g_inverse<double, 0, 1>,
                                (const void *)k_g_inverse<double, 1, 0>, (const void *)k_g_inverse<double, 1, 1>,
                                (const void *)k_g_inverse<float, 0, 0>, (const void *)k_g_inverse<float, 0, 1>,
                                (const void *)k_g_inverse<float, 1, 0>, (const void *)k_g_inverse<float, 1, 1>};
            for (const void *f : gi)   // n² doubles beside its 9 KB of static shared memory
                if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(sizeof(double) * GDENSE_LDS * GDENSE_LDS)) != hipSuccess) {
                    s->err = "hipFuncSetAttribute (dense coarsest inverse)";
                    return fail(NK_DEVICE_ERROR);
                }
            if (ok && start >= 1 && (int)s->lev.size() - start >= 2 && !(nt && nt[0] == '1')) s->gtail_start = start;
        }
    }
    {
        long nb = (long)((s->g.Nx + BX - 1) / BX) * (long)((s->g.Ny + BY - 1) / BY);
        if ((st = alloc_bytes(s, (void **)&s->partials, 2 * sizeof(double) * MAXRED * (size_t)nb)) != NK_OK) return fail(st);
        s->partials_xr = s->partials + MAXRED * nb;
        const char *nf = std::getenv("NK_NO_FOLD");
        const char *fbk = std::getenv("NK_FOLD_BLOCKS");   // tuning
        if (fbk) s->fold_blocks = std::strtol(fbk, nullptr, 10);
        s->fold = !(nf && nf[0] == '1') && nb <= s->fold_blocks;
        if (options->coarse_operator == NK_COARSE_GALERKIN && (st = alloc_bytes(s, (void **)&s->gerr, 64)) != NK_OK) return fail(st);
        if (options->coarse_operator == NK_COARSE_GALERKIN && s->lev.size() > 1) {
            const KG &gc = s->lev.back().g;
            const size_t n = 2 * (size_t)gc.Nx * (size_t)gc.Ny;
            if (n <= (size_t)GDENSE_MAX && (st = alloc_bytes(s, &s->cinv, n * n * sizeof(double))) != NK_OK) return fail(st);
        }
        if ((st = alloc_bytes(s, (void **)&s->S, sizeof(Scal))) != NK_OK) return fail(st);
        if (hipHostMalloc((void **)&s->S_host, sizeof(Scal), hipHostMallocDefault) != hipSuccess) { s->err = "hipHostMalloc failed"; return fail(NK_OUT_OF_MEMORY); }
    }
    if (hipStreamSynchronize(s->stream) != hipSuccess) { s->err = "clearing device buffers failed"; return fail(NK_DEVICE_ERROR); }
    *out = s;
    return NK_OK;
}

namespace {

// Device bytes nk_create allocates for these arguments: the same planes, levels and work vectors, counted.
size_t planned_bytes(const nk_grid *grid, const nk_solver_options *o) {
    const KG g0 = make_kg(grid->Nx, grid->Ny, grid->topology_x, grid->topology_y, grid->dx, grid->dy);
    const size_t p0 = (size_t)plane_of(g0), Ns = (size_t)grid->Nsigma;
    const bool mixed = o->precision != NK_PRECISION_FP64;
    size_t b = 0;
    for (int id = 0; id < NK_FIELD_COUNT; id++) b += sizeof(double) * p0 * (LOC[id].sigma ? Ns : 1);
    b += 2 * p0;                                                              // face classes
    b += sizeof(double) * ((Ns + 1) + 3 * Ns + Ns * Ns);                      // σ tables
    std::vector<KG> lev{g0};
    if (o->preconditioner == NK_PRECONDITIONER_MULTIGRID) {
        KG cur = g0;
        for (;;) {
            if (o->coarse_operator == NK_COARSE_GALERKIN && lev.size() >= 2 && 2 * cur.Nx * cur.Ny <= GDENSE_STOP) break;
            int cxh = coarsening(cur.Nx, grid->halo, o->minimum_coarse_size);
            int cyh = coarsening(cur.Ny, grid->halo, o->minimum_coarse_size);
            if (!cxh && !cyh) break;
            cur = make_kg(cur.Nx / (cxh ? 2 : 1), cur.Ny / (cyh ? 2 : 1), cur.tx, cur.ty, cur.dx, cur.dy);
            lev.push_back(cur);
        }
    }
    for (size_t l = 0; l < lev.size(); l++) {
        const size_t p = (size_t)plane_of(lev[l]);
        if (l > 0) b += 4 * sizeof(double) * p;
        b += p + 9 * sizeof(double) * p;                                      // activity flags, FP64 vectors
        if (mixed) b += (4 + 9) * sizeof(float) * p;
        if (o->coarse_operator == NK_COARSE_GALERKIN) {
            const size_t es = mixed ? sizeof(float) : sizeof(double);
            b += GS_ALLOC * es * p;
            if (wants_gcst(lev[l], (int)l, es)) b += (size_t)(lev[l].Nx + 1) * (size_t)(lev[l].Ny + 1) * GS_ALLOC * es;
            if (wants_gpack(lev[l], (int)l)) b += GP_PLANES * es * (size_t)gp_plane(lev[l]);
        }
    }
    b += (24 + 16) * sizeof(double) * p0 + 3 * Ns * sizeof(double) * p0;
    b += sizeof(uint16_t) * p0;                                               // neighbour masks of the finest level
    if (mixed) b += 14 * sizeof(float) * p0;
    if (o->precision == NK_PRECISION_MIXED_REFINED) b += 2 * sizeof(float) * p0;
    const size_t nb = (size_t)((g0.Nx + BX - 1) / BX) * (size_t)((g0.Ny + BY - 1) / BY);
    b += 2 * sizeof(double) * MAXRED * nb + sizeof(Scal);                     // two slot buffers (partials, partials_xr)
    if (o->coarse_operator == NK_COARSE_GALERKIN) {
        b += 64;
        const size_t n = 2 * (size_t)lev.back().Nx * (size_t)lev.back().Ny;
        if (lev.size() > 1 && n <= (size_t)GDENSE_MAX) b += n * n * sizeof(double);
    }
    return b;
}

}  // namespace

nk_status nk_device_bytes(const nk_solver *s, int64_t *bytes) {
    if (!s || !bytes) return NK_INVALID_ARGUMENT;
    *bytes = (int64_t)s->device_bytes;
    return NK_OK;
}

nk_status nk_estimate_device_bytes(const nk_grid *grid, const nk_physics *physics, const nk_solver_options *options,
                                   int64_t *bytes) {
    if (!grid || !physics || !options || !bytes) return NK_INVALID_ARGUMENT;
    std::string err;
    nk_status st = validate(grid, physics, options, err);
    if (st != NK_OK) return st;
    *bytes = (int64_t)planned_bytes(grid, options);
    return NK_OK;
}

nk_status nk_destroy(nk_solver *s) {
    if (!s) return NK_OK;
    std::lock_guard<std::mutex> lk(g_device_mu);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (void *p : s->allocs) (void)hipFree(p);
    if (s->tape) (void)hipFree(s->tape);
    if (s->S_host) (void)hipHostFree(s->S_host);
    if (s->chunk_graph) (void)hipGraphExecDestroy(s->chunk_graph);
    for (hipEvent_t e : s->prof_events) (void)hipEventDestroy(e);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return NK_OK;
}

const char *nk_last_error(const nk_solver *s) { return s ? s->err.c_str() : g_free_err.c_str(); }

nk_status nk_set_physics(nk_solver *s, const nk_physics *physics) {
    if (!s || !physics) return NK_INVALID_ARGUMENT;
    nk_status st = validate(&s->grid, physics, &s->op, s->err);
    if (st != NK_OK) return st;
    // The captured PCG chunk holds the wall factor in its kernel arguments: a change of physics discards it.
    if (std::memcmp(&s->ph, physics, sizeof(nk_physics)) != 0) drop_chunk_graph(s);
    s->ph = *physics;
    return NK_OK;
}

nk_status nk_set_options(nk_solver *s, const nk_solver_options *o) {
    if (!s || !o) return NK_INVALID_ARGUMENT;
    nk_status st = validate(&s->grid, &s->ph, o, s->err);
    if (st != NK_OK) return st;
    // nk_create sized the multigrid hierarchy, the precision's buffers, the Galerkin stencils and the LDS tails for
    // its options. A handle created with the multigrid preconditioner may switch to the diagonal one (or none) and
    // back; every other change of these needs a new handle.
    const nk_solver_options &c = s->op0;
    const bool mg0 = c.preconditioner == NK_PRECONDITIONER_MULTIGRID, mg = o->preconditioner == NK_PRECONDITIONER_MULTIGRID;
    if ((mg && !mg0) || o->precision != c.precision ||
        (mg0 && (o->minimum_coarse_size != c.minimum_coarse_size || o->smoother != c.smoother ||
                 o->coarse_operator != c.coarse_operator)) ||
        (!mg && o->coarse_operator != NK_COARSE_REDISCRETIZED)) {
        s->err = "the multigrid hierarchy, smoother, coarse operators and precision are fixed at nk_create";
        return NK_UNSUPPORTED;
    }
    s->op = *o;
    drop_chunk_graph(s);
    return NK_OK;
}

nk_status nk_field_shape(const nk_solver *s, nk_field_id id, int32_t shape[3]) {
    if (!s || !shape || check_field(s, id) != NK_OK) return NK_INVALID_ARGUMENT;
    field_shape(s, id, shape);
    return NK_OK;
}

nk_status nk_set_field(nk_solver *s, nk_field_id id, const double *data, int64_t ld, int64_t lds) {
    if (!s || !data || check_field(s, id) != NK_OK) return NK_INVALID_ARGUMENT;
    int32_t sh[3];
    field_shape(s, id, sh);
    if (ld == 0) ld = sh[0];
    if (lds == 0) lds = ld * sh[1];
    if (ld < sh[0] || lds < ld * sh[1]) { s->err = "leading dimension smaller than the field extent"; return NK_INVALID_ARGUMENT; }
    for (int k = 0; k < sh[2]; k++)
        HIPCHK(hipMemcpy2DAsync(s->f[id] + (long)k * s->plane, sizeof(double) * s->g.pitch, data + (long)k * lds, sizeof(double) * ld,
                                sizeof(double) * sh[0], sh[1], hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (id == NK_FIELD_ETAH) s->flags_valid = false;
    return NK_OK;
}

nk_status nk_get_field(nk_solver *s, nk_field_id id, double *data, int64_t ld, int64_t lds) {
    if (!s || !data || check_field(s, id) != NK_OK) return NK_INVALID_ARGUMENT;
    int32_t sh[3];
    field_shape(s, id, sh);
    if (ld == 0) ld = sh[0];
    if (lds == 0) lds = ld * sh[1];
    if (ld < sh[0] || lds < ld * sh[1]) { s->err = "leading dimension smaller than the field extent"; return NK_INVALID_ARGUMENT; }
    for (int k = 0; k < sh[2]; k++)
        HIPCHK(hipMemcpy2DAsync(data + (long)k * lds, sizeof(double) * ld, s->f[id] + (long)k * s->plane, sizeof(double) * s->g.pitch,
                                sizeof(double) * sh[0], sh[1], hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

void nk_internal_field_extents(const nk_grid *grid, const nk_physics *physics, int id, int32_t shape[3]) {
    const LocInfo &L = LOC[id];
    shape[0] = grid->Nx + ((L.lx && grid->topology_x == NK_BOUNDED) ? 1 : 0);
    shape[1] = grid->Ny + ((L.ly && grid->topology_y == NK_BOUNDED) ? 1 : 0);
    shape[2] = (L.sigma == 1 || (L.sigma == 2 && physics->rate_factor_depth == NK_RATE_FACTOR_3D)) ? grid->Nsigma : 1;
}

nk_status nk_internal_reset_fields(nk_solver *s, const uint8_t *keep) {
    for (int id = 0; id < NK_FIELD_COUNT; id++) {
        if (keep && keep[id]) continue;
        const int nz = LOC[id].sigma ? s->Ns : 1;
        for (int k = 0; k < nz; k++) {
            k_fill2<double><<<blocks_for(s->g.Nx + 1, s->g.Ny + 1), THREADS, 0, s->stream>>>(s->g, s->f[id] + (long)k * s->plane, nullptr,
                                                                                             0.0, nullptr);
            CHK(launch_check(s));
        }
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    s->flags_valid = false;
    return NK_OK;
}

nk_status nk_fill_field(nk_solver *s, nk_field_id id, double value) {
    if (!s || check_field(s, id) != NK_OK) return NK_INVALID_ARGUMENT;
    int32_t sh[3];
    field_shape(s, id, sh);
    for (int k = 0; k < sh[2]; k++) {
        k_fill2<double><<<blocks_for(s->g.Nx + 1, s->g.Ny + 1), THREADS, 0, s->stream>>>(s->g, s->f[id] + (long)k * s->plane, nullptr,
                                                                                         value, nullptr);
        CHK(launch_check(s));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    if (id == NK_FIELD_ETAH) s->flags_valid = false;
    return NK_OK;
}

nk_status nk_get_classes(nk_solver *s, uint8_t *cu, int64_t ldu, uint8_t *cv, int64_t ldv) {
    if (!s || !cu || !cv) return NK_INVALID_ARGUMENT;
    int32_t su[3], sv[3];
    field_shape(s, NK_FIELD_U, su);
    field_shape(s, NK_FIELD_V, sv);
    if (ldu == 0) ldu = su[0];
    if (ldv == 0) ldv = sv[0];
    HIPCHK(hipMemcpy2DAsync(cu, ldu, s->cls_u, s->g.pitch, su[0], su[1], hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpy2DAsync(cv, ldv, s->cls_v, s->g.pitch, sv[0], sv[1], hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_field_device_pointer(nk_solver *s, nk_field_id id, double **ptr, int64_t *pitch, int64_t *layer) {
    if (!s || check_field(s, id) != NK_OK) return NK_INVALID_ARGUMENT;
    if (ptr) *ptr = s->f[id];
    if (pitch) *pitch = s->g.pitch;
    if (layer) *layer = s->plane;
    return NK_OK;
}

void *nk_stream(nk_solver *s) { return s ? (void *)s->stream : nullptr; }

nk_status nk_synchronize(nk_solver *s) {
    if (!s) return NK_INVALID_ARGUMENT;
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_update_geometry(nk_solver *s) {
    if (!s) return NK_INVALID_ARGUMENT;
    double **f = s->f;
    k_geometry<<<blocks_for(s->g.Nx + 1, s->g.Ny + 1), THREADS, 0, s->stream>>>(s->g, physk(s), f[NK_FIELD_H], f[NK_FIELD_BED],
                                                                                f[NK_FIELD_S], f[NK_FIELD_PHI], f[NK_FIELD_PHIU],
                                                                                f[NK_FIELD_PHIV], f[NK_FIELD_N], f[NK_FIELD_MU]);
    CHK(launch_check(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_compute_driving_stress(nk_solver *s) {
    if (!s) return NK_INVALID_ARGUMENT;
    CHK(driving_stress(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_update_coefficients(nk_solver *s) {
    if (!s) return NK_INVALID_ARGUMENT;
    CHK(update_coefficients(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_apply_operator(nk_solver *s) {
    if (!s) return NK_INVALID_ARGUMENT;
    CHK(corners_flags(s, 0, false));
    Level &L = s->lev[0];
    launch_apply0<double>(s, coef<double>(L), s->ph.walls == NK_NO_SLIP ? 2.0 : 0.0, s->f[NK_FIELD_PU], s->f[NK_FIELD_PV],
                          s->f[NK_FIELD_QU], s->f[NK_FIELD_QV], nullptr, nullptr, red(s, FIN_NONE), 0, nullptr);
    CHK(launch_check(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_precondition(nk_solver *s) {
    if (!s) return NK_INVALID_ARGUMENT;
    if (s->op.precision != NK_PRECISION_FP64) { s->err = "nk_precondition runs the FP64 preconditioner only"; return NK_UNSUPPORTED; }
    CHK(corners_flags(s, 0, false));
    const nk_control fixed0 = {NK_FIXED_ITERATIONS, 0, 0, 0, 0.0, 0.0};
    CHK(write_scalars(s, fixed0));  // done = 0: no gating
    CHK(update_preconditioner(s));
    Level &L = s->lev[0];
    const KG &g = s->g;
    const dim3 B = blocks_for(g.Nx, g.Ny);
    double *ru = s->f[NK_FIELD_PU], *rv = s->f[NK_FIELD_PV], *zu = s->f[NK_FIELD_QU], *zv = s->f[NK_FIELD_QV];
    const int pre = s->op.preconditioner;
    if (pre == NK_PRECONDITIONER_MULTIGRID) {
        Pair<double> z;
        CHK(v_cycle<double>(s, 0, Pair<double>{s->w64.zu, s->w64.zv}, Pair<double>{s->w64.z2u, s->w64.z2v}, ru, rv, false, nullptr, &z));
        k_copy2<double><<<B, THREADS, 0, s->stream>>>(g, z.u, z.v, zu, zv);
        CHK(launch_check(s));
    } else {
        NK_LAUNCH_TOPO(g, k_update_diag, double, B, THREADS, 0, s->stream, g, coef<double>(L), s->ph.walls == NK_NO_SLIP ? 2.0 : 0.0, pre, nullptr, nullptr,
                                                             nullptr, nullptr, nullptr, nullptr, ru, rv, zu, zv, s->w64.pu, s->w64.pv, 1,
                                                             red(s, FIN_NONE), nullptr, nullptr, 0);
        CHK(launch_check(s));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_solve(nk_solver *s, nk_info *info) {
    if (!s) return NK_INVALID_ARGUMENT;
    Timer t(s, info);
    if (info) info->total_linear_iterations = 0;
    nk_status st = corners_flags(s, 0, false);
    if (st == NK_OK) st = solve_linear(s, info);
    if (st == NK_OK && s->op.report_not_converged && s->op.linear.kind == NK_CONVERGENCE && !s->S_host->done_by_tol) {
        s->err = "PCG stopped at maxiter before meeting its tolerance";
        st = NK_NOT_CONVERGED;
    }
    return t.done(st);
}

nk_status nk_picard_solve(nk_solver *s, nk_info *info) {
    if (!s) return NK_INVALID_ARGUMENT;
    Timer t(s, info);
    return t.done(picard(s, info));
}

nk_status nk_compute_velocity_profile(nk_solver *s) {
    if (!s) return NK_INVALID_ARGUMENT;
    CHK(post_process(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_update_velocities(nk_solver *s, int32_t geometry, nk_info *info) {
    if (!s) return NK_INVALID_ARGUMENT;
    Timer t(s, info);
    nk_status st = NK_OK;
    if (geometry) {
        double **f = s->f;
        k_geometry<<<blocks_for(s->g.Nx + 1, s->g.Ny + 1), THREADS, 0, s->stream>>>(s->g, physk(s), f[NK_FIELD_H], f[NK_FIELD_BED],
                                                                                    f[NK_FIELD_S], f[NK_FIELD_PHI], f[NK_FIELD_PHIU],
                                                                                    f[NK_FIELD_PHIV], f[NK_FIELD_N], f[NK_FIELD_MU]);
        st = launch_check(s);
    }
    if (st == NK_OK) st = driving_stress(s);
    if (st == NK_OK) st = picard(s, info);
    if (st == NK_OK) st = post_process(s);
    return t.done(st);
}

nk_status nk_compute_thickness_tendency(nk_solver *s, int32_t advection, int32_t grounded_melt_coupling) {
    if (!s) return NK_INVALID_ARGUMENT;
    if (advection < NK_ADVECTION_UPWIND1 || advection > NK_ADVECTION_WENO5_BOUNDED) {
        s->err = "thickness advection: NK_ADVECTION_UPWIND1, _CENTERED2, _WENO5 or _WENO5_BOUNDED";
        return NK_INVALID_ARGUMENT;
    }
    double **f = s->f;
    const double *theta = nullptr;
    if (advection == NK_ADVECTION_WENO5_BOUNDED) {
        if (!s->weno_theta) {
            std::lock_guard<std::mutex> lk(g_device_mu);
            nk_status st = alloc_plane(s, &s->weno_theta, s->g);
            if (st != NK_OK) return st;
        }
        // update_bounds_preserving_limiter! before the tendency (src/Models/time_stepping.jl:37)
        k_weno_limiter<<<blocks_for(s->g.Nx, s->g.Ny), THREADS, 0, s->stream>>>(s->g, f[NK_FIELD_H], s->thick_lower, s->thick_upper,
                                                                                   s->weno_theta);
        CHK(launch_check(s));
        theta = s->weno_theta;
    }
    k_thickness_tendency<<<blocks_for(s->g.Nx, s->g.Ny), THREADS, 0, s->stream>>>(
        s->g, advection, s->ph.grounding != NK_GROUNDING_NONE ? 1 : 0, grounded_melt_coupling ? 1 : 0, f[NK_FIELD_U], f[NK_FIELD_V],
        f[NK_FIELD_H], f[NK_FIELD_SMB], f[NK_FIELD_BASAL_MELT], f[NK_FIELD_PHI], f[NK_FIELD_GROUNDED_MELT], f[NK_FIELD_GH], theta);
    CHK(launch_check(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_set_thickness_bounds(nk_solver *s, double lower, double upper) {
    if (!s) return NK_INVALID_ARGUMENT;
    if (!(lower <= upper) || std::isnan(lower) || std::isnan(upper)) {
        s->err = "thickness bounds: lower ≤ upper";
        return NK_INVALID_ARGUMENT;
    }
    s->thick_lower = lower;
    s->thick_upper = upper;
    return NK_OK;
}

nk_status nk_thickness_substep(nk_solver *s, double dtau, int32_t final_stage) {
    if (!s) return NK_INVALID_ARGUMENT;
    double **f = s->f;
    k_thickness_substep<<<blocks_for(s->g.Nx, s->g.Ny), THREADS, 0, s->stream>>>(s->g, dtau, final_stage ? 1 : 0, f[NK_FIELD_HPREV],
                                                                                   f[NK_FIELD_GH], f[NK_FIELD_H], f[NK_FIELD_HCLIP]);
    CHK(launch_check(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_cache_thickness(nk_solver *s) {
    if (!s) return NK_INVALID_ARGUMENT;
    k_copy_plane<<<blocks_for(s->g.Nx + 1, s->g.Ny + 1), THREADS, 0, s->stream>>>(s->g, s->f[NK_FIELD_H], s->f[NK_FIELD_HPREV]);
    CHK(launch_check(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_cell_advection_timescale(nk_solver *s, double *timescale) {
    if (!s || !timescale) return NK_INVALID_ARGUMENT;
    const dim3 B = blocks_for(s->g.Nx, s->g.Ny);
    const size_t nb = (size_t)B.x * B.y;
    k_advection_timescale<<<B, THREADS, 0, s->stream>>>(s->g, s->f[NK_FIELD_U], s->f[NK_FIELD_V], s->partials);
    CHK(launch_check(s));
    std::vector<double> mins(nb);
    HIPCHK(hipMemcpyAsync(mins.data(), s->partials, nb * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    double t = INFINITY;
    for (double b : mins) t = std::fmin(t, b);
    *timescale = t;
    return NK_OK;
}

nk_status nk_compute_strain_heating(nk_solver *s) {
    if (!s) return NK_INVALID_ARGUMENT;
    double **f = s->f;
    const bool diva = s->ph.stress_balance == NK_DIVA;
    k_strain_heating<<<blocks_for(s->g.Nx, s->g.Ny), THREADS, 0, s->stream>>>(
        s->g, sig_tables(s), s->Ns, s->plane, s->ph.walls == NK_NO_SLIP ? 2.0 : 0.0, diva ? 1 : 0, f[NK_FIELD_U], f[NK_FIELD_V],
        f[NK_FIELD_ETASIGMA], f[NK_FIELD_UB], f[NK_FIELD_VB], f[NK_FIELD_BETABU], f[NK_FIELD_BETABV], f[NK_FIELD_STRAIN_HEATING]);
    CHK(launch_check(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_till_effective_pressure(nk_solver *s, const nk_till *till) {
    if (!s || !till) return NK_INVALID_ARGUMENT;
    double **f = s->f;
    const PhysK p = physk(s);
    k_till_pressure<<<blocks_for(s->g.Nx, s->g.Ny), THREADS, 0, s->stream>>>(s->g, p.rig, p.rwg, *till, f[NK_FIELD_H], f[NK_FIELD_BED],
                                                                             f[NK_FIELD_HW], f[NK_FIELD_N]);
    CHK(launch_check(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_energy_step(nk_solver *s, double dt, int32_t grounded_melt_coupling, const nk_till *till) {
    if (!s) return NK_INVALID_ARGUMENT;
    if (!(dt > 0.0) || s->Ns < 1) { s->err = "nk_energy_step: dt must be positive and the grid needs σ layers"; return NK_INVALID_ARGUMENT; }
    if (!s->en_W) {
        std::lock_guard<std::mutex> lk(g_device_mu);
        nk_status st;
        if ((st = alloc_plane(s, &s->en_W, s->g, s->Ns + 1)) != NK_OK) return st;
        double **w[] = {&s->en_a, &s->en_b, &s->en_c, &s->en_f, &s->en_t};
        for (auto p : w) if ((st = alloc_plane(s, p, s->g, s->Ns)) != NK_OK) return st;
        if ((st = alloc_plane(s, &s->en_bas, s->g, 4)) != NK_OK) return st;
    }
    double **f = s->f;
    const bool diva = s->ph.stress_balance == NK_DIVA;
    EnergyArgs e;
    e.rho = s->ph.ice_density;
    e.grav = s->ph.gravitational_acceleration;
    e.dt = dt;
    e.grounding = s->ph.grounding != NK_GROUNDING_NONE ? 1 : 0;
    e.coupled = grounded_melt_coupling ? 1 : 0;
    e.Ns = s->Ns;
    e.a3d = s->ph.rate_factor_depth == NK_RATE_FACTOR_3D ? 1 : 0;
    e.plane = s->plane;
    e.H = f[NK_FIELD_H]; e.Hn = f[NK_FIELD_HN]; e.us = f[NK_FIELD_USIGMA]; e.vs = f[NK_FIELD_VSIGMA];
    e.ub = f[NK_FIELD_UB]; e.vb = f[NK_FIELD_VB];
    e.bu = diva ? f[NK_FIELD_BETABU] : f[NK_FIELD_BETAU];
    e.bv = diva ? f[NK_FIELD_BETABV] : f[NK_FIELD_BETAV];
    e.phi = f[NK_FIELD_PHI]; e.m = f[NK_FIELD_BASAL_MELT]; e.G = f[NK_FIELD_GEOTHERMAL]; e.Ts = f[NK_FIELD_TSURF];
    e.To = f[NK_FIELD_TOCEAN]; e.Phi = f[NK_FIELD_STRAIN_HEATING];
    e.E = f[NK_FIELD_ENTHALPY]; e.Hw = f[NK_FIELD_HW]; e.mg = f[NK_FIELD_GROUNDED_MELT]; e.T = f[NK_FIELD_TEMPERATURE];
    e.W = f[NK_FIELD_WATER_FRACTION]; e.A = f[NK_FIELD_A];
    e.W3 = s->en_W; e.a = s->en_a; e.b = s->en_b; e.c = s->en_c; e.f = s->en_f; e.t = s->en_t; e.bas = s->en_bas;
    e.till = till ? 1 : 0;
    e.till_rate = till ? till->drainage_rate : 0.0;
    e.till_wmax = till ? till->maximum_water_thickness : 0.0;
    const dim3 B = blocks_for(s->g.Nx, s->g.Ny);
    k_energy_assemble<<<B, THREADS, 0, s->stream>>>(s->g, sig_tables(s), e);
    CHK(launch_check(s));
    k_energy_solve<<<B, THREADS, 0, s->stream>>>(s->g, sig_tables(s), e, f[NK_FIELD_HN]);
    CHK(launch_check(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_thermal_diagnostics(nk_solver *s) {
    if (!s) return NK_INVALID_ARGUMENT;
    if (s->ph.rate_factor_depth != NK_RATE_FACTOR_3D) { s->err = "the rate factor must be on the σ grid (NK_RATE_FACTOR_3D)"; return NK_INVALID_ARGUMENT; }
    k_thermal_diagnostics<<<blocks_for(s->g.Nx, s->g.Ny), THREADS, 0, s->stream>>>(
        s->g, sig_tables(s), s->Ns, s->plane, s->ph.ice_density * s->ph.gravitational_acceleration, s->f[NK_FIELD_ENTHALPY],
        s->f[NK_FIELD_H], s->f[NK_FIELD_TEMPERATURE], s->f[NK_FIELD_WATER_FRACTION], s->f[NK_FIELD_A]);
    CHK(launch_check(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

nk_status nk_rate_factor_from_temperature(nk_solver *s) {
    if (!s) return NK_INVALID_ARGUMENT;
    if (s->ph.rate_factor_depth != NK_RATE_FACTOR_3D) { s->err = "the rate factor must be on the σ grid (NK_RATE_FACTOR_3D)"; return NK_INVALID_ARGUMENT; }
    k_rate_factor<<<blocks_for(s->g.Nx, s->g.Ny), THREADS, 0, s->stream>>>(s->g, sig_tables(s), s->Ns, s->plane,
                                                                           s->ph.ice_density * s->ph.gravitational_acceleration,
                                                                           s->f[NK_FIELD_TEMPERATURE], s->f[NK_FIELD_H], s->f[NK_FIELD_A],
                                                                           s->f[NK_FIELD_ENTHALPY]);
    CHK(launch_check(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    return NK_OK;
}

namespace {

// Colour classes of one direction for the scatter kernels of nk_adjoint.hpp, whose thread at point i adds into
// points i − 1 .. i + 1 (after the clamp or wrap), over points 1 .. extent. Bounded: the points ≡ 1, 2, 0 (mod 3).
// Periodic (only points 1 .. N write anything): the first M = 3⌊N/3⌋ points in three classes of stride 3 — two points
// of one class are then at least 3 apart also across the wrap, the N − M remaining points lying between the class's
// last and first — and each of the N − M remaining points in a class of its own.
extern "C++" {
struct DirClasses {
    int n = 0;
    int o[5], s[5], c[5];
};
DirClasses adjoint_classes(int N, int extent, bool periodic) {
    DirClasses d;
    const int span = periodic ? 3 * (N / 3) : extent;
    for (int k = 0; k < 3 && k < span; k++) {
        d.o[d.n] = 1 + k;
        d.s[d.n] = 3;
        d.c[d.n] = (span - 1 - k) / 3 + 1;
        d.n++;
    }
    if (periodic)
        for (int k = span; k < N; k++) {
            d.o[d.n] = k + 1;
            d.s[d.n] = 1;
            d.c[d.n] = 1;
            d.n++;
        }
    return d;
}

// Runs launch(blocks, colour) once per colour class of the points 1..ex × 1..ey, in a fixed order on the handle's
// stream.
template <typename Launch>
nk_status coloured(nk_solver *s, int ex, int ey, Launch launch) {
    const DirClasses X = adjoint_classes(s->g.Nx, ex, s->g.tx == NK_PERIODIC);
    const DirClasses Y = adjoint_classes(s->g.Ny, ey, s->g.ty == NK_PERIODIC);
    for (int b = 0; b < Y.n; b++)
        for (int a = 0; a < X.n; a++) {
            const Colour col{X.o[a], X.s[a], X.c[a], Y.o[b], Y.s[b], Y.c[b]};
            launch(blocks_for(col.nx, col.ny), col);
            CHK(launch_check(s));
        }
    return NK_OK;
}
}   // extern "C++"

// One transposed coefficient step: ȳ = (∂q/∂x)ᵀ q̄ (velocity) and/or C̄, Ā (parameters), with q̄ in ad_q. q(x) =
// A(c(x)) xs is linearized at x = (U, V) with xs = (xsu, xsv): the same vector at the fixed point (null: U, V), the step's
// result x_k at the input x_{k−1} of an unrolled Picard step.
nk_status adjoint_step(nk_solver *s, bool velocity, bool parameters, const double *xsu = nullptr, const double *xsv = nullptr) {
    const KG &g = s->g;
    const dim3 Bc = blocks_for(g.Nx, g.Ny), Be = blocks_for(g.Nx + 1, g.Ny + 1);
    double **f = s->f;
    if (!xsu) {
        xsu = f[NK_FIELD_U];
        xsv = f[NK_FIELD_V];
    }
    const PhysK p = physk(s);
    const uint8_t *act = s->lev[0].act;
    k_fill2<double><<<Be, THREADS, 0, s->stream>>>(g, s->ad_eta, s->ad_exy, 0.0, nullptr);
    CHK(launch_check(s));
    if (velocity) {
        k_fill2<double><<<Be, THREADS, 0, s->stream>>>(g, s->ad_yu, s->ad_yv, 0.0, nullptr);
        CHK(launch_check(s));
    }
    k_adj_centres<<<Bc, THREADS, 0, s->stream>>>(g, act, xsu, xsv, s->ad_qu, s->ad_qv, s->ad_eta);
    CHK(launch_check(s));
    CHK(coloured(s, g.Nx + 1, g.Ny + 1, [&](dim3 B, Colour col) {
        k_adj_corners<<<B, THREADS, 0, s->stream>>>(g, col, act, p.wf, xsu, xsv, s->ad_qu, s->ad_qv, s->ad_eta);
    }));
    CHK(coloured(s, g.Nx, g.Ny, [&](dim3 B, Colour col) {
        k_adj_viscosity<<<B, THREADS, 0, s->stream>>>(g, col, p, sig_tables(s), f[NK_FIELD_U], f[NK_FIELD_V], f[NK_FIELD_ETAH],
                                                      f[NK_FIELD_ETASIGMA], s->plane, f[NK_FIELD_A], f[NK_FIELD_H],
                                                      s->ph.masking ? f[NK_FIELD_MU] : nullptr, s->ad_eta, s->ad_yu, s->ad_yv,
                                                      s->ad_exy, f[NK_FIELD_ABAR], velocity ? 1 : 0, parameters ? 1 : 0);
    }));
    if (velocity)
        CHK(coloured(s, g.Nx + 1, g.Ny + 1, [&](dim3 B, Colour col) {
            k_adj_shear<<<B, THREADS, 0, s->stream>>>(g, col, p.wf, s->ad_exy, s->ad_yu, s->ad_yv);
        }));
    return coloured(s, g.Nx + 1, g.Ny + 1, [&](dim3 B, Colour col) {
        k_adj_drag<<<B, THREADS, 0, s->stream>>>(g, col, p, act, f[NK_FIELD_U], f[NK_FIELD_V], s->ad_qu, s->ad_qv, f[NK_FIELD_C],
                                                 f[NK_FIELD_N], f[NK_FIELD_PHIU], f[NK_FIELD_PHIV], s->ad_yu, s->ad_yv, f[NK_FIELD_CBAR],
                                                 velocity ? 1 : 0, parameters ? 1 : 0, nullptr, nullptr, xsu, xsv);
    });
}

// The lagged DIVA state a step read (βᵇ at u and v faces, F₂, ησ): null members are the handle's fields, which hold it
// at the fixed point; an unrolled step passes its input state from the tape.
struct DivaLag {
    const double *bbu = nullptr, *bbv = nullptr, *F2 = nullptr, *es = nullptr;
};

// One transposed DIVA coefficient step (update_diva_coefficients!, src/StressBalances/diva.jl:60-75), q̄ in ad_q and the
// shadows of the step's state (βᵇ′, F₂′, ησ′) in st_*: ȳ in ad_y, the lagged state's shadows in out_*, and C̄, Ā in the
// parameter pass. The whole chain runs in both passes (C̄ needs βᵇ̄ from the basal-shear terms of the viscosity).
// The step is linearized at (U, V) with the lagged state lag, and (xsu, xsv) is the vector the operator multiplies
// (null: U, V); the step's own outputs (βᵇ′, F₂′, uᵇ) are read from the fields.
nk_status adjoint_step_diva(nk_solver *s, bool parameters, const double *xsu = nullptr, const double *xsv = nullptr,
                            DivaLag lag = DivaLag{}) {
    const KG &g = s->g;
    const dim3 Bc = blocks_for(g.Nx, g.Ny), Be = blocks_for(g.Nx + 1, g.Ny + 1);
    double **f = s->f;
    if (!xsu) {
        xsu = f[NK_FIELD_U];
        xsv = f[NK_FIELD_V];
    }
    if (!lag.bbu) lag = DivaLag{f[NK_FIELD_BETABU], f[NK_FIELD_BETABV], f[NK_FIELD_F2], f[NK_FIELD_ETASIGMA]};
    const PhysK p = physk(s);
    const uint8_t *act = s->lev[0].act;
    auto fill = [&](double *a, double *b) { k_fill2<double><<<Be, THREADS, 0, s->stream>>>(g, a, b, 0.0, nullptr); return launch_check(s); };
    CHK(fill(s->ad_eta, s->ad_exy));
    CHK(fill(s->ad_yu, s->ad_yv));
    CHK(fill(s->ad_ubb, s->ad_vbb));
    CHK(fill(s->out_bbu, s->out_bbv));
    CHK(fill(s->out_F2, nullptr));
    k_copy_plane<<<Be, THREADS, 0, s->stream>>>(g, s->st_bbu, s->ad_bbu);
    CHK(launch_check(s));
    k_copy_plane<<<Be, THREADS, 0, s->stream>>>(g, s->st_bbv, s->ad_bbv);
    CHK(launch_check(s));
    k_copy_plane<<<Be, THREADS, 0, s->stream>>>(g, s->st_F2, s->ad_F2n);
    CHK(launch_check(s));
    k_adj_centres<<<Bc, THREADS, 0, s->stream>>>(g, act, xsu, xsv, s->ad_qu, s->ad_qv, s->ad_eta);
    CHK(launch_check(s));
    CHK(coloured(s, g.Nx + 1, g.Ny + 1, [&](dim3 B, Colour col) {
        k_adj_corners<<<B, THREADS, 0, s->stream>>>(g, col, act, p.wf, xsu, xsv, s->ad_qu, s->ad_qv, s->ad_eta);
    }));
    CHK(coloured(s, g.Nx + 1, g.Ny + 1, [&](dim3 B, Colour col) {
        k_adj_effective_drag<<<B, THREADS, 0, s->stream>>>(g, col, act, xsu, xsv, s->ad_qu, s->ad_qv,
                                                           f[NK_FIELD_BETABU], f[NK_FIELD_BETABV], f[NK_FIELD_F2], s->ad_bbu,
                                                           s->ad_bbv, s->ad_F2n);
    }));
    CHK(coloured(s, g.Nx, g.Ny, [&](dim3 B, Colour col) {
        k_adj_viscosity_diva<<<B, THREADS, 0, s->stream>>>(g, col, p, sig_tables(s), f[NK_FIELD_U], f[NK_FIELD_V], lag.es,
                                                           s->plane, f[NK_FIELD_A], f[NK_FIELD_H],
                                                           s->ph.masking ? f[NK_FIELD_MU] : nullptr, f[NK_FIELD_BETABU],
                                                           f[NK_FIELD_BETABV], f[NK_FIELD_UB], f[NK_FIELD_VB], s->ad_eta, s->ad_F2n,
                                                           s->st_es, s->ad_yu, s->ad_yv, s->ad_exy, s->ad_bbu, s->ad_bbv, s->ad_ubb,
                                                           s->ad_vbb, s->out_es, f[NK_FIELD_ABAR], 1, parameters ? 1 : 0);
    }));
    CHK(coloured(s, g.Nx + 1, g.Ny + 1, [&](dim3 B, Colour col) {
        k_adj_shear<<<B, THREADS, 0, s->stream>>>(g, col, p.wf, s->ad_exy, s->ad_yu, s->ad_yv);
    }));
    CHK(coloured(s, g.Nx + 1, g.Ny + 1, [&](dim3 B, Colour col) {
        k_adj_drag<<<B, THREADS, 0, s->stream>>>(g, col, p, act, f[NK_FIELD_UB], f[NK_FIELD_VB], s->ad_qu, s->ad_qv, f[NK_FIELD_C],
                                                 f[NK_FIELD_N], f[NK_FIELD_PHIU], f[NK_FIELD_PHIV], s->ad_ubb, s->ad_vbb,
                                                 f[NK_FIELD_CBAR], 1, parameters ? 1 : 0, s->ad_bbu, s->ad_bbv, f[NK_FIELD_UB],
                                                 f[NK_FIELD_VB]);
    }));
    return coloured(s, g.Nx + 1, g.Ny + 1, [&](dim3 B, Colour col) {
        k_adj_basal_velocity<<<B, THREADS, 0, s->stream>>>(g, col, f[NK_FIELD_UB], f[NK_FIELD_VB], lag.bbu, lag.bbv, lag.F2,
                                                           s->ad_ubb, s->ad_vbb, s->ad_yu, s->ad_yv, s->out_bbu, s->out_bbv,
                                                           s->out_F2);
    });
}

// The cost's pullback through DIVA's post-processing (compute_velocity_profile!, after the Picard loop) at the current
// state: x̄ = (UBAR, VBAR) + ū from the layer velocities' shadows (USIGMABAR, VSIGMABAR) into ad_xb, and the shadows
// s̄ = (βᵇ̄, F̄₂, ησ̄) the cost gives the state into cs_*.
nk_status diva_profile_pullback(nk_solver *s) {
    const KG &g = s->g;
    const dim3 B = blocks_for(g.Nx, g.Ny), Be = blocks_for(g.Nx + 1, g.Ny + 1);
    double **f = s->f;
    // the cost's pullback through the layer velocities (compute_velocity_profile!, after the Picard loop): ū joins x̄,
    // and the DIVA state gets the shadows s̄ that every fixed-point step adds to its sweep
    double *z[] = {s->cs_bbu, s->cs_bbv, s->cs_F2, s->ad_ubb, s->ad_vbb, s->ad_yu, s->ad_yv};
    for (double *a : z) {
        k_fill2<double><<<Be, THREADS, 0, s->stream>>>(g, a, nullptr, 0.0, nullptr);
        CHK(launch_check(s));
    }
    for (int k = 0; k < s->Ns; k++) {
        k_fill2<double><<<Be, THREADS, 0, s->stream>>>(g, s->cs_es + (long)k * s->plane, s->out_es + (long)k * s->plane, 0.0, nullptr);
        CHK(launch_check(s));
    }
    // uᵇ of the post-processing, at x* and the state there, into out_bbu/out_bbv (free until the first step clears them)
    double *ubp = s->out_bbu, *vbp = s->out_bbv;
    k_basal_velocity<<<Be, THREADS, 0, s->stream>>>(g, f[NK_FIELD_U], f[NK_FIELD_V], f[NK_FIELD_BETABU], f[NK_FIELD_BETABV],
                                                     f[NK_FIELD_F2], ubp, vbp);
    CHK(launch_check(s));
    CHK(coloured(s, g.Nx + 1, g.Ny + 1, [&](dim3 Bk, Colour col) {
        k_adj_velocity_profile<<<Bk, THREADS, 0, s->stream>>>(g, col, sig_tables(s), s->Ns, s->plane, f[NK_FIELD_ETASIGMA],
                                                              f[NK_FIELD_H], ubp, vbp, f[NK_FIELD_BETABU], f[NK_FIELD_BETABV],
                                                              f[NK_FIELD_USIGMABAR], f[NK_FIELD_VSIGMABAR], s->ad_ubb, s->ad_vbb,
                                                              s->cs_bbu, s->cs_bbv, s->cs_es);
    }));
    CHK(coloured(s, g.Nx + 1, g.Ny + 1, [&](dim3 Bk, Colour col) {
        k_adj_basal_velocity<<<Bk, THREADS, 0, s->stream>>>(g, col, ubp, vbp, f[NK_FIELD_BETABU], f[NK_FIELD_BETABV],
                                                            f[NK_FIELD_F2], s->ad_ubb, s->ad_vbb, s->ad_yu, s->ad_yv, s->cs_bbu,
                                                            s->cs_bbv, s->cs_F2);
    }));
    k_add<<<B, THREADS, 0, s->stream>>>(g, f[NK_FIELD_UBAR], f[NK_FIELD_VBAR], s->ad_yu, s->ad_yv, s->ad_xbu, s->ad_xbv);
    CHK(launch_check(s));
    return NK_OK;
}

// Geometry pullback of one coefficient step: H̄, b̄ from τ̄ = (lu, lv), q̄ in ad_q and η̄H̄ in ad_eta (two passes,
// faces then cells; the adjoint loop's spare planes carry the face sends), added to HBAR/BEDBAR when accumulate.
// (xsu, xsv): the vector the operator multiplies (null: U, V).
nk_status geometry_pullback(nk_solver *s, const double *lu, const double *lv, const double *xsu, const double *xsv, bool accumulate) {
    const KG &g = s->g;
    const dim3 B = blocks_for(g.Nx, g.Ny), Be = blocks_for(g.Nx + 1, g.Ny + 1);
    double **f = s->f;
    const bool diva = s->ph.stress_balance == NK_DIVA;
    const PhysK p = physk(s);
    if (!xsu) {
        xsu = diva ? f[NK_FIELD_UB] : f[NK_FIELD_U];
        xsv = diva ? f[NK_FIELD_VB] : f[NK_FIELD_V];
    }
    k_adj_geometry_faces<<<Be, THREADS, 0, s->stream>>>(
        g, p, s->lev[0].act, f[NK_FIELD_H], f[NK_FIELD_BED], f[NK_FIELD_S], s->ph.masking ? f[NK_FIELD_MU] : nullptr,
        f[NK_FIELD_N], f[NK_FIELD_C], f[NK_FIELD_PHIU], f[NK_FIELD_PHIV], diva ? f[NK_FIELD_UB] : f[NK_FIELD_U],
        diva ? f[NK_FIELD_VB] : f[NK_FIELD_V], lu, lv, s->ad_qu, s->ad_qv, diva ? s->ad_bbu : nullptr, diva ? s->ad_bbv : nullptr,
        xsu, xsv, s->ph.grounding == NK_GROUNDING_SUBGRID_DRIVING_STRESS ? 1 : 0, s->ad_yu, s->ad_ru, s->ad_lu, s->xmu, s->ad_yv,
        s->ad_rv, s->ad_lv, s->xmv);
    CHK(launch_check(s));
    k_adj_geometry_cells<<<B, THREADS, 0, s->stream>>>(g, p, sig_tables(s), f[NK_FIELD_H], f[NK_FIELD_ETASIGMA], s->plane, s->ad_eta,
                                                      diva ? s->ad_F2n : nullptr, s->ad_yu, s->ad_ru, s->ad_lu, s->xmu, s->ad_yv,
                                                      s->ad_rv, s->ad_lv, s->xmv, f[NK_FIELD_HBAR], f[NK_FIELD_BEDBAR],
                                                      accumulate ? 1 : 0);
    return launch_check(s);
}

// EnzymeRules.reverse for converged picard_solve! (ext/NunatakEnzymeExt.jl:175-240), SSA coefficients.
nk_status adjoint(nk_solver *s, const nk_control &ctl, nk_info *info) {
    const KG &g = s->g;
    const dim3 B = blocks_for(g.Nx, g.Ny), Be = blocks_for(g.Nx + 1, g.Ny + 1);
    double **f = s->f;
    double *lu = f[NK_FIELD_LAMBDAU], *lv = f[NK_FIELD_LAMBDAV];
    CHK(update_coefficients(s));   // coefficients at the fixed point x*
    CHK(update_preconditioner(s));
    k_fill2<double><<<Be, THREADS, 0, s->stream>>>(g, lu, lv, 0.0, nullptr);
    CHK(launch_check(s));
    k_fill2<double><<<Be, THREADS, 0, s->stream>>>(g, s->ad_lu, s->ad_lv, 0.0, nullptr);
    CHK(launch_check(s));
    k_fill2<double><<<Be, THREADS, 0, s->stream>>>(g, f[NK_FIELD_CBAR], nullptr, 0.0, nullptr);
    CHK(launch_check(s));
    k_fill2<double><<<Be, THREADS, 0, s->stream>>>(g, s->ad_qu, s->ad_qv, 0.0, nullptr);
    CHK(launch_check(s));
    for (int k = 0; k < (s->ph.rate_factor_depth == NK_RATE_FACTOR_3D ? s->Ns : 1); k++) {
        k_fill2<double><<<Be, THREADS, 0, s->stream>>>(g, f[NK_FIELD_ABAR] + (long)k * s->plane, nullptr, 0.0, nullptr);
        CHK(launch_check(s));
    }
    const bool diva = s->ph.stress_balance == NK_DIVA;
    const double *xbu = f[NK_FIELD_UBAR], *xbv = f[NK_FIELD_VBAR];
    if (diva) {
        CHK(diva_profile_pullback(s));
        xbu = s->ad_xbu;
        xbv = s->ad_xbv;
        // first step's state shadows: s̄ + 0
        k_copy_plane<<<Be, THREADS, 0, s->stream>>>(g, s->cs_bbu, s->st_bbu);
        CHK(launch_check(s));
        k_copy_plane<<<Be, THREADS, 0, s->stream>>>(g, s->cs_bbv, s->st_bbv);
        CHK(launch_check(s));
        k_copy_plane<<<Be, THREADS, 0, s->stream>>>(g, s->cs_F2, s->st_F2);
        CHK(launch_check(s));
        for (int k = 0; k < s->Ns; k++) {
            k_copy_plane<<<Be, THREADS, 0, s->stream>>>(g, s->cs_es + (long)k * s->plane, s->st_es + (long)k * s->plane);
            CHK(launch_check(s));
        }
    }
    const bool fixed = ctl.kind == NK_FIXED_ITERATIONS;
    double D = INFINITY, M = 0.0;
    int it = 0;
    long total = 0;
    for (;;) {
        if (fixed) { if (it >= ctl.iterations) break; }
        else if (!(D > jmax_h(ctl.reltol * M, ctl.abstol) && it < ctl.maxiter)) break;
        k_negate<<<B, THREADS, 0, s->stream>>>(g, lu, lv, s->ad_qu, s->ad_qv);   // q̄ = −λ
        CHK(launch_check(s));
        if (diva) {
            CHK(adjoint_step_diva(s, false));                                    // ȳ and the lagged state's shadows
            // next step's state shadows: s̄ + this step's sweep (ext/NunatakEnzymeExt.jl:205-212)
            double *pairs[][3] = {{s->cs_bbu, s->out_bbu, s->st_bbu}, {s->cs_bbv, s->out_bbv, s->st_bbv}, {s->cs_F2, s->out_F2, s->st_F2}};
            for (auto &q : pairs) {
                k_add_plane<<<Be, THREADS, 0, s->stream>>>(g, q[0], q[1], q[2]);
                CHK(launch_check(s));
            }
            for (int k = 0; k < s->Ns; k++) {
                const long ko = (long)k * s->plane;
                k_add_plane<<<Be, THREADS, 0, s->stream>>>(g, s->cs_es + ko, s->out_es + ko, s->st_es + ko);
                CHK(launch_check(s));
            }
        } else {
            CHK(adjoint_step(s, true, false));                                   // ȳ
        }
        k_add<<<B, THREADS, 0, s->stream>>>(g, xbu, xbv, s->ad_yu, s->ad_yv, s->ad_ru, s->ad_rv);
        CHK(launch_check(s));
        k_copy2<double><<<B, THREADS, 0, s->stream>>>(g, lu, lv, s->ad_lu, s->ad_lv);     // λ⁺ = λ (warm start)
        CHK(launch_check(s));
        int cg = 0;
        double res = 0.0;
        CHK(pcg<double>(s, s->ad_lu, s->ad_lv, s->ad_ru, s->ad_rv, s->op.linear, false, &cg, &res));   // λ⁺ = A⁻ᵀ r
        total += cg;
        k_picard_change<<<B, THREADS, 0, s->stream>>>(g, lu, lv, s->ad_lu, s->ad_lv, red(s, FIN_PICARD));   // Δ, M
        CHK(launch_check(s));
        CHK(finish_red(s, B, 2, FIN_PICARD, nullptr));
        k_copy2<double><<<B, THREADS, 0, s->stream>>>(g, s->ad_lu, s->ad_lv, lu, lv);     // λ = λ⁺
        CHK(launch_check(s));
        CHK(read_scalars(s));
        D = s->S_host->picard_D;
        M = s->S_host->picard_X;
        if (!std::isfinite(D) || !std::isfinite(M)) { s->err = "the adjoint fixed point's change or norm is not finite"; return NK_NONFINITE; }
        it++;
        if (s->verbose) {
            std::fprintf(stderr, "[nunatak_hip] adjoint %d: linear %d, change %.3e, norm %.3e\n", it, cg, D, M);
            std::fflush(stderr);
        }
    }
    if (!fixed && D > jmax_h(ctl.reltol * M, ctl.abstol))
        std::fprintf(stderr, "nunatak_hip: Picard adjoint did not converge in %d iterations: |dλ| = %g, |λ| = %g\n", ctl.maxiter, D, M);
    // parameter pullback with the converged λ: p̄ = (∂q/∂p)ᵀ(−λ)
    k_negate<<<B, THREADS, 0, s->stream>>>(g, lu, lv, s->ad_qu, s->ad_qv);
    CHK(launch_check(s));
    if (diva) CHK(adjoint_step_diva(s, true));
    else CHK(adjoint_step(s, false, true));
    CHK(geometry_pullback(s, lu, lv, nullptr, nullptr, false));
    if (info) {
        info->picard_iterations = it;
        info->total_linear_iterations = total;
        info->last_picard_change = D;
        info->last_picard_norm = M;
    }
    return NK_OK;
}


// Reverse of picard_solve! with FixedIterations(n) (src/Solvers/picard.jl:63-69), differentiated step by step as
// Enzyme does where the converged-Picard rule does not apply (that rule is only for Convergence controls,
// ext/NunatakEnzymeExt.jl:121). Step k is x_k = A(c_k)⁻¹ τ with c_k = c(x_{k−1}) (DIVA: c_k, s_k = c(x_{k−1}, s_{k−1}) with
// the lagged state s = (βᵇ, F₂, ησ)), warm-started at x_{k−1}; its solve has the implicit reverse rule of solve!
// (ext/NunatakEnzymeExt.jl:68-103): λ_k = A(c_k)⁻¹ x̄_k from zero, τ̄ += λ_k, and the coefficient pullback of
// q = A(c(x_{k−1})) x_k with q̄ = −λ_k gives x̄_{k−1} (DIVA: also s̄_{k−1} from s̄_k) and the shadows of C, A, H and b;
// the initial guess gets no shadow. The iterates come from replaying the forward steps from the initial guess that
// picard() recorded, into a tape of the n + 1 iterates (and states), allocated on first use.
nk_status adjoint_unrolled(nk_solver *s, int n, nk_info *info) {
    const KG &g = s->g;
    const dim3 B = blocks_for(g.Nx, g.Ny), Be = blocks_for(g.Nx + 1, g.Ny + 1);
    double **f = s->f;
    const bool diva = s->ph.stress_balance == NK_DIVA;
    if (!s->x0_valid) {
        s->err = "the unrolled adjoint of FixedIterations Picard replays the last FixedIterations Picard solve of this handle: run one first";
        return NK_INVALID_ARGUMENT;
    }
    const long P = plane_of(g);
    const long SP = diva ? 5L + s->Ns : 2L;   // planes per tape step: x (u, v) and, DIVA, βᵇu, βᵇv, F₂, ησ
    const long need = SP * (n + 1) + 2;
    if (s->tape_planes < need) {
        if (s->tape) {
            HIPCHK(hipStreamSynchronize(s->stream));
            (void)hipFree(s->tape);
            s->device_bytes -= (size_t)s->tape_planes * (size_t)P * sizeof(double);
            s->tape = nullptr;
            s->tape_planes = 0;
        }
        if (hipMalloc((void **)&s->tape, (size_t)need * (size_t)P * sizeof(double)) != hipSuccess) {
            s->tape = nullptr;
            s->err = "hipMalloc of the unrolled adjoint's tape failed";
            return NK_OUT_OF_MEMORY;
        }
        s->tape_planes = need;
        s->device_bytes += (size_t)need * (size_t)P * sizeof(double);
    }
    auto step = [&](int k) { return s->tape + SP * k * P; };
    auto tu = [&](int k) { return step(k); };
    auto tv = [&](int k) { return step(k) + P; };
    auto lag_of = [&](int k) { return DivaLag{step(k) + 2 * P, step(k) + 3 * P, step(k) + 4 * P, step(k) + 5 * P}; };
    double *tlu = s->tape + SP * (n + 1) * P, *tlv = tlu + P;
    auto set_state = [&](int k) -> nk_status {   // (U, V) = x_k and, DIVA, the fields of the lagged state = s_k
        k_copy2<double><<<B, THREADS, 0, s->stream>>>(g, tu(k), tv(k), f[NK_FIELD_U], f[NK_FIELD_V]);
        CHK(launch_check(s));
        if (diva) {
            const DivaLag L = lag_of(k);
            CHK(copy_diva_state(s, L.bbu, L.bbv, L.F2, L.es, f[NK_FIELD_BETABU], f[NK_FIELD_BETABV], f[NK_FIELD_F2],
                                f[NK_FIELD_ETASIGMA]));
        }
        return NK_OK;
    };
    auto record = [&](int k) -> nk_status {      // tape step k = (U, V) and, DIVA, the fields' state
        k_copy2<double><<<B, THREADS, 0, s->stream>>>(g, f[NK_FIELD_U], f[NK_FIELD_V], tu(k), tv(k));
        CHK(launch_check(s));
        if (diva) {
            const DivaLag L = lag_of(k);
            CHK(copy_diva_state(s, f[NK_FIELD_BETABU], f[NK_FIELD_BETABV], f[NK_FIELD_F2], f[NK_FIELD_ETASIGMA],
                                (double *)L.bbu, (double *)L.bbv, (double *)L.F2, (double *)L.es));
        }
        return NK_OK;
    };
    // forward replay from the recorded initial guess (and lagged state)
    k_copy2<double><<<B, THREADS, 0, s->stream>>>(g, s->x0u, s->x0v, tu(0), tv(0));
    CHK(launch_check(s));
    if (diva) {
        const DivaLag L = lag_of(0);
        CHK(copy_diva_state(s, s->s0_bbu, s->s0_bbv, s->s0_F2, s->s0_es, (double *)L.bbu, (double *)L.bbv, (double *)L.F2,
                            (double *)L.es));
    }
    CHK(set_state(0));
    nk_info fwd{};
    for (int k = 1; k <= n; k++) {
        CHK(update_coefficients(s));
        CHK(solve_linear(s, &fwd));
        CHK(record(k));
    }
    // reverse sweep
    double *lu = f[NK_FIELD_LAMBDAU], *lv = f[NK_FIELD_LAMBDAV];
    double *zero2[][2] = {{lu, lv}, {f[NK_FIELD_CBAR], nullptr}, {f[NK_FIELD_HBAR], f[NK_FIELD_BEDBAR]}};
    for (auto &z : zero2) {
        k_fill2<double><<<Be, THREADS, 0, s->stream>>>(g, z[0], z[1], 0.0, nullptr);
        CHK(launch_check(s));
    }
    for (int k = 0; k < (s->ph.rate_factor_depth == NK_RATE_FACTOR_3D ? s->Ns : 1); k++) {
        k_fill2<double><<<Be, THREADS, 0, s->stream>>>(g, f[NK_FIELD_ABAR] + (long)k * s->plane, nullptr, 0.0, nullptr);
        CHK(launch_check(s));
    }
    if (diva) {
        // the post-processing's pullback at the final state: x̄_n = x̄ + ū, and s̄_n from the layer velocities
        CHK(diva_profile_pullback(s));
        CHK(copy_diva_state(s, s->cs_bbu, s->cs_bbv, s->cs_F2, s->cs_es, s->st_bbu, s->st_bbv, s->st_F2, s->st_es));
    } else {
        k_copy2<double><<<B, THREADS, 0, s->stream>>>(g, f[NK_FIELD_UBAR], f[NK_FIELD_VBAR], s->ad_xbu, s->ad_xbv);   // x̄_n
        CHK(launch_check(s));
    }
    long total = 0;
    double res = 0.0;
    for (int k = n; k >= 1; k--) {
        CHK(set_state(k - 1));
        CHK(update_coefficients(s));                                              // c_k (and s_k) from x_{k−1} (and s_{k−1})
        CHK(update_preconditioner(s));
        k_fill2<double><<<Be, THREADS, 0, s->stream>>>(g, tlu, tlv, 0.0, nullptr);
        CHK(launch_check(s));
        int cg = 0;
        CHK(pcg<double>(s, tlu, tlv, s->ad_xbu, s->ad_xbv, s->op.linear, false, &cg, &res));   // λ_k = A(c_k)⁻¹ x̄_k
        total += cg;
        k_add<<<B, THREADS, 0, s->stream>>>(g, lu, lv, tlu, tlv, lu, lv);      // τ̄ += λ_k
        CHK(launch_check(s));
        k_negate<<<B, THREADS, 0, s->stream>>>(g, tlu, tlv, s->ad_qu, s->ad_qv);   // q̄ = −λ_k
        CHK(launch_check(s));
        if (diva) {
            CHK(adjoint_step_diva(s, true, tu(k), tv(k), lag_of(k - 1)));       // x̄_{k−1}, s̄_{k−1} (out_*), C̄ and Ā += ...
            CHK(copy_diva_state(s, s->out_bbu, s->out_bbv, s->out_F2, s->out_es, s->st_bbu, s->st_bbv, s->st_F2, s->st_es));
        } else {
            CHK(adjoint_step(s, true, true, tu(k), tv(k)));                      // x̄_{k−1}, C̄ and Ā += ...
        }
        k_copy2<double><<<B, THREADS, 0, s->stream>>>(g, s->ad_yu, s->ad_yv, s->ad_xbu, s->ad_xbv);
        CHK(launch_check(s));
        CHK(geometry_pullback(s, tlu, tlv, diva ? nullptr : tu(k), diva ? nullptr : tv(k), true));   // H̄, b̄ += ...
        if (s->verbose) {
            std::fprintf(stderr, "[nunatak_hip] unrolled adjoint step %d: linear %d\n", k, cg);
            std::fflush(stderr);
        }
    }
    // the handle as the forward solve left it: the coefficients (and state) of step n, (U, V) = x_n, post-processing
    if (n >= 1) {
        CHK(set_state(n - 1));
        CHK(update_coefficients(s));
    }
    k_copy2<double><<<B, THREADS, 0, s->stream>>>(g, tu(n), tv(n), f[NK_FIELD_U], f[NK_FIELD_V]);
    CHK(launch_check(s));
    CHK(post_process(s));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (info) {
        info->picard_iterations = n;
        info->total_linear_iterations = total;
        info->last_residual_norm = res;
    }
    return NK_OK;
}

}  // namespace

nk_status nk_adjoint(nk_solver *s, const nk_control *control, nk_info *info) {
    if (!s) return NK_INVALID_ARGUMENT;
    Timer t(s, info);
    if (s->op.picard.kind == NK_FIXED_ITERATIONS) return t.done(adjoint_unrolled(s, s->op.picard.iterations, info));
    const nk_control ctl = control ? *control : s->op.picard;
    return t.done(adjoint(s, ctl, info));
}

nk_status nk_multigrid_levels(const nk_solver *s, int32_t *levels, int32_t *shapes, int32_t max_levels) {
    if (!s || !levels) return NK_INVALID_ARGUMENT;
    int n = (int)s->lev.size() - 1;
    *levels = n;
    for (int l = 0; l < n && l < max_levels && shapes; l++) {
        shapes[2 * l] = s->lev[l + 1].g.Nx;
        shapes[2 * l + 1] = s->lev[l + 1].g.Ny;
    }
    return NK_OK;
}

// ---- allocation-free operator apply ------------------------------------------------------------------------------
// q = A p on caller buffers with the activity flags and neighbour masks in a fixed device array instead of a handle's
// planes. The array is per device (each GPU the library runs on holds its own copy until the process ends); its device
// address is taken once per device with hipGetSymbolAddress and only that address is used on the host side.

}  // extern "C"

namespace {

static __device__ __align__(256) uint8_t scratch[NK_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

nk_status free_fail(nk_status st, const std::string &msg) {
    g_free_err = msg;
    return st;
}

nk_status scratch_for_current_device(void **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return free_fail(NK_DEVICE_ERROR, "nk_apply_operator_noalloc: no current HIP device");
    if (dev >= NOALLOC_MAX_DEVICES) return free_fail(NK_UNSUPPORTED, "nk_apply_operator_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return free_fail(NK_DEVICE_ERROR, std::string("nk_apply_operator_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return NK_OK;
}

}  // namespace

extern "C" {

nk_status nk_apply_operator_noalloc(const nk_grid *grid, int32_t walls, const double *eta, const double *etaff,
                                    const double *betau, const double *betav, const double *pu, const double *pv, double *qu,
                                    double *qv, int64_t pitch, void *stream) {
    if (!grid || !eta || !etaff || !betau || !betav || !pu || !pv || !qu || !qv)
        return free_fail(NK_INVALID_ARGUMENT, "nk_apply_operator_noalloc: null argument");
    if (grid->Nx < 1 || grid->Ny < 1 || !(grid->dx > 0) || !(grid->dy > 0) ||
        (grid->topology_x != NK_PERIODIC && grid->topology_x != NK_BOUNDED) ||
        (grid->topology_y != NK_PERIODIC && grid->topology_y != NK_BOUNDED))
        return free_fail(NK_INVALID_ARGUMENT, "nk_apply_operator_noalloc: grid sizes, spacings or topologies invalid");
    if (walls != NK_FREE_SLIP && walls != NK_NO_SLIP)
        return free_fail(NK_INVALID_ARGUMENT, "nk_apply_operator_noalloc: walls must be NK_FREE_SLIP or NK_NO_SLIP");
    const KG g = make_kg(grid->Nx, grid->Ny, grid->topology_x, grid->topology_y, grid->dx, grid->dy);
    if (pitch != g.pitch)
        return free_fail(NK_INVALID_ARGUMENT, "nk_apply_operator_noalloc: pitch must be Nx + 1 rounded up to 64 elements (" +
                                                  std::to_string(g.pitch) + " here)");
    const size_t plane = (size_t)plane_of(g);
    const size_t nm_off = (plane + 255) / 256 * 256;
    const size_t need = nm_off + 2 * plane;
    if (need > (size_t)NK_NOALLOC_SCRATCH_BYTES)
        return free_fail(NK_INVALID_ARGUMENT, "nk_apply_operator_noalloc: grid too large for the fixed scratch array (needs " +
                                                  std::to_string(need) + " bytes, NK_NOALLOC_SCRATCH_BYTES = " +
                                                  std::to_string((size_t)NK_NOALLOC_SCRATCH_BYTES) + ")");
    void *base = nullptr;
    nk_status st = scratch_for_current_device(&base);
    if (st != NK_OK) return st;
    uint8_t *act = (uint8_t *)base;
    uint16_t *nm = (uint16_t *)((uint8_t *)base + nm_off);
    hipStream_t hs = (hipStream_t)stream;
    // every flag and mask word the operator reads is rewritten here, so the previous call's contents never leak in
    k_corners_flags<<<blocks_for(g.Nx + 1, g.Ny + 1), THREADS, 0, hs>>>(g, eta, nullptr, act);
    NK_LAUNCH_TOPO2(g, k_nmask, blocks_for(g.Nx, g.Ny), THREADS, 0, hs, g, act, nm);
    Coef<double> c;
    c.eta = eta; c.etaff = etaff; c.bu = betau; c.bv = betav; c.act = act;
    Red r;
    r.partials = nullptr; r.S = nullptr; r.fin = FIN_NONE;
    NK_LAUNCH_TOPO(g, k_apply_mt, double, blocks_for(g.Nx, g.Ny), THREADS, 0, hs, g, pu, pv, c, (const uint16_t *)nm,
                   walls == NK_NO_SLIP ? 2.0 : 0.0, qu, qv, (const double *)nullptr, (const double *)nullptr, r, 0, (const Scal *)nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return free_fail(NK_DEVICE_ERROR, std::string("nk_apply_operator_noalloc: launch: ") + hipGetErrorString(e));
    return NK_OK;
}

}  // extern "C"
