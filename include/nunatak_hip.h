/*
 * nunatak_hip.h — C ABI of libnunatak_hip.so, the MI355X (gfx950) velocity solver for Nunatak.jl.
 *
 * The library replaces the reference's depth-integrated velocity solve:
 *   update_velocities!  (src/Models/ice_sheet_model_methods.jl:116-147)
 *   -> compute_driving_stress!  (src/StressBalances/driving_stress.jl:43-53)
 *   -> picard_solve!            (src/Solvers/picard.jl:71-90)
 *      -> update_stress_balance_coefficients! / update_diva_coefficients!
 *                               (src/StressBalances/stress_balance_coefficients.jl:10-20, src/StressBalances/diva.jl:61-75)
 *      -> solve! (PCG)          (src/Solvers/conjugate_gradient.jl:128-223)
 *         -> ssa_operator!      (src/StressBalances/ssa_operator.jl:47-89)
 *         -> precondition!      (src/StressBalances/diagonal_preconditioner.jl:11-18,
 *                                src/StressBalances/multigrid_preconditioner.jl:122-128)
 *   and the geometry diagnostics feeding it, update_geometry! (src/Models/ice_sheet_model_methods.jl:103-111).
 *
 * Model of the boundary. A handle owns one grid's device state: every field the reference keeps in
 * `model.geometry`, `model.auxiliary` and `model.velocities` (src/Models/ice_sheet_fields.jl) lives in HBM under a
 * field id (nk_field_id). The host copies fields in and out with nk_set_field / nk_get_field and then calls
 * operations that mirror the reference functions one to one. Device pointers of the FP64 fields are exposed for
 * zero-copy callers. All entry points return an nk_status; nothing aborts and no C++ exception crosses the ABI.
 *
 * Host array convention (the reference's column-major layout): element (i, j[, k]) of a field, 1-based as in
 * Julia, is at data[(i-1) + (j-1)*ld + (k-1)*lds]. ld = 0 means ld = the field's x extent; lds = 0 means
 * lds = ld * (y extent). For an Oceananigans parent array with halo (Hx, Hy) pass
 * pointer(parent(f), Hx + 1 + Hy * size(parent(f), 1)) with ld = size(parent(f), 1).
 * Field extents follow Oceananigans `size(field)` on a RectilinearGrid with Flat z: a Face location on a
 * Bounded direction has N + 1 points, every other location has N.
 *
 * Threading: a handle is used by one host thread at a time; distinct handles are independent and each owns its
 * own HIP stream.
 */
#ifndef NUNATAK_HIP_H
#define NUNATAK_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NK_ABI_VERSION 1

typedef enum nk_status {
    NK_OK = 0,
    NK_INVALID_ARGUMENT = 1,  /* the reference throws ArgumentError (e.g. src/StressBalances/ssa_operator.jl:56) */
    NK_NOT_CONVERGED = 2,     /* reported only when the caller asks for it; the reference is silent */
    NK_NONFINITE = 3,         /* NaN or Inf found in a solution */
    NK_OUT_OF_MEMORY = 4,
    NK_DEVICE_ERROR = 5,
    NK_UNSUPPORTED = 6
} nk_status;

/* Topology of a horizontal direction: Oceananigans Periodic or Bounded. */
enum { NK_PERIODIC = 0, NK_BOUNDED = 1 };

/* Grid: a uniform RectilinearGrid (Flat z) plus the σ column of SigmaGrid (src/Grids/sigma_grid.jl:20-62). */
typedef struct nk_grid {
    int32_t Nx, Ny;                 /* cells */
    int32_t topology_x, topology_y; /* NK_PERIODIC or NK_BOUNDED */
    int32_t halo;                   /* halo of the host grid; enters the multigrid coarsening rule */
    int32_t Nsigma;                 /* σ layers, default 10 (src/Models/ice_sheet_model.jl:119) */
    double dx, dy;                  /* uniform cell widths (m) */
    double bed_refinement;          /* σ faces ζ_k = (k/Nσ)^bed_refinement, default 2 */
} nk_grid;

enum { NK_GROUNDING_NONE = 0, NK_GROUNDING_SUBGRID = 1, NK_GROUNDING_SUBGRID_DRIVING_STRESS = 2 };
enum { NK_FRICTION_NONE = 0, NK_FRICTION_WEERTMAN = 1, NK_FRICTION_REGULARIZED_COULOMB = 2 };
enum { NK_FREE_SLIP = 0, NK_NO_SLIP = 1 };
enum { NK_SSA = 0, NK_DIVA = 1 };
enum { NK_RATE_FACTOR_2D = 0, NK_RATE_FACTOR_3D = 1 };

/* Physics of the model: IceSheetConstants, SmoothIceMask, GlenFlowLaw, the friction law, the grounding treatment
 * and DepthIntegrated (src/Models/ice_sheet_constants.jl:33, src/Models/smooth_ice_mask.jl:41,
 * src/Rheologies/glen_flow_law.jl:36, src/Frictions/frictions.jl:34,65, src/Groundings/subgrid_grounded_fraction.jl:23,
 * src/StressBalances/depth_integrated.jl:29). */
typedef struct nk_physics {
    double ice_density, ocean_density, gravitational_acceleration;   /* 918, 1028, 9.81 */
    int32_t grounding;                                               /* NK_GROUNDING_* */
    int32_t masking;                                                 /* 1 = SmoothIceMask, 0 = none (μ ≡ 1 not used) */
    double mask_minimum_thickness, mask_transition_thickness;        /* 1, 1 */
    double glen_exponent, strain_rate_regularization;                /* 3, 1e-13 */
    int32_t rate_factor_depth;                                       /* NK_RATE_FACTOR_2D or _3D */
    int32_t friction;                                                /* NK_FRICTION_* */
    double friction_exponent, coulomb_coefficient, velocity_regularization;  /* 1/3, α², 1e-8 */
    int32_t walls;                                                   /* NK_FREE_SLIP or NK_NO_SLIP */
    int32_t stress_balance;                                          /* NK_SSA or NK_DIVA */
} nk_physics;

/* Iteration control: Convergence(reltol, abstol, maxiter) or FixedIterations(n) (src/Solvers/iteration_controls.jl). */
enum { NK_CONVERGENCE = 0, NK_FIXED_ITERATIONS = 1 };
typedef struct nk_control {
    int32_t kind;
    int32_t maxiter;     /* Convergence */
    int32_t iterations;  /* FixedIterations */
    int32_t pad;
    double reltol, abstol;
} nk_control;

enum { NK_PRECONDITIONER_NONE = 0, NK_PRECONDITIONER_DIAGONAL = 1, NK_PRECONDITIONER_MULTIGRID = 2 };
enum { NK_POINT_JACOBI = 0, NK_LINE_JACOBI = 1 };
/* NK_PRECISION_MIXED: FP64 PCG (operator, residual, α/β, stopping rule) preconditioned by an FP32 multigrid V-cycle
 * whose coarsest-grid solve is FP64. NK_PRECISION_MIXED_REFINED: FP64 iterative refinement — the defect r = b − A x
 * and its norm in FP64, the stopping rule on it — around an FP32 PCG (FP32 operator, vectors and V-cycle, FP64 dots
 * and α/β) that solves A e = r to a relative 1e-4 from zero each step. Both need the multigrid preconditioner. */
enum { NK_PRECISION_FP64 = 0, NK_PRECISION_MIXED = 1, NK_PRECISION_MIXED_REFINED = 2 };
/* Coarse-level operators of the multigrid preconditioner: rediscretized from coarsened coefficients, as the reference
 * builds them (src/StressBalances/multigrid_preconditioner.jl:128-141), or Galerkin products Pᵀ A P of the level
 * above with the same transfers (no reference counterpart; keeps a calving front free on coarse levels where a
 * coarse cell straddles the front and an ice-free strip narrower than the cell). */
enum { NK_COARSE_REDISCRETIZED = 0, NK_COARSE_GALERKIN = 1 };

/* StressBalanceSolver options (src/StressBalances/stress_balance_solver.jl:43-52) and the multigrid defaults
 * (src/StressBalances/multigrid_preconditioner.jl:57). */
typedef struct nk_solver_options {
    nk_control linear;                /* Convergence(reltol = 1e-8, maxiter = 10_000) */
    nk_control picard;                /* Convergence(reltol = 1e-6, maxiter = 100) */
    int32_t preconditioner;           /* NK_PRECONDITIONER_DIAGONAL */
    int32_t smoother;                 /* NK_POINT_JACOBI */
    int32_t smoothing_sweeps;         /* 2 */
    int32_t coarsest_sweeps;          /* 20 */
    int32_t minimum_coarse_size;      /* 2 */
    int32_t precision;                /* NK_PRECISION_FP64; NK_PRECISION_MIXED = FP64 PCG (operator, residual, α/β,
                                         stopping rule) preconditioned by an FP32 multigrid V-cycle (multigrid only) */
    int32_t coarse_operator;          /* NK_COARSE_REDISCRETIZED; NK_COARSE_GALERKIN (multigrid, point Jacobi) */
    int32_t check_interval;           /* PCG iterations launched between host checks of the on-device stop flag */
    int32_t profile_kernel;           /* NK_PROFILE_*: bracket every launch of that kernel with HIP events
                                         on the handle's stream and report the summed device time in nk_info */
    int32_t graphs;                   /* 1: capture the launches of check_interval PCG iterations in a hipGraph once
                                         and replay it (not while profiling); 0: launch kernel by kernel */
    int32_t report_not_converged;     /* 0 (the reference: silent); 1: nk_solve returns NK_NOT_CONVERGED when PCG stops at
                                         maxiter above its tolerance, the Picard entry points when Picard does */
} nk_solver_options;

/* Kernels that nk_solver_options.profile_kernel can time (finest grid unless stated). */
enum {
    NK_PROFILE_NONE = 0,
    NK_PROFILE_OPERATOR = 1,   /* q = A p of the PCG iteration */
    NK_PROFILE_SWEEP = 2,      /* fused ℓ1-Jacobi sweep x + (b − A x)/s (not the first sweep from zero, nor the mixed
                                  form's last one, which also writes the FP64 z and r·z) */
    NK_PROFILE_UPDATE = 3,     /* PCG vector update x += αp, r −= αq (+ Jacobi z = r/d) */
    NK_PROFILE_GALERKIN_SWEEP = 4   /* ℓ1-Jacobi sweep with the Galerkin stencil on the first coarse level (not the sweeps
                                       from zero); its launches count that level's points */
};

/* Field ids. Location (x, y[, σ]) in brackets; Face on a Bounded direction has N + 1 points. */
typedef enum nk_field_id {
    NK_FIELD_U = 0,        /* depth-averaged u            (Face, Center)      */
    NK_FIELD_V,            /* depth-averaged v            (Center, Face)      */
    NK_FIELD_H,            /* thickness                   (Center, Center)    */
    NK_FIELD_BED,          /* bed elevation b             (Center, Center)    */
    NK_FIELD_S,            /* surface elevation           (Center, Center)    */
    NK_FIELD_PHI,          /* grounded fraction φ         (Center, Center)    */
    NK_FIELD_PHIU,         /* φᵘ                          (Face, Center)      */
    NK_FIELD_PHIV,         /* φᵛ                          (Center, Face)      */
    NK_FIELD_N,            /* effective pressure          (Center, Center)    */
    NK_FIELD_MU,           /* ice mask μ                  (Center, Center)    */
    NK_FIELD_TAUU,         /* driving stress τᵘ           (Face, Center)      */
    NK_FIELD_TAUV,         /* driving stress τᵛ           (Center, Face)      */
    NK_FIELD_ETAH,         /* η̄H                          (Center, Center)    */
    NK_FIELD_ETAHFF,       /* η̄Hᶠᶠ                        (Face, Face)        */
    NK_FIELD_BETAU,        /* operator drag βᵘ            (Face, Center)      */
    NK_FIELD_BETAV,        /* operator drag βᵛ            (Center, Face)      */
    NK_FIELD_F2,           /* DIVA integral F₂            (Center, Center)    */
    NK_FIELD_ETASIGMA,     /* ησ                          (Center, Center, σ) */
    NK_FIELD_BETABU,       /* friction-law drag βᵇᵘ       (Face, Center)      */
    NK_FIELD_BETABV,       /* βᵇᵛ                         (Center, Face)      */
    NK_FIELD_UB,           /* basal velocity uᵇ           (Face, Center)      */
    NK_FIELD_VB,           /* vᵇ                          (Center, Face)      */
    NK_FIELD_USIGMA,       /* uσ                          (Face, Center, σ)   */
    NK_FIELD_VSIGMA,       /* vσ                          (Center, Face, σ)   */
    NK_FIELD_A,            /* rate factor A               (Center, Center[, σ]) per rate_factor_depth */
    NK_FIELD_C,            /* friction coefficient C      (Center, Center)    */
    NK_FIELD_TEMPERATURE,  /* temperature (K) input of nk_rate_factor_from_temperature (Center, Center, σ) */
    NK_FIELD_PU,           /* operand vector p / r        (Face, Center)      */
    NK_FIELD_PV,           /*                             (Center, Face)      */
    NK_FIELD_QU,           /* result vector q / z         (Face, Center)      */
    NK_FIELD_QV,           /*                             (Center, Face)      */
    NK_FIELD_UBAR,         /* adjoint input x̄ = ∂J/∂u     (Face, Center)      */
    NK_FIELD_VBAR,         /* ∂J/∂v                       (Center, Face)      */
    NK_FIELD_LAMBDAU,      /* adjoint state λ = ∂J/∂τᵘ    (Face, Center)      */
    NK_FIELD_LAMBDAV,      /* ∂J/∂τᵛ                      (Center, Face)      */
    NK_FIELD_CBAR,         /* ∂J/∂C                       (Center, Center)    */
    NK_FIELD_ABAR,         /* ∂J/∂A                       (Center, Center[, σ]) like NK_FIELD_A */
    NK_FIELD_USIGMABAR,    /* adjoint input ∂J/∂uσ (DIVA) (Face, Center, σ)   */
    NK_FIELD_VSIGMABAR,    /* ∂J/∂vσ                      (Center, Face, σ)   */
    NK_FIELD_ENTHALPY,     /* enthalpy E (J/kg), input of nk_thermal_diagnostics (Center, Center, σ) */
    NK_FIELD_WATER_FRACTION, /* liquid water fraction ω    (Center, Center, σ)   */
    NK_FIELD_SMB,          /* surface mass balance a (m/s) (Center, Center)    */
    NK_FIELD_BASAL_MELT,   /* basal melt rate m (m/s)     (Center, Center)    */
    NK_FIELD_GROUNDED_MELT,/* grounded melt mᵍ (m/s), written by nk_energy_step (Center, Center) */
    NK_FIELD_HPREV,        /* H⁻: thickness at the start of the time step (Center, Center) */
    NK_FIELD_GH,           /* thickness tendency G        (Center, Center)    */
    NK_FIELD_HCLIP,        /* clipped thickness Hᶜ (thickness added by the clip at zero) (Center, Center) */
    NK_FIELD_HW,           /* basal water layer thickness Hw (m) (Center, Center) */
    NK_FIELD_HN,           /* thickness at the last energy step Hⁿ (Center, Center) */
    NK_FIELD_GEOTHERMAL,   /* geothermal heat flux G (W m⁻²) (Center, Center) */
    NK_FIELD_TSURF,        /* surface temperature Tₛ (K)  (Center, Center)    */
    NK_FIELD_TOCEAN,       /* ocean temperature under floating ice Tₒ (K) (Center, Center) */
    NK_FIELD_STRAIN_HEATING, /* strain heating Φ (W m⁻³) (Center, Center, σ)  */
    NK_FIELD_HBAR,         /* ∂J/∂H from nk_adjoint (Center, Center): through geometry, driving stress, coefficients */
    NK_FIELD_BEDBAR,       /* ∂J/∂b from nk_adjoint (Center, Center)          */
    NK_FIELD_COUNT
} nk_field_id;

/* Face classes written by nk_compute_driving_stress, one byte per velocity point, bit-exact with the reference's
 * branch selection in _compute_x/_y_driving_stress! (src/StressBalances/driving_stress.jl:109-125). */
enum {
    NK_CLASS_WALL = 1,       /* peripheral_node: Bounded wall face */
    NK_CLASS_ICE_FREE = 2,   /* μ = 0 on both sides */
    NK_CLASS_FRONT = 4,      /* μ > 0 on exactly one side: calving-front force */
    NK_CLASS_GROUNDING = 8   /* flotation changes sign across the face: grounding-line split */
};

typedef struct nk_info {
    int32_t picard_iterations;      /* solver.iteration of StressBalanceSolver */
    int32_t linear_iterations;      /* iterations of the last linear solve (VelocityConjugateGradientSolver.iteration) */
    int64_t total_linear_iterations;
    int32_t refinement_steps;       /* NK_PRECISION_MIXED_REFINED: FP64 defect corrections in the last linear solve */
    int32_t pad;
    double last_picard_change;      /* Δ = ‖x − x⁻‖ */
    double last_picard_norm;        /* X = ‖x‖ */
    double last_residual_norm;      /* ‖r‖ at the end of the last linear solve */
    double elapsed_ms;              /* device time of the call, from HIP events */
    double profiled_ms;             /* summed device time of the profiled kernel's launches (profile_kernel) */
    int64_t profiled_launches;      /* launches of the profiled kernel */
    int64_t profiled_points;        /* Σ Nx·Ny of those launches (finest grid) */
    int32_t profiled_precision;     /* bytes per element of the profiled launches (8 FP64, 4 FP32) */
    int32_t pad2;
} nk_info;

typedef struct nk_solver nk_solver;

/* Version of this ABI (NK_ABI_VERSION). */
int32_t nk_abi_version(void);

/* Fill defaults of the reference's constructors. */
void nk_default_grid(nk_grid *grid);
void nk_default_physics(nk_physics *physics);
void nk_default_options(nk_solver_options *options);

/* StressBalanceSolver(grid; ...) + the field set of IceSheetModel (src/StressBalances/stress_balance_solver.jl:43-52,
 * src/Models/ice_sheet_fields.jl:8-59), and StressBalanceMultigridPreconditioner(grid; ...) when options ask for
 * it (src/StressBalances/multigrid_preconditioner.jl:57). Allocates all device state of the handle. */
nk_status nk_create(const nk_grid *grid, const nk_physics *physics, const nk_solver_options *options,
                    nk_solver **out);
nk_status nk_destroy(nk_solver *solver);
const char *nk_last_error(const nk_solver *solver);   /* NULL: the last handle-free call's error on this thread */

/* Change physics or options of an existing handle (same grid). */
nk_status nk_set_physics(nk_solver *solver, const nk_physics *physics);
nk_status nk_set_options(nk_solver *solver, const nk_solver_options *options);

/* Field extents (nx, ny, nz) of a field id on this handle's grid. */
nk_status nk_field_shape(const nk_solver *solver, nk_field_id id, int32_t shape[3]);
/* Copy a host array into / out of a device field (FP64), synchronously. */
nk_status nk_set_field(nk_solver *solver, nk_field_id id, const double *data, int64_t ld, int64_t lds);
nk_status nk_get_field(nk_solver *solver, nk_field_id id, double *data, int64_t ld, int64_t lds);
nk_status nk_fill_field(nk_solver *solver, nk_field_id id, double value);
/* Face classes (NK_CLASS_* bits) of the last nk_compute_driving_stress. */
nk_status nk_get_classes(nk_solver *solver, uint8_t *classu, int64_t ldu, uint8_t *classv, int64_t ldv);
/* Device address, pitch (elements between rows) and layer stride of an FP64 field. */
nk_status nk_field_device_pointer(nk_solver *solver, nk_field_id id, double **ptr, int64_t *pitch, int64_t *layer);
/* hipStream_t of the handle, as void*. */
void *nk_stream(nk_solver *solver);
nk_status nk_synchronize(nk_solver *solver);

/* update_geometry!: s, φ, φᵘ, φᵛ, N, μ from H and bed (src/Models/ice_sheet_model_methods.jl:103-111). */
nk_status nk_update_geometry(nk_solver *solver);
/* compute_driving_stress!(τ, H, s, ρᵢ, ρw, g, μ, b) (src/StressBalances/driving_stress.jl:43-53); the bed is used
 * for the grounding-line split only with NK_GROUNDING_SUBGRID_DRIVING_STRESS, μ only when masking is on. */
nk_status nk_compute_driving_stress(nk_solver *solver);
/* update_stress_balance_coefficients! (SSA) or update_diva_coefficients! (DIVA) at the current (u, v). */
nk_status nk_update_coefficients(nk_solver *solver);
/* ssa_operator!(q, p, η̄H, η̄Hᶠᶠ, βᵘ, βᵛ, walls): (QU, QV) = A (PU, PV). */
nk_status nk_apply_operator(nk_solver *solver);

/* Allocation-free variant of nk_apply_operator. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * q = A p (ssa_operator!, src/StressBalances/ssa_operator.jl:81-89) on caller-owned device arrays in the library's
 * plane layout: element (i, j) of every field at (j − 1)·pitch + (i − 1), pitch = Nx + 1 rounded up to 64 elements,
 * (Ny + 1) rows (nk_field_device_pointer of a handle with the same grid returns such planes). eta: η̄H at centres,
 * etaff: η̄Hᶠᶠ at corners, betau/betav: β at u/v points, (pu, pv) in, (qu, qv) out over the Nx × Ny :xyz extent;
 * walls is NK_FREE_SLIP or NK_NO_SLIP. The activity flags and neighbour masks live in one fixed device array of
 * NK_NOALLOC_SCRATCH_BYTES per GPU (3 bytes per plane point: grids up to about 2048² cells); larger grids return
 * NK_INVALID_ARGUMENT. The call makes no allocation of any kind and does not synchronize: it enqueues three kernels
 * on `stream` (a hipStream_t; NULL = the null stream) of the current device. Calls that overlap on one GPU share the
 * array and race; serialize them (one stream, or events). Errors are read with nk_last_error(NULL) on the calling
 * thread. Results are nk_apply_operator's bit for bit. */
#define NK_NOALLOC_SCRATCH_BYTES (16u << 20)
nk_status nk_apply_operator_noalloc(const nk_grid *grid, int32_t walls, const double *eta, const double *etaff,
                                    const double *betau, const double *betav, const double *pu, const double *pv,
                                    double *qu, double *qv, int64_t pitch, void *stream);
/* precondition!(z, P, r, coefficients...): (QU, QV) = M⁻¹ (PU, PV) with the handle's preconditioner; the multigrid
 * coarse coefficients are rebuilt first (update_preconditioner!). */
nk_status nk_precondition(nk_solver *solver);
/* solve!(x, solver, b, coefficients...): PCG on (U, V) with right-hand side (TAUU, TAUV), frozen coefficients. */
nk_status nk_solve(nk_solver *solver, nk_info *info);
/* picard_solve!(x, picard, solver, update!, τ, coefficients...) with the current τ. */
nk_status nk_picard_solve(nk_solver *solver, nk_info *info);
/* update_velocities!(model, solver, ::SSA or ::DIVA): driving stress, Picard, then extend_depth_average! (SSA) or
 * compute_velocity_profile! (DIVA). With geometry != 0 update_geometry! runs first. */
nk_status nk_update_velocities(nk_solver *solver, int32_t geometry, nk_info *info);
/* The post-processing of update_velocities! alone: compute_velocity_profile! (src/StressBalances/diva.jl:89-109:
 * UB, VB and the layer velocities USIGMA, VSIGMA from U, V and the DIVA state) or extend_depth_average!
 * (src/Models/ice_sheet_model_methods.jl:167-180) for SSA. */
nk_status nk_compute_velocity_profile(nk_solver *solver);
/* set!(model; T = ...): ENTHALPY = enthalpy(T, 0, p) and the Arrhenius rate factor A on the σ grid from TEMPERATURE via the enthalpy round trip
 * (src/Models/ice_sheet_model_methods.jl:47-57, src/Thermodynamics/column_energy_solve.jl:105-117,
 * src/Rheologies/arrhenius_rate_factor.jl:102-110) with the reference's default parameters. */
nk_status nk_rate_factor_from_temperature(nk_solver *solver);
/* update_thermal_diagnostics! (src/Thermodynamics/column_energy_solve.jl:85-117) with the reference's default
 * parameters: TEMPERATURE = min(Tₚₘ, Tᵣ + E/cᵢ) and WATER_FRACTION = max(0, E − Eₚₘ)/L from ENTHALPY at each layer's
 * pressure (enthalpy_conversions.jl), then A = arrhenius_rate_factor(T − Tₚₘ + T₀, ω) on the σ grid
 * (NK_RATE_FACTOR_3D). */
nk_status nk_thermal_diagnostics(nk_solver *solver);

/* ---- Thickness transport: the step after the velocity solve in time_step! ---------------------------------------- */
/* Advection schemes of the thickness flux: Oceananigans UpwindBiased(order = 1), Centered(order = 2), WENO(order = 5)
 * and WENO(order = 5, bounds) with the bounds-preserving limiter (the reference's default for SSA and DIVA). */
enum { NK_ADVECTION_UPWIND1 = 0, NK_ADVECTION_CENTERED2 = 1, NK_ADVECTION_WENO5 = 2, NK_ADVECTION_WENO5_BOUNDED = 3 };
/* compute_thickness_tendency! (src/Models/time_stepping.jl:34-63): GH = a − m·(1 − φ if grounding else 1) − mᵍ
 * − ∇·(ū H) with the face thickness of the scheme; U, V are the advecting depth-averaged velocities; mᵍ
 * (GROUNDED_MELT) enters only with grounded_melt_coupling (coupled_grounded_melt, :54-55). */
nk_status nk_compute_thickness_tendency(nk_solver *solver, int32_t advection, int32_t grounded_melt_coupling);
/* Bounds of NK_ADVECTION_WENO5_BOUNDED (WENO(order = 5, bounds = (lower, upper)); default (0, 1e5), the reference's
 * default thickness advection, src/Models/ice_sheet_model.jl:123). */
nk_status nk_set_thickness_bounds(nk_solver *solver, double lower, double upper);
/* _thickness_substep! (time_stepping.jl:65-72): H★ = HPREV + Δτ·GH; H = max(0, H★); in the final stage
 * HCLIP += max(0, −H★). */
nk_status nk_thickness_substep(nk_solver *solver, double dtau, int32_t final_stage);
/* cache_current_fields! (time_stepping.jl:1-5): HPREV = H. */
nk_status nk_cache_thickness(nk_solver *solver);
/* cell_advection_timescale(model) (time_stepping.jl:78, Oceananigans): min over cells of 1/(|u|/Δx + |v|/Δy). */
nk_status nk_cell_advection_timescale(nk_solver *solver, double *timescale);

/* ---- Enthalpy step: the energy equation once per ice time step -------------------------------------------------- */
/* compute_strain_heating! (src/Thermodynamics/strain_heating.jl): STRAIN_HEATING = 4 ησ ε̇ₑ² (SSA); DIVA adds the
 * vertical shear τᵇ² w₂ₖ / Δζₖ / ησ. Reads U, V, ETASIGMA (and UB, VB, BETABU, BETABV for DIVA). */
nk_status nk_compute_strain_heating(nk_solver *solver);
/* advance_enthalpy! (src/Thermodynamics/column_energy_solve.jl:48-84) over dt with the reference's default
 * EnthalpyParameters: layer volume fluxes from USIGMA, VSIGMA, H, HN, BASAL_MELT (and GROUNDED_MELT when
 * grounded_melt_coupling), the basal state from ENTHALPY, HW, GEOTHERMAL and the basal drag (BETAU/BETAV with UB/VB
 * under SSA, BETABU/BETABV under DIVA), the implicit column system with STRAIN_HEATING, TSURF and TOCEAN, the
 * tridiagonal solve, drainage and the basal water (HW, GROUNDED_MELT = mᵍ), then update_thermal_diagnostics!
 * (TEMPERATURE, WATER_FRACTION; A with NK_RATE_FACTOR_3D) and HN = H. Needs Nσ ≥ 1 and the σ grid of the handle;
 * the first call allocates the column system's workspace (≈ (6 Nσ + 5) planes of FP64). With a till law, the basal
 * water then drains and is capped (drain_till); till == NULL is a bed connected to the ocean (no till). */
typedef struct nk_till {
    /* TillEffectivePressure (src/Thermodynamics/till_effective_pressure.jl:16-61); nk_default_till gives its defaults */
    double reference_effective_pressure;   /* N₀ = 1000 Pa */
    double overburden_fraction;            /* δ = 0.02 */
    double reference_void_ratio;           /* e₀ = 0.69 */
    double compressibility;                /* Cc = 0.12 */
    double maximum_water_thickness;        /* Wᵐᵃˣ = 2 m */
    double drainage_rate;                  /* 1 mm per year, in m/s */
} nk_till;
void nk_default_till(nk_till *till);
nk_status nk_energy_step(nk_solver *solver, double dt, int32_t grounded_melt_coupling, const nk_till *till);
/* compute_effective_pressure! with a till law (src/Groundings/subgrid_grounded_fraction.jl:168-183, effective_pressure
 * of till_effective_pressure.jl:66-73): N = min(N₀ (δ P̃/N₀)ˢ 10^((e₀/Cc)(1 − s)), Nᵒ), s = clamp(HW/Wᵐᵃˣ, 0, 1),
 * P₀ = ρᵢ g H, P̃ = P₀ (N₀/δ where P₀ = 0), Nᵒ = max(P₀ + ρw g min(BED, 0), 0). Replaces the N of update_geometry. */
nk_status nk_till_effective_pressure(nk_solver *solver, const nk_till *till);

/* Fixed-point adjoint of the converged SSA Picard solve (the EnzymeRules reverse rule of picard_solve!,
 * ext/NunatakEnzymeExt.jl:175-240, Christianson 1994), hand-written. Inputs: the converged solution in (U, V) and the
 * cost gradient x̄ in (UBAR, VBAR). Iterates λ = A⁻¹(x̄ + (∂q/∂x)ᵀ(−λ)) with q(x) = A(c(x)) x* until
 * ‖Δλ‖ ≤ max(reltol‖λ‖, abstol) under `control` (NULL: the Picard control, as the reference defaults), each linear
 * solve being the handle's PCG. Outputs: λ in (LAMBDAU, LAMBDAV) — the gradient with respect to the driving stress —
 * and the parameter gradients ∂J/∂C (CBAR) and ∂J/∂A (ABAR). DIVA carries the lagged state (βᵇ, F₂, ησ) through the
 * fixed point as the reference rule does; a cost on the layer velocities enters through (USIGMABAR, VSIGMABAR),
 * whose pullback through compute_velocity_profile! (src/StressBalances/diva.jl:89-109) gives the state's shadows.
 * Geometry: ∂J/∂H (HBAR) and ∂J/∂b (BEDBAR) through update_geometry! (s, μ, φᵘ, φᵛ, N;
 * src/Models/ice_sheet_model_methods.jl:103-111), the driving stress with its calving-front and grounding-line terms
 * (src/StressBalances/driving_stress.jl:56-125), the H μ factor of η̄H, the H factor of F₂ (DIVA) and the drag's φ
 * and N — the gradient of the whole update_velocities!(model) with geometry computed on the device (nk_update_velocities
 * with geometry = 1), as the reference's test differentiates update_state! (test/test_enzyme.jl:341-356).
 * With a FixedIterations(n) Picard control the converged-Picard rule does not apply (ext/NunatakEnzymeExt.jl:121 matches
 * Convergence only) and the reference's Enzyme differentiates the n unrolled steps: so does this call, SSA and DIVA.
 * It replays the handle's last FixedIterations Picard solve from the initial guess (and DIVA lagged state) that solve
 * recorded (NK_INVALID_ARGUMENT if there was none), keeps the n + 1 iterates (and states) on a device tape (2 planes
 * per step, 5 + Nσ for DIVA, allocated on first use and counted by nk_device_bytes), and runs each step's reverse:
 * λ_k = A(c(x_{k−1}))⁻¹ x̄_k (solve!'s rule, ext/NunatakEnzymeExt.jl:68-103), LAMBDA += λ_k, and the coefficient
 * pullback of A(c(x_{k−1})) x_k with −λ_k into x̄_{k−1} (DIVA: the lagged state's shadows too), CBAR, ABAR, HBAR and
 * BEDBAR; `control` is not used. The handle is left as the forward solve left it ((U, V) = x_n, its coefficients and
 * post-processing). info: picard_iterations = n, total_linear_iterations of the n adjoint solves. */
nk_status nk_adjoint(nk_solver *solver, const nk_control *control, nk_info *info);

/* Multigrid level shapes of the handle (StressBalanceMultigridPreconditioner(grid).levels). Writes up to
 * max_levels (Nx, Ny) pairs; returns the level count in *levels. */
nk_status nk_multigrid_levels(const nk_solver *solver, int32_t *levels, int32_t *shapes, int32_t max_levels);

/* Device memory a handle holds: nk_device_bytes of an existing handle, nk_estimate_device_bytes of one that
 * nk_create would build from these arguments (the same figure, without allocating). No reference counterpart:
 * the serving form's HBM admission uses them. */
nk_status nk_device_bytes(const nk_solver *solver, int64_t *bytes);
nk_status nk_estimate_device_bytes(const nk_grid *grid, const nk_physics *physics, const nk_solver_options *options,
                                   int64_t *bytes);

/* ---- Serving: continuous batching of velocity solves ---------------------------------------------------------
 * No reference counterpart (the reference solves one model at a time through update_velocities!,
 * src/Models/ice_sheet_model_methods.jl:119-147); each request is that call on its own inputs.
 *
 * A server owns `slots` worker threads, each driving one request at a time on its own HIP stream, and a pool of
 * solver handles keyed by (grid, options): a request reuses an idle handle of its size class (device buffers and
 * captured PCG graphs stay resident) or, when the HBM budget allows, a new one is built; idle handles of other
 * classes are evicted (least recently used first) to make room. Requests are admitted in arrival order when a slot
 * is free and their handle fits the budget, so requests that arrive while others run join the device work as slots
 * and memory free up. A reused handle has every field not given by the request reset to zero, so a request's
 * result does not depend on what ran before it. */
typedef struct nk_server nk_server;

typedef struct {
    int32_t slots;               /* requests in flight at once (worker threads, one HIP stream each), ≥ 1 */
    int32_t pad;
    int64_t hbm_budget_bytes;    /* device memory the server's handles may hold; 0 = 80 % of free device memory */
} nk_server_config;

typedef struct {
    int64_t id;
    int32_t status;              /* nk_status of the solve */
    int32_t reused_handle;       /* 1 when an idle handle of the size class was reused */
    int32_t picard_iterations;
    int32_t total_linear_iterations;
    double submit_ms;            /* server clock (ms since nk_server_create) */
    double admit_ms;             /* slot and memory granted */
    double start_ms;             /* inputs on the device, solve started */
    double finish_ms;            /* result copied back to host memory */
    int64_t device_bytes;        /* device memory of the handle that served the request */
} nk_request_info;

nk_status nk_server_create(const nk_server_config *config, nk_server **server);
/* Waits for every submitted request to finish, then frees the handles and joins the workers. */
nk_status nk_server_destroy(nk_server *server);
const char *nk_server_last_error(const nk_server *server);
/* Queue one update_velocities! (geometry on the device). fields[k] holds field field_ids[k] packed to the extents
 * of nk_field_shape (column-major, σ layers last); the server copies them before returning. U and V default to
 * zero (the initial guess); TEMPERATURE, when given, sets A through nk_rate_factor_from_temperature. */
nk_status nk_server_submit(nk_server *server, const nk_grid *grid, const nk_physics *physics,
                           const nk_solver_options *options, int32_t nfields, const int32_t *field_ids,
                           const double *const *fields, int64_t *request_id);
/* Non-blocking: *done = 1 when the request has finished. */
nk_status nk_server_poll(nk_server *server, int64_t request_id, int32_t *done);
/* Blocks until the request finishes; writes u and v packed to their nk_field_shape extents (either may be NULL)
 * and the request's record, then forgets the request. Returns the solve's status. */
nk_status nk_server_wait(nk_server *server, int64_t request_id, double *u, double *v, nk_request_info *info);

#ifdef __cplusplus
}
#endif
#endif /* NUNATAK_HIP_H */
