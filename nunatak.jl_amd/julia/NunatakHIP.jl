"""
    NunatakHIP

Julia side of the drop-in: an `HIPStressBalanceSolver` that Nunatak's `update_velocities!` dispatches to, backed by
libnunatak_hip.so through `ccall`. It is the binding a maintainer adds next to `StressBalanceSolver`
(src/StressBalances/stress_balance_solver.jl); the model keeps its Oceananigans fields and the library keeps a
device copy of the solve's state between calls.

Usage:

    using Nunatak, NunatakHIP
    solver = HIPStressBalanceSolver(grid; preconditioner = :multigrid, precision = :mixed)
    model = IceSheetModel(grid; velocity_solver = solver, rheology, friction, grounding, masking)
    set!(model; H)          # update_state! → update_velocities!(model, solver, model.stress_balance)
"""
module NunatakHIP

using Oceananigans
using Oceananigans.Fields: interior
using Oceananigans.Grids: topology, halo_size
using Nunatak
import Nunatak: update_velocities!

const libnunatak_hip = get(ENV, "NUNATAK_HIP_LIB", joinpath(@__DIR__, "..", "libnunatak_hip.so"))

# ---- mirrors of include/nunatak_hip.h ---------------------------------------------------------------------------

struct NKGrid
    Nx::Int32; Ny::Int32; topology_x::Int32; topology_y::Int32; halo::Int32; Nsigma::Int32
    dx::Float64; dy::Float64; bed_refinement::Float64
end

struct NKPhysics
    ice_density::Float64; ocean_density::Float64; gravitational_acceleration::Float64
    grounding::Int32; masking::Int32
    mask_minimum_thickness::Float64; mask_transition_thickness::Float64
    glen_exponent::Float64; strain_rate_regularization::Float64
    rate_factor_depth::Int32; friction::Int32
    friction_exponent::Float64; coulomb_coefficient::Float64; velocity_regularization::Float64
    walls::Int32; stress_balance::Int32
end

struct NKControl
    kind::Int32; maxiter::Int32; iterations::Int32; pad::Int32
    reltol::Float64; abstol::Float64
end

struct NKOptions
    linear::NKControl; picard::NKControl
    preconditioner::Int32; smoother::Int32; smoothing_sweeps::Int32; coarsest_sweeps::Int32
    minimum_coarse_size::Int32; precision::Int32
    coarse_operator::Int32; check_interval::Int32; profile_kernel::Int32; graphs::Int32; report_not_converged::Int32
end

mutable struct NKInfo
    picard_iterations::Int32; linear_iterations::Int32; total_linear_iterations::Int64
    refinement_steps::Int32; pad::Int32
    last_picard_change::Float64; last_picard_norm::Float64; last_residual_norm::Float64; elapsed_ms::Float64
    profiled_ms::Float64; profiled_launches::Int64; profiled_points::Int64; profiled_precision::Int32; pad2::Int32
    NKInfo() = new(0, 0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, 0.0, 0, 0, 0, 0)
end

# field ids (nk_field_id)
const U, V, H, BED, S, PHI, PHIU, PHIV, N, MU, TAUU, TAUV, ETAH, ETAHFF, BETAU, BETAV, F2, ETASIGMA,
      BETABU, BETABV, UB, VB, USIGMA, VSIGMA, A, C, TEMPERATURE, PU, PV, QU, QV,
      UBAR, VBAR, LAMBDAU, LAMBDAV, CBAR, ABAR, USIGMABAR, VSIGMABAR, ENTHALPY, WATER_FRACTION,
      SMB, BASAL_MELT, GROUNDED_MELT, HPREV, GH, HCLIP, HW, HN, GEOTHERMAL, TSURF, TOCEAN, STRAIN_HEATING,
      HBAR, BEDBAR = 0:54

control(c::Convergence) = NKControl(0, c.maxiter, 0, 0, c.reltol, c.abstol)
control(c::FixedIterations) = NKControl(1, c.iterations, c.iterations, 0, 0.0, 0.0)

"""
    HIPStressBalanceSolver(grid; linear_iteration_control, picard_iteration_control, preconditioner, precision, schedule)

Drop-in for `StressBalanceSolver` whose velocity solve runs in libnunatak_hip.so on an MI355X. `preconditioner` is
`:diagonal` (the reference default) or `:multigrid`; `precision` is `:fp64` or `:mixed` (FP64 PCG — operator, residual,
α/β and stopping rule — preconditioned by an FP32 multigrid V-cycle whose coarsest-grid solve is FP64; multigrid
only, best with `coarse_operator = :galerkin`). `Nσ` and `bed_refinement` are the model's σ grid (IceSheetModel's
defaults); the first solve checks them against `model.σgrid` and rebuilds the handle if they differ. `iteration`
holds the Picard count of the last solve, like the reference's.
"""
mutable struct HIPStressBalanceSolver{Sc}
    handle::Ptr{Cvoid}
    schedule::Sc
    iteration::Int
    grid::NKGrid
    options::NKOptions
    info::NKInfo
end

function check(status, handle, what)
    status == 0 && return nothing
    message = handle == C_NULL ? "" : unsafe_string(ccall((:nk_last_error, libnunatak_hip), Cstring, (Ptr{Cvoid},), handle))
    status == 1 && throw(ArgumentError("$what: $message"))
    error("$what failed with status $status: $message")
end

function HIPStressBalanceSolver(grid; linear_iteration_control = Convergence(eltype(grid); reltol = 1e-8, maxiter = 10_000),
                                picard_iteration_control = Convergence(eltype(grid); reltol = 1e-6, maxiter = 100),
                                preconditioner = :diagonal, precision = :fp64, coarse_operator = :rediscretized,
                                schedule = IterationInterval(1), Nσ = 10, bed_refinement = 2,
                                physics = nothing)
    Nx, Ny, _ = size(grid)
    TX, TY, _ = topology(grid)
    Hx, _, _ = halo_size(grid)
    g = NKGrid(Nx, Ny, TX === Bounded, TY === Bounded, Hx, Nσ, grid.Δxᶜᵃᵃ, grid.Δyᵃᶜᵃ, bed_refinement)
    o = NKOptions(control(linear_iteration_control), control(picard_iteration_control),
                  preconditioner === :multigrid ? 2 : 1, 0, 2, 20, 2, precision === :mixed ? 1 : 0,
                  coarse_operator === :galerkin ? 1 : 0, 8, 0, 1, 0)
    solver = HIPStressBalanceSolver(create_handle(g, o), schedule, 0, g, o, NKInfo())
    finalizer(s -> ccall((:nk_destroy, libnunatak_hip), Cint, (Ptr{Cvoid},), s.handle), solver)
    return solver
end

function create_handle(g, o)
    h = Ref{Ptr{Cvoid}}(C_NULL)   # physics are set from the model at every solve
    check(ccall((:nk_create, libnunatak_hip), Cint, (Ref{NKGrid}, Ref{NKPhysics}, Ref{NKOptions}, Ref{Ptr{Cvoid}}),
                g, default_physics(), o, h), h[], "nk_create")
    return h[]
end

# Nσ and bed_refinement of the model's SigmaGrid (src/Grids/sigma_grid.jl:20-22,43: faces (k/Nσ)^bed_refinement); the
# handle is rebuilt when they differ from the ones it was created with, so the column weights always match the model.
function match_sigma_grid!(solver::HIPStressBalanceSolver, model)
    σgrid = model.σgrid
    Nσ = size(σgrid, 3)
    ζ = znodes(σgrid, Face())
    refinement = Nσ > 1 ? log(ζ[2]) / log(1 / Nσ) : solver.grid.bed_refinement
    refinement = isapprox(refinement, round(refinement); rtol = 1e-12) ? round(refinement) : refinement
    (Nσ == solver.grid.Nsigma && refinement == solver.grid.bed_refinement) && return nothing
    g = solver.grid
    solver.grid = NKGrid(g.Nx, g.Ny, g.topology_x, g.topology_y, g.halo, Nσ, g.dx, g.dy, refinement)
    ccall((:nk_destroy, libnunatak_hip), Cint, (Ptr{Cvoid},), solver.handle)
    solver.handle = create_handle(solver.grid, solver.options)
    return nothing
end

function default_physics()
    p = Ref{NKPhysics}()
    ccall((:nk_default_physics, libnunatak_hip), Cvoid, (Ref{NKPhysics},), p)
    return p[]
end

friction_kind(::Nothing) = 0
friction_kind(::WeertmanFriction) = 1
friction_kind(::RegularizedCoulombFriction) = 2
coulomb(f) = 0.0
coulomb(f::RegularizedCoulombFriction) = f.coulomb_coefficient
walls_kind(::FreeSlipWalls) = 0
walls_kind(::NoSlipWalls) = 1

function physics(model)
    c, r, f, sb = model.constants, model.rheology, model.friction, model.stress_balance
    grounding = model.grounding === nothing ? 0 : (model.grounding isa SubgridGroundedFraction{true} ? 2 : 1)
    A = Nunatak.rate_factor_field(r.rate_factor)
    depth = A isa Field{<:Any, <:Any, Nothing} ? 0 : 1
    return NKPhysics(c.ice_density, c.ocean_density, c.gravitational_acceleration, grounding, 1,
                     model.masking.minimum_thickness, model.masking.transition_thickness,
                     r.exponent, r.strain_rate_regularization, depth, friction_kind(f),
                     f === nothing ? 1/3 : f.exponent, coulomb(f), f === nothing ? 1e-8 : f.velocity_regularization,
                     walls_kind(sb.side_walls), sb.shear === nothing ? 0 : 1)
end

# Host pointer of interior element (1, 1[, 1]) of a field's parent array and its leading dimension / layer stride.
function strided(f)
    P = parent(f)
    Hx, Hy = halo_size(f.grid)[1:2]
    offset = Hx + Hy * size(P, 1) + (ndims(P) == 3 && size(P, 3) > 1 ? size(P, 1) * size(P, 2) : 0)
    return pointer(P, offset + 1), size(P, 1), size(P, 1) * size(P, 2)
end

function set_field!(s, id, f)
    p, ld, lds = strided(f)
    check(ccall((:nk_set_field, libnunatak_hip), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Int64, Int64), s.handle, id, p, ld, lds),
          s.handle, "nk_set_field")
end

function get_field!(f, s, id)
    p, ld, lds = strided(f)
    check(ccall((:nk_get_field, libnunatak_hip), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Int64, Int64), s.handle, id, p, ld, lds),
          s.handle, "nk_get_field")
end

# update_velocities!(model, solver, ::SSA / ::DIVA) (src/Models/ice_sheet_model_methods.jl:119-147): the geometry
# fields come from update_geometry!, the velocities are the initial guess and are written back with the solution.
function update_velocities!(model, solver::HIPStressBalanceSolver, stress_balance)
    match_sigma_grid!(solver, model)
    check(ccall((:nk_set_physics, libnunatak_hip), Cint, (Ptr{Cvoid}, Ref{NKPhysics}), solver.handle, physics(model)),
          solver.handle, "nk_set_physics")
    (; H, s, φᵘ, φᵛ, N, ice_mask) = model.geometry
    aux, vel = model.auxiliary, model.velocities
    set_field!(solver, NunatakHIP.H, H); set_field!(solver, S, s); set_field!(solver, PHIU, φᵘ); set_field!(solver, PHIV, φᵛ)
    set_field!(solver, NunatakHIP.N, N); set_field!(solver, MU, ice_mask); set_field!(solver, BED, model.bed)
    set_field!(solver, NunatakHIP.A, Nunatak.rate_factor_field(model.rheology.rate_factor))
    model.friction === nothing || set_field!(solver, NunatakHIP.C, model.friction.coefficient)
    set_field!(solver, U, vel.u); set_field!(solver, V, vel.v)
    for (id, f) in ((ETASIGMA, aux.ησ), (F2, aux.F₂), (BETABU, aux.βᵇᵘ), (BETABV, aux.βᵇᵛ), (BETAU, aux.βᵘ), (BETAV, aux.βᵛ))
        set_field!(solver, id, f)
    end
    info = solver.info
    check(ccall((:nk_update_velocities, libnunatak_hip), Cint, (Ptr{Cvoid}, Int32, Ref{NKInfo}), solver.handle, 0, info),
          solver.handle, "nk_update_velocities")
    solver.iteration = info.picard_iterations
    for (f, id) in ((vel.u, U), (vel.v, V), (vel.uᵇ, UB), (vel.vᵇ, VB), (vel.uσ, USIGMA), (vel.vσ, VSIGMA),
                    (aux.ησ, ETASIGMA), (aux.η̄H, ETAH), (aux.η̄Hᶠᶠ, ETAHFF), (aux.F₂, F2), (aux.βᵘ, BETAU), (aux.βᵛ, BETAV),
                    (aux.βᵇᵘ, BETABU), (aux.βᵇᵛ, BETABV), (aux.τᵘ, TAUU), (aux.τᵛ, TAUV))
        get_field!(f, solver, id)
    end
    return nothing
end

update_velocities!(model, solver::HIPStressBalanceSolver) = update_velocities!(model, solver, model.stress_balance)

# Reverse pass of a converged SSA or DIVA velocity solve (the fixed-point rule of ext/NunatakEnzymeExt.jl:175-240): given the
# cost gradient with respect to the velocities (ū, v̄) and, for DIVA, the layer velocities (ūσ, v̄σ, pulled back through
# compute_velocity_profile!), accumulate ∂J/∂τ into (τ̄ᵘ, τ̄ᵛ), ∂J/∂C into C̄, ∂J/∂A into Ā, and ∂J/∂H, ∂J/∂b into H̄, b̄
# (through the driving stress, the geometry diagnostics s, μ, φ, N of update_geometry!, the H μ factor of η̄H and the
# drag). Any output may be `nothing`. The handle must still hold the converged solution of the forward call.
function hip_velocity_adjoint!(τ̄ᵘ, τ̄ᵛ, C̄, Ā, solver::HIPStressBalanceSolver, ū, v̄; ūσ = nothing, v̄σ = nothing,
                               H̄ = nothing, b̄ = nothing, control = nothing)
    set_field!(solver, UBAR, ū); set_field!(solver, VBAR, v̄)
    for (id, f) in ((USIGMABAR, ūσ), (VSIGMABAR, v̄σ))
        f === nothing ? check(ccall((:nk_fill_field, libnunatak_hip), Cint, (Ptr{Cvoid}, Cint, Float64), solver.handle, id, 0.0),
                              solver.handle, "nk_fill_field") : set_field!(solver, id, f)
    end
    info = solver.info
    c = control === nothing ? C_NULL : Ref(NunatakHIP.control(control))
    check(ccall((:nk_adjoint, libnunatak_hip), Cint, (Ptr{Cvoid}, Ptr{NKControl}, Ref{NKInfo}), solver.handle, c, info),
          solver.handle, "nk_adjoint")
    for (f, id) in ((τ̄ᵘ, LAMBDAU), (τ̄ᵛ, LAMBDAV), (C̄, CBAR), (Ā, ABAR), (H̄, HBAR), (b̄, BEDBAR))
        f === nothing && continue
        g = similar(f); get_field!(g, solver, id)
        interior(f) .+= interior(g)
    end
    return info.picard_iterations
end

# The library's code of the model's thickness advection: UpwindBiased(order = 1) 0, Centered(order = 2) 1, WENO(order = 5)
# 2, WENO(order = 5, bounds) 3 (the reference's default, src/Models/ice_sheet_model.jl:123; its bounds are passed on).
advection_code(a) = throw(ArgumentError("thickness advection $(summary(a)) is not built in libnunatak_hip"))
advection_code(::Oceananigans.Advection.UpwindBiased{1}) = 0
advection_code(::Oceananigans.Advection.Centered{1}) = 1
advection_code(::Oceananigans.Advection.WENO{3}) = 2
advection_code(::Oceananigans.Advection.BoundsPreservingWENO{3}) = 3

# (lower, upper) of a bounds-preserving scheme's bounds object (which also carries the limiter field)
thickness_bounds(b::Tuple) = (b[1], b[2])
thickness_bounds(b) = hasproperty(b, :lower) ? (b.lower, b.upper) : (b.bounds[1], b.bounds[2])

# One split Runge-Kutta stage of the thickness on the device (rk_substep!, src/Models/time_stepping.jl:15-72) for a
# model whose velocities are current: the tendency with the model's advection scheme (advection_code), then
# H = max(0, H⁻ + Δτ G), the clip accumulated in the final stage.
function hip_rk_substep!(model, solver::HIPStressBalanceSolver, Δτ, final_stage; advection = advection_code(model.advection))
    if advection == 3
        lower, upper = thickness_bounds(model.advection.bounds)
        check(ccall((:nk_set_thickness_bounds, libnunatak_hip), Cint, (Ptr{Cvoid}, Float64, Float64), solver.handle, lower, upper),
              solver.handle, "nk_set_thickness_bounds")
    end
    geometry, vel = model.geometry, model.velocities
    set_field!(solver, NunatakHIP.H, geometry.H); set_field!(solver, HPREV, model.timestepper.Ψ⁻.H)
    set_field!(solver, U, vel.u); set_field!(solver, V, vel.v); set_field!(solver, PHI, geometry.φ)
    set_field!(solver, SMB, model.surface_mass_balance); set_field!(solver, BASAL_MELT, model.basal_melt_rate)
    set_field!(solver, HCLIP, geometry.clipped_thickness)
    coupled = model.thermodynamics !== nothing && model.thermodynamics.grounded_melt_coupling
    coupled && set_field!(solver, GROUNDED_MELT, model.thermodynamics.state.mᵍ)
    check(ccall((:nk_compute_thickness_tendency, libnunatak_hip), Cint, (Ptr{Cvoid}, Int32, Int32), solver.handle, advection, coupled),
          solver.handle, "nk_compute_thickness_tendency")
    check(ccall((:nk_thickness_substep, libnunatak_hip), Cint, (Ptr{Cvoid}, Float64, Int32), solver.handle, Δτ, final_stage),
          solver.handle, "nk_thickness_substep")
    get_field!(geometry.H, solver, NunatakHIP.H); get_field!(geometry.clipped_thickness, solver, HCLIP)
    return nothing
end

Base.summary(s::HIPStressBalanceSolver) = "HIPStressBalanceSolver(libnunatak_hip, iteration = $(s.iteration))"

export HIPStressBalanceSolver, hip_velocity_adjoint!, hip_rk_substep!

end # module
