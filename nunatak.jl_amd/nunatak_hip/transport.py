"""Thickness time step on the device: the split Runge-Kutta stage loop of time_step! around the library's transport
kernels (src/Models/time_stepping.jl:1-72).

The stage loop follows Oceananigans' SplitRungeKutta3 time stepper, which the reference selects by default
(src/Models/ice_sheet_model.jl:95): the thickness at the start of the step is cached (cache_current_fields!), then
each stage m = 1, 2, 3 sets H = H⁻ + (Δt/βₘ)·G(H) with β = (3, 2, 1) and runs update_state! (geometry, then the
velocity solve when the model solves for velocities). The clip's thickness is accumulated in the final stage only.
Oceananigans (0.117) is a dependency outside the reference tree: the stage coefficients are restated from its
published scheme, and the tests pin the loop by the reference's own transport properties (volume conservation,
source budget, centre-of-mass speed, clip budget, floating-only melt).
"""
from . import abi

SPLIT_RK3 = (3.0, 2.0, 1.0)


def update_state(model, solve_velocities):
    """update_state! (src/Models/ice_sheet_model_methods.jl:80-86): geometry, then the velocities if solved."""
    if solve_velocities:
        return model.update_velocities(geometry=True)
    model.update_geometry()
    return None


def time_step(model, dt, advection=abi.ADVECTION_UPWIND1, solve_velocities=False, first=False):
    """One time_step!(model, Δt) of the thickness. `first`: the model's first step, which starts with update_state!."""
    if first:
        update_state(model, solve_velocities)
    model.cache_thickness()
    for stage, beta in enumerate(SPLIT_RK3, start=1):
        model.compute_thickness_tendency(advection)
        model.thickness_substep(dt / beta, stage == len(SPLIT_RK3))
        update_state(model, solve_velocities)
