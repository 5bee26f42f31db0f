"""Thickness (and enthalpy) time step on the device: the split Runge-Kutta stage loop of time_step! around the
library's transport kernels (src/Models/time_stepping.jl:1-72) and, with thermodynamics, the energy step of the final
stage (maybe_step_thermodynamics!, src/Thermodynamics/column_energy_solve.jl:1-35).

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


class Thermodynamics:
    """EnthalpyThermodynamics switches of the step: the energy is advanced in the final stage of every time step
    (schedule IterationInterval(1)); `grounded_melt_coupling` takes the basal water's change from the thickness;
    `till` (an abi.Till) sets N from the basal water and drains the till (TillEffectivePressure)."""

    def __init__(self, grounded_melt_coupling=False, till=None):
        self.grounded_melt_coupling = grounded_melt_coupling
        self.till = till


def update_state(model, solve_velocities, thermodynamics=None, energy_dt=None):
    """update_state! (src/Models/ice_sheet_model_methods.jl:80-86): geometry, then the velocities if solved, then in
    the final stage of a step the energy (strain heating, then advance_enthalpy! over the step)."""
    till = thermodynamics.till if thermodynamics is not None else None
    if solve_velocities and till is None:
        model.update_velocities(geometry=True)
    else:
        model.update_geometry()
        if till is not None:
            model.till_effective_pressure(till)
        if solve_velocities:
            model.update_velocities(geometry=False)
    if thermodynamics is not None and energy_dt:
        model.compute_strain_heating()
        model.energy_step(energy_dt, thermodynamics.grounded_melt_coupling, till)


def time_step(model, dt, advection=abi.ADVECTION_UPWIND1, solve_velocities=False, first=False, thermodynamics=None):
    """One time_step!(model, Δt). `first`: the model's first step, which starts with update_state! (no energy step:
    no time has elapsed). `thermodynamics`: a Thermodynamics to advance the enthalpy in the final stage."""
    coupled = thermodynamics is not None and thermodynamics.grounded_melt_coupling
    if first:
        update_state(model, solve_velocities)
    model.cache_thickness()
    for stage, beta in enumerate(SPLIT_RK3, start=1):
        final = stage == len(SPLIT_RK3)
        model.compute_thickness_tendency(advection, coupled)
        model.thickness_substep(dt / beta, final)
        update_state(model, solve_velocities, thermodynamics, dt if final else None)
