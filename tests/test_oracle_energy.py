"""The enthalpy step in the oracle (src/Thermodynamics/column_energy_solve.jl, layer_transport.jl,
strain_heating.jl), driven through the product's stage loop `nunatak_hip.transport.time_step` with the energy step in
the final stage, against the reference's own column tests (test/test_thermodynamics.jl:1-232): steady conduction
(linear in height on uniform and refined layers, 1e-8), surface warming against the Fourier series (second order),
the energy budget with geothermal and frictional heating (1e-11), a warm base melting at the analytic rate and then
refreezing its water with a closed budget (1e-10), a floating base at the ocean temperature, ice-free columns at the
surface enthalpy. The column solve follows Oceananigans' BatchedTridiagonalSolver (a dependency outside the reference
tree: restated as the Thomas algorithm); these closed forms pin it."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import P, abi, grid
from nunatak_hip import transport

CI, K, TR, T0, BETA, L = 2009.0, 2.1, 223.15, 273.15, 7.9e-8, 3.34e5
TH = transport.Thermodynamics(grounded_melt_coupling=False)


def column_model(H0=1000.0, Ns=10, bed_refinement=1.0, G=0.042, Ts=243.15, grounding=abi.GROUNDING_NONE, bed=0.0,
                 cls=O.OracleModel):
    """column_model of test/test_thermodynamics.jl:3-15: one periodic 1 km column, ice density 910, the enthalpy
    starting at the surface value."""
    g = grid(1, 1, Lx=1e3, Ly=1e3, topology=(P, P), Nsigma=Ns, bed_refinement=bed_refinement)
    phys = abi.default_physics(masking=0, grounding=grounding, ice_density=910.0, rate_factor_depth=abi.RATE_FACTOR_3D)
    m = cls(g, phys, abi.default_options())
    m.set("H", np.full((1, 1), H0))
    m.set("HN", np.full((1, 1), H0))
    m.set("BED", np.full((1, 1), bed))
    m.set("GEOTHERMAL", np.full((1, 1), G))
    m.set("TSURF", np.full((1, 1), Ts))
    m.set("TOCEAN", np.full((1, 1), 271.15))
    m.set("ENTHALPY", np.full((Ns, 1, 1), CI * (min(Ts, T0) - TR)))
    m.update_geometry()
    return m, g


def centers(m):
    return np.array(m.sigma_metrics()[1])


def faces(m):
    return np.array(m.sigma_metrics()[0])


def col(m, name):
    return m.get(name)[:, 0, 0].copy()


def steps(m, dt, n, first=False):
    for k in range(n):
        transport.time_step(m, dt, thermodynamics=TH, first=first and k == 0)


def set_temperature(m, T):
    m.set("TEMPERATURE", T.reshape(-1, 1, 1))
    m.rate_factor_from_temperature()   # set!(model; T): E = enthalpy(T, 0, p)


def column_energy(m, H0):
    return 910.0 * H0 * float(np.sum(np.diff(faces(m)) * col(m, "ENTHALPY")))


def surface_heat_flux(m, H0, Ts):
    zN = centers(m)[-1]
    EN = col(m, "ENTHALPY")[-1]
    return K / CI * (EN - CI * (Ts - TR)) / (H0 * (1 - zN))


@pytest.mark.parametrize("bed_refinement", [1.0, 2.0])
def test_steady_conduction_is_linear_in_height(bed_refinement):
    H0, G, Ts = 1000.0, 0.042, 243.15
    m, g = column_model(H0=H0, G=G, Ts=Ts, bed_refinement=bed_refinement)
    steps(m, 1e14, 4, first=True)
    exact = Ts + G * H0 * (1 - centers(m)) / K
    assert np.max(np.abs(col(m, "TEMPERATURE") - exact)) < 1e-8
    assert m.get("HW")[0, 0] == 0


def test_surface_warming_converges_at_second_order_to_the_fourier_series():
    H0, G, Ts, dT, t_end = 1000.0, 0.042, 243.15, 10.0, 6.3e10
    errors = []
    for Ns in (10, 20, 40):
        m, g = column_model(H0=H0, Ns=Ns, G=G, Ts=Ts)
        z = centers(m)
        set_temperature(m, Ts + G * H0 * (1 - z) / K)
        m.set("TSURF", np.full((1, 1), Ts + dT))
        Nt = 4 * Ns ** 2
        steps(m, t_end / Nt, Nt, first=True)
        kappa = K / (910.0 * CI)
        zz = z * H0
        steady = Ts + dT + G * (H0 - zz) / K
        transient = np.zeros_like(zz)
        for n in range(201):
            lam = (2 * n + 1) * math.pi / (2 * H0)
            transient += -dT * 4 / math.pi * (-1) ** n / (2 * n + 1) * np.cos(lam * zz) * math.exp(-kappa * lam ** 2 * t_end)
        errors.append(float(np.max(np.abs(col(m, "TEMPERATURE") - (steady + transient)))))
    assert errors[2] < 0.01
    assert errors[0] / errors[1] > 3 and errors[1] / errors[2] > 3


def test_energy_budget_closes_with_geothermal_and_frictional_heating():
    H0, G = 800.0, 0.05
    m, g = column_model(H0=H0, G=G, Ns=12, bed_refinement=2.0)
    set_temperature(m, 250 - 5 * centers(m))
    beta, ub = 1e8, 1e-6
    m.set("BETAU", np.full(m.shape("BETAU")[1::-1], beta))
    m.set("UB", np.full(m.shape("UB")[1::-1], ub))
    F = beta * ub ** 2
    dt = 1e9
    m.update_geometry()
    for k in range(50):
        e0 = column_energy(m, H0)
        transport.time_step(m, dt, thermodynamics=TH, first=(k == 0))
        budget = column_energy(m, H0) - e0 - dt * (G + F - surface_heat_flux(m, H0, 243.15))
        assert abs(budget) < 1e-11 * abs(e0)


def test_warm_base_melts_at_the_analytic_rate_then_refreezes_the_stored_water():
    H0, G, Ts = 1000.0, 0.1, 263.15
    m, g = column_model(H0=H0, G=G, Ts=Ts, Ns=20)
    rho = 910.0
    Tpm = T0 - BETA * rho * 9.81 * H0
    steps(m, 1e13, 20, first=True)
    exact = (G - K * (Tpm - Ts) / H0) / (rho * L)
    assert exact > 0
    assert math.isclose(m.get("GROUNDED_MELT")[0, 0], exact, rel_tol=1e-3)
    assert m.get("HW")[0, 0] > 0
    m.set("HW", np.full((1, 1), 1e-2))
    m.set("GEOTHERMAL", np.zeros((1, 1)))
    dt = 1e9

    def total_energy():
        return column_energy(m, H0) + rho * L * m.get("HW")[0, 0]

    froze_out = False
    for _ in range(200):
        e0 = total_energy()
        transport.time_step(m, dt, thermodynamics=TH)
        q = K / CI * (col(m, "ENTHALPY")[-1] - CI * (Ts - TR)) / (H0 * (1 - centers(m)[-1]))
        assert abs(total_energy() - e0 + dt * q) < 1e-10 * abs(e0)
        froze_out |= m.get("HW")[0, 0] == 0
    assert froze_out
    assert m.get("GROUNDED_MELT")[0, 0] == 0


def test_floating_base_sits_at_the_ocean_temperature_with_no_grounded_melt():
    H0, Ts = 500.0, 253.15
    m, g = column_model(H0=H0, Ts=Ts, G=0.05, Ns=10, grounding=abi.GROUNDING_SUBGRID, bed=-3000.0)
    assert m.get("PHI")[0, 0] == 0
    steps(m, 1e14, 4, first=True)
    To = 271.15
    exact = To + (Ts - To) * centers(m)
    assert np.max(np.abs(col(m, "TEMPERATURE") - exact)) < 1e-8
    assert m.get("GROUNDED_MELT")[0, 0] == 0


def test_ice_free_columns_hold_the_surface_enthalpy():
    Ts = 253.15
    m, g = column_model(H0=0.5, Ts=Ts)
    set_temperature(m, np.full(10, 240.0))
    steps(m, 1e9, 1, first=True)
    assert np.allclose(col(m, "TEMPERATURE"), Ts, rtol=1e-14)
    assert m.get("HW")[0, 0] == 0
