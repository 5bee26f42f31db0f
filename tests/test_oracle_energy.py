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


# ---- further reference column and patch tests (test/test_thermodynamics.jl:293-470) ------------------------------

def test_thickening_or_thinning_column_follows_the_heat_equation_in_fixed_height():
    H0, G, Tb, Bamp = 1000.0, 0.04, 240.0, 5.0
    lam, t_end = math.pi / 1000, 2e10
    for a0 in (1e-8, -1e-8):
        errors = []
        for Ns in (10, 20, 40):
            m, g = column_model(H0=H0, Ns=Ns, G=G)
            m.set("SMB", np.full((1, 1), a0))
            kappa = K / (910.0 * CI)

            def exact(z, t):
                return Tb - G * z / K + Bamp * math.exp(-kappa * lam ** 2 * t) * np.cos(lam * z)

            set_temperature(m, exact(centers(m) * H0, 0.0))
            Nt = 4 * Ns ** 2
            dt = t_end / Nt
            for n in range(1, Nt + 1):
                m.set("TSURF", np.full((1, 1), exact(H0 + a0 * n * dt, n * dt)))
                transport.time_step(m, dt, thermodynamics=TH, first=(n == 1))
            H = m.get("H")[0, 0]
            assert math.isclose(H, H0 + a0 * t_end, rel_tol=1e-12)
            errors.append(float(np.max(np.abs(col(m, "TEMPERATURE") - exact(centers(m) * H, t_end)))))
        assert errors[2] < 0.03
        assert errors[0] / errors[1] > 2.5 and errors[1] / errors[2] > 2.5


def patch_model(topology, T=250.0, Ts=250.0, G=0.0, cls=O.OracleModel):
    """patch_model of test/test_thermodynamics.jl:44-58 (thickness advected upwind here): 8 × 8, variable thickness,
    flow shearing with depth, accumulation and melt."""
    L = 8e4
    g = grid(8, 8, Lx=L, Ly=L, topology=topology, Nsigma=10, bed_refinement=2.0)
    phys = abi.default_physics(masking=0, grounding=abi.GROUNDING_NONE, ice_density=910.0, rate_factor_depth=abi.RATE_FACTOR_3D)
    m = cls(g, phys, abi.default_options())
    from helpers import field_values
    ubar = lambda x, y: 1e-6 * (1 + np.sin(2 * np.pi * y / L) / 2)  # noqa: E731
    vbar = lambda x, y: 5e-7 * np.cos(2 * np.pi * x / L)  # noqa: E731
    z = np.array(O.OracleModel(g, phys, abi.default_options()).sigma_metrics()[1])
    H = field_values(g, "H", lambda x, y: 1000 + 200 * np.sin(2 * np.pi * x / L) * np.cos(2 * np.pi * y / L))
    m.set("H", H)
    m.set("HN", H)
    m.set("BED", field_values(g, "BED", -5000.0))
    m.set("U", field_values(g, "U", ubar))
    m.set("V", field_values(g, "V", vbar))
    m.set("USIGMA", np.stack([field_values(g, "U", ubar) * (0.5 + zk) for zk in z]))
    m.set("VSIGMA", np.stack([field_values(g, "V", vbar) * (0.5 + zk) for zk in z]))
    m.set("SMB", field_values(g, "SMB", lambda x, y: 1e-8 * (1 + np.cos(2 * np.pi * x / L))))
    m.set("BASAL_MELT", field_values(g, "BASAL_MELT", 2e-9))
    m.set("GEOTHERMAL", field_values(g, "H", G))
    m.set("TSURF", field_values(g, "H", Ts))
    m.set("TOCEAN", field_values(g, "H", 271.15))
    Tf = T(g, z) if callable(T) else np.full((10, 8, 8), T)
    m.set("TEMPERATURE", Tf)
    m.rate_factor_from_temperature()
    m.update_geometry()
    return m, g, z


def test_uniform_enthalpy_stays_uniform_under_variable_thickness_shear_accumulation_and_melt():
    m, g, z = patch_model((P, P))
    E0 = m.get("ENTHALPY")[0, 0, 0]
    H0 = m.get("H").copy()
    for k in range(20):
        transport.time_step(m, 1e9, thermodynamics=TH, first=(k == 0))
    assert np.max(np.abs(m.get("H") - H0)) > 10
    assert np.max(np.abs(m.get("ENTHALPY") - E0)) < 1e-9 * E0


def test_strong_accumulation_over_a_sharp_front_creates_no_new_extrema():
    m, g = column_model(H0=1000.0, Ns=10, G=0.0, Ts=250.0)
    m.set("SMB", np.full((1, 1), 2e-7))
    set_temperature(m, np.where(centers(m) < 0.5, 230.0, 260.0))
    for k in range(20):
        transport.time_step(m, 1e8, thermodynamics=TH, first=(k == 0))
        T = col(m, "TEMPERATURE")
        assert T.min() >= 230 - 1e-8 and T.max() <= 260 + 1e-8


def test_temperate_layers_drain_towards_the_drainage_water_fraction():
    H0, wd, Rmax = 1000.0, 0.01, 1.58443823077064e-9
    m, g = column_model(H0=H0, G=0.0, Ts=263.15, Ns=10)
    m.set("ENTHALPY", np.full((10, 1, 1), CI * (T0 - TR) + 5 * wd * L))
    m.thermal_diagnostics()
    assert col(m, "WATER_FRACTION").min() > 3 * wd
    transport.time_step(m, 1.0, thermodynamics=TH, first=True)
    assert math.isclose(m.get("HW")[0, 0], H0 * Rmax * 1.0, rel_tol=1e-8)
    for _ in range(100):
        transport.time_step(m, 1e9, thermodynamics=TH)
    assert col(m, "WATER_FRACTION").max() <= wd + 1e-12
    assert m.get("HW")[0, 0] > 0.1 * H0 * wd


def test_coupled_grounded_melt_moves_mass_between_the_ice_and_the_basal_water():
    H0, G, Ts = 1000.0, 0.1, 263.15
    coupled, _ = column_model(H0=H0, G=G, Ts=Ts, Ns=20)
    fixed, _ = column_model(H0=H0, G=G, Ts=Ts, Ns=20)
    TC = transport.Thermodynamics(grounded_melt_coupling=True)
    for mm in (coupled, fixed):
        set_temperature(mm, Ts + 10 * (1 - centers(mm)))
    dt = 1e10
    for k in range(30):
        transport.time_step(coupled, dt, thermodynamics=TC, first=(k == 0))
        transport.time_step(fixed, dt, thermodynamics=TH, first=(k == 0))
        total = coupled.get("H")[0, 0] + coupled.get("HW")[0, 0] - dt * coupled.get("GROUNDED_MELT")[0, 0]
        assert math.isclose(total, H0, rel_tol=1e-12)
        assert fixed.get("H")[0, 0] == H0
    assert coupled.get("GROUNDED_MELT")[0, 0] > 0
    assert coupled.get("H")[0, 0] < H0


def test_a_cold_base_leaves_the_thickness_as_a_model_without_thermodynamics():
    from helpers import B, field_values
    g = grid(8, 1, Lx=8e3, Ly=1e3, topology=(B, P), Nsigma=10)
    phys = abi.default_physics(masking=0, grounding=abi.GROUNDING_NONE, rate_factor_depth=abi.RATE_FACTOR_3D)
    models = []
    for thermal in (False, True):
        m = O.OracleModel(g, phys, abi.default_options())
        m.set("H", field_values(g, "H", lambda x, y: 100 + x / 100))
        m.set("U", field_values(g, "U", lambda x, y: np.where((x > 0) & (x < 8e3), 1e-6, 0.0)))
        m.set("BED", field_values(g, "BED", 0.0))
        if thermal:
            m.set("HN", m.get("H"))
            m.set("GEOTHERMAL", field_values(g, "H", 0.05))
            m.set("TSURF", field_values(g, "H", 243.15))
            m.set("TOCEAN", field_values(g, "H", 271.15))
            m.set("ENTHALPY", np.full((10, 1, 8), CI * (243.15 - TR)))
        for k in range(5):
            transport.time_step(m, 1e7, thermodynamics=transport.Thermodynamics(True) if thermal else None, first=(k == 0))
        models.append(m)
    assert np.array_equal(models[0].get("H"), models[1].get("H"))


def dissipation_mismatch(stress_balance, cls=O.OracleModel):
    """dissipation_mismatch of test/test_thermodynamics.jl:60-96: one time step of a 16 × 16 doubly periodic sheet
    with the velocity solve in every stage; the frictional heat qᵍ plus the column strain heat H Σ Δζ Φ against the
    power of the driving stress Σ τ·u."""
    from helpers import field_values
    Nx, L = 16, 3.2e5
    g = grid(Nx, Nx, Lx=L, Ly=L, topology=(P, P), Nsigma=10)
    phys = abi.default_physics(masking=0, grounding=abi.GROUNDING_NONE, ice_density=910.0, stress_balance=stress_balance,
                               friction=abi.FRICTION_WEERTMAN, friction_exponent=1.0, strain_rate_regularization=1e-16,
                               rate_factor_depth=abi.RATE_FACTOR_2D)
    opts = abi.default_options(linear=abi.convergence(1e-12, 0.0, 10_000), picard=abi.convergence(1e-11, 0.0, 500))
    m = cls(g, phys, opts)
    m.set("H", field_values(g, "H", lambda x, y: 1000 + 100 * np.sin(2 * np.pi * x / L) * np.cos(2 * np.pi * y / L)))
    m.set("HN", m.get("H"))
    m.set("BED", field_values(g, "BED", 0.0))
    m.set("A", field_values(g, "H", 1e-24))
    m.set("C", field_values(g, "H", 1e10))
    m.set("TSURF", field_values(g, "H", 243.15))
    m.set("TOCEAN", field_values(g, "H", 271.15))
    m.set("ENTHALPY", np.full((10, Nx, Nx), CI * (243.15 - TR)))
    transport.time_step(m, 1e6, solve_velocities=True, first=True, thermodynamics=transport.Thermodynamics(True))
    zf = np.array(O.OracleModel(g, phys, opts).sigma_metrics()[0])
    strain = np.tensordot(np.diff(zf), m.get("STRAIN_HEATING"), axes=1)
    diva = stress_balance == abi.DIVA
    bu, bv = m.get("BETABU" if diva else "BETAU"), m.get("BETABV" if diva else "BETAV")
    ub, vb = m.get("UB"), m.get("VB")
    fu, fv = bu * ub * ub, bv * vb * vb
    F = (fu + np.roll(fu, -1, axis=1) + fv + np.roll(fv, -1, axis=0)) / 2
    heat = (L / Nx) ** 2 * float(np.sum(F + m.get("H") * strain))
    power = float(np.sum(m.get("TAUU") * m.get("U")) + np.sum(m.get("TAUV") * m.get("V")))
    return abs(heat - power) / power


@pytest.mark.parametrize("stress_balance", [abi.SSA, abi.DIVA])
def test_friction_and_strain_heat_account_for_the_work_of_the_driving_stress(stress_balance):
    assert dissipation_mismatch(stress_balance) < 1e-8


# ---- till effective pressure (test/test_till_effective_pressure.jl) ------------------------------------------------

def till_formula(t, P0, Hw):
    s = min(max(Hw / t.maximum_water_thickness, 0.0), 1.0)
    Nt = t.reference_effective_pressure * (t.overburden_fraction * P0 / t.reference_effective_pressure) ** s * \
        10 ** (t.reference_void_ratio / t.compressibility * (1 - s))
    return min(Nt, P0)


def test_dry_partly_wet_saturated_and_floating_till():
    from helpers import B, field_values
    g = grid(5, 1, Lx=5.0, Ly=1.0, topology=(B, P), Nsigma=2)
    m = O.OracleModel(g, abi.default_physics(masking=0), abi.default_options())
    t = abi.default_till()
    m.set("H", np.full((1, 5), 1000.0))
    m.set("BED", field_values(g, "BED", lambda x, y: np.where(x < 4, 0.0, -2000.0)))
    m.set("HW", np.array([[0.0, 1.5, 2.0, 3.0, 1.0]]))
    m.till_effective_pressure(t)
    P0 = 918.0 * 9.81 * 1000.0
    N = m.get("N")[0]
    assert math.isclose(N[0], P0, rel_tol=1e-14)
    assert math.isclose(N[1], till_formula(t, P0, 1.5), rel_tol=1e-13) and N[1] < P0
    assert math.isclose(N[2], t.overburden_fraction * P0, rel_tol=1e-13)
    assert math.isclose(N[3], t.overburden_fraction * P0, rel_tol=1e-13)
    assert N[4] == 0
    assert N[0] > N[1] > N[2]


def test_ice_free_cells_have_zero_effective_pressure():
    from helpers import B
    g = grid(2, 1, Lx=2.0, Ly=1.0, topology=(B, P), Nsigma=2)
    m = O.OracleModel(g, abi.default_physics(masking=0), abi.default_options())
    m.set("H", np.array([[0.0, 100.0]]))
    m.set("HW", np.ones((1, 2)))
    m.till_effective_pressure(abi.default_till())
    N = m.get("N")[0]
    assert N[0] == 0 and N[1] > 0


def till_column_model(till, cls=O.OracleModel):
    g = grid(2, 1, Lx=2e3, Ly=1e3, topology=(P, P), Nsigma=10)
    phys = abi.default_physics(masking=0, ice_density=910.0, rate_factor_depth=abi.RATE_FACTOR_3D)
    m = cls(g, phys, abi.default_options())
    m.set("H", np.full((1, 2), 1000.0))
    m.set("HN", np.full((1, 2), 1000.0))
    m.set("TSURF", np.full((1, 2), 243.15))
    m.set("TOCEAN", np.full((1, 2), 271.15))
    m.set("ENTHALPY", np.full((10, 1, 2), CI * (243.15 - TR)))
    m.set("HW", np.array([[1.5, 3.0]]))
    return m, transport.Thermodynamics(grounded_melt_coupling=True, till=till)


def test_the_till_drains_and_is_capped_and_the_model_sees_the_till_pressure(cls=O.OracleModel):
    dt = 365.25 * 86400
    law = abi.default_till()
    ref, tr = till_column_model(None, cls)
    till, tt = till_column_model(law, cls)
    transport.time_step(ref, dt, thermodynamics=tr, first=True)
    transport.time_step(till, dt, thermodynamics=tt, first=True)
    Hwr, Hwt = ref.get("HW")[0], till.get("HW")[0]
    assert math.isclose(Hwt[0], Hwr[0] - dt * law.drainage_rate, rel_tol=1e-12)
    assert Hwt[1] == law.maximum_water_thickness
    assert Hwr[1] > law.maximum_water_thickness
    assert np.array_equal(till.get("GROUNDED_MELT"), ref.get("GROUNDED_MELT"))
    P0 = 910.0 * 9.81 * till.get("H")[0]
    N = till.get("N")[0]
    # the geometry, and so N, is diagnosed before the energy step updates the water
    assert np.allclose(N, [till_formula(law, P0[0], 1.5), till_formula(law, P0[1], 3.0)], rtol=1e-13)
    transport.update_state(till, False, tt)
    assert np.allclose(till.get("N")[0], [till_formula(law, P0[k], Hwt[k]) for k in range(2)], rtol=1e-13)
    assert np.allclose(ref.get("N")[0], P0, rtol=1e-14)
