"""The enthalpy step on the device (nk_compute_strain_heating, nk_energy_step) against the oracle: one step on a
random state with flow, thinning, melt, temperate layers and ice-free cells, bit for bit in the enthalpy, basal water
and grounded melt (the kernels do the oracle's arithmetic in the same order), A within 1e-13 (device exp/log); and
the reference's column tests (test/test_thermodynamics.jl) through the device."""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import B, P, abi, grid
from nunatak_hip import transport
from nunatak_hip.model import HIPModel
import test_oracle_energy as TE

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ulps(a, b):
    ia, ib = np.asarray(a, np.float64).view(np.int64), np.asarray(b, np.float64).view(np.int64)
    return int(np.max(np.abs(ia - ib)))


@pytest.mark.parametrize("stress_balance", [abi.SSA, abi.DIVA])
@pytest.mark.parametrize("grounding,coupled", [(abi.GROUNDING_NONE, False), (abi.GROUNDING_SUBGRID, True)])
@pytest.mark.parametrize("topology", [(B, P), (P, B)])
def test_energy_step_bit_exact(stress_balance, grounding, coupled, topology):
    say(f"energy step {stress_balance} grounding {grounding} coupled {coupled} {topology}")
    rng = np.random.default_rng(31)
    Nx, Ny, Ns = 23, 17, 8
    g = grid(Nx, Ny, Lx=23e3, Ly=17e3, topology=topology, Nsigma=Ns)
    phys = abi.default_physics(masking=0, grounding=grounding, stress_balance=stress_balance, walls=abi.NO_SLIP,
                               rate_factor_depth=abi.RATE_FACTOR_3D)
    h, o = HIPModel(g, phys, abi.default_options()), O.OracleModel(g, phys, abi.default_options())
    H = 100 + 2000 * rng.random((Ny, Nx))
    H[rng.random((Ny, Nx)) < 0.1] = 0.5   # ice-free columns
    p = (Ns, Ny, Nx)
    Epm = 2009.0 * (273.15 - 223.15)
    E = Epm - 2009.0 * 40 * rng.random(p)
    E[rng.random(p) < 0.25] = Epm + 0.03 * 3.34e5 * rng.random()
    f = {"H": H, "HN": H * (1 + 0.01 * rng.standard_normal((Ny, Nx))), "BED": -800 + 1200 * rng.random((Ny, Nx)),
         "ENTHALPY": E, "HW": np.where(rng.random((Ny, Nx)) < 0.5, 0.0, 1e-3 * rng.random((Ny, Nx))),
         "GEOTHERMAL": 0.05 + 0.05 * rng.random((Ny, Nx)), "TSURF": 240 + 30 * rng.random((Ny, Nx)),
         "TOCEAN": np.full((Ny, Nx), 271.15), "BASAL_MELT": 1e-9 * rng.random((Ny, Nx)),
         "GROUNDED_MELT": 1e-10 * rng.random((Ny, Nx)), "ETASIGMA": 1e13 * (1 + rng.random(p))}
    for name in ("U", "V", "UB", "VB", "BETAU", "BETAV", "BETABU", "BETABV"):
        s = h.shape(name)[1::-1]
        f[name] = (1e-5 * rng.standard_normal(s)) if name in ("U", "V", "UB", "VB") else 1e8 * rng.random(s)
    for name in ("USIGMA", "VSIGMA"):
        s = h.shape(name)[2::-1]
        f[name] = 1e-5 * rng.standard_normal(s)
    for m in (h, o):
        for k, v in f.items():
            m.set(k, v)
        m.update_geometry()
        m.compute_strain_heating()
        m.energy_step(3e8, coupled)
    a, b = h.get("STRAIN_HEATING"), o.get("STRAIN_HEATING")
    assert np.allclose(a, b, rtol=1e-13, atol=0)
    for name in ("ENTHALPY", "HW", "GROUNDED_MELT", "TEMPERATURE", "WATER_FRACTION", "HN"):
        assert ulps(h.get(name), o.get(name)) == 0, name
    assert np.allclose(h.get("A"), o.get("A"), rtol=1e-13, atol=0)
    assert o.get("WATER_FRACTION").max() > 0 and o.get("HW").max() > 0


def test_column_steady_conduction_on_the_device():
    say("energy: steady conduction column")
    H0, G, Ts = 1000.0, 0.042, 243.15
    m, g = TE.column_model(H0=H0, G=G, Ts=Ts, cls=HIPModel)
    TE.steps(m, 1e14, 4, first=True)
    z = np.array(O.OracleModel(g, abi.default_physics(ice_density=910.0), abi.default_options()).sigma_metrics()[1])
    exact = Ts + G * H0 * (1 - z) / TE.K
    assert np.max(np.abs(TE.col(m, "TEMPERATURE") - exact)) < 1e-8


def test_column_energy_budget_on_the_device():
    say("energy: budget column")
    H0, G = 800.0, 0.05
    m, g = TE.column_model(H0=H0, G=G, Ns=12, bed_refinement=2.0, cls=HIPModel)
    ref = O.OracleModel(g, abi.default_physics(ice_density=910.0), abi.default_options())
    z, zf = np.array(ref.sigma_metrics()[1]), np.array(ref.sigma_metrics()[0])
    TE.set_temperature(m, 250 - 5 * z)
    m.set("BETAU", np.full(m.shape("BETAU")[1::-1], 1e8))
    m.set("UB", np.full(m.shape("UB")[1::-1], 1e-6))
    F, dt = 1e8 * 1e-12, 1e9

    def energy():
        return 910.0 * H0 * float(np.sum(np.diff(zf) * TE.col(m, "ENTHALPY")))

    for k in range(20):
        e0 = energy()
        transport.time_step(m, dt, thermodynamics=TE.TH, first=(k == 0))
        q = TE.K / TE.CI * (TE.col(m, "ENTHALPY")[-1] - TE.CI * (243.15 - TE.TR)) / (H0 * (1 - z[-1]))
        assert abs(energy() - e0 - dt * (G + F - q)) < 1e-11 * abs(e0)


@pytest.mark.parametrize("stress_balance", [abi.SSA, abi.DIVA])
def test_friction_and_strain_heat_account_for_the_work_of_the_driving_stress_on_the_device(stress_balance):
    """test/test_thermodynamics.jl:392-398 through the device: velocity solves in every stage, strain heating and the
    energy step; frictional plus strain heat against the driving power, bar 1e-8."""
    say(f"energy: dissipation {stress_balance}")
    assert TE.dissipation_mismatch(stress_balance, cls=HIPModel) < 1e-8


def test_uniform_enthalpy_stays_uniform_on_the_device():
    say("energy: uniform enthalpy patch")
    m, g, z = TE.patch_model((P, P), cls=HIPModel)
    E0 = m.get("ENTHALPY")[0, 0, 0]
    H0 = m.get("H").copy()
    for k in range(20):
        transport.time_step(m, 1e9, thermodynamics=TE.TH, first=(k == 0))
    assert np.max(np.abs(m.get("H") - H0)) > 10
    assert np.max(np.abs(m.get("ENTHALPY") - E0)) < 1e-9 * E0


def test_till_pressure_and_drainage_on_the_device():
    """test/test_till_effective_pressure.jl:54-79 through the device, and nk_till_effective_pressure against the
    oracle on a random sheet (device pow: 1e-13)."""
    say("energy: till")
    TE.test_the_till_drains_and_is_capped_and_the_model_sees_the_till_pressure(cls=HIPModel)
    rng = np.random.default_rng(4)
    g = grid(19, 13, Lx=19e3, Ly=13e3, topology=(B, P), Nsigma=2)
    phys = abi.default_physics(masking=0)
    h, o = HIPModel(g, phys, abi.default_options()), O.OracleModel(g, phys, abi.default_options())
    f = {"H": 3000 * rng.random((13, 19)) * (rng.random((13, 19)) > 0.1), "BED": -1500 + 2000 * rng.random((13, 19)),
         "HW": 3 * rng.random((13, 19))}
    for m in (h, o):
        for k, v in f.items():
            m.set(k, v)
        m.till_effective_pressure(abi.default_till())
    assert np.allclose(h.get("N"), o.get("N"), rtol=1e-13, atol=0)
