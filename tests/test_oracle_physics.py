"""Pins the CPU oracle's pointwise physics against the reference's own tests and doctests.

Each test names the reference test or doctest it ports. Values come from closed forms, the reference's stated
doctest outputs, or textbook limits; none restates the oracle's formula.
"""
import math

import numpy as np
import pytest

import oracle as O
from helpers import B, P, abi, grid, set_field, field_values


def model(g, **phys):
    return O.OracleModel(g, abi.default_physics(**phys), abi.default_options())


# ---- ice mask: src/Models/smooth_ice_mask.jl:32-39 doctest, test/test_ice_mask.jl:28-35 ----------------------------

def test_ice_mask_doctest_and_values():
    assert [O.ice_mask_value(H, 1, 2) for H in (0.5, 1, 2, 3, 10)] == [0.0, 0.0, 0.5, 1.0, 1.0]
    assert O.ice_mask_value(10.0, 10, 20) == 0 and O.ice_mask_value(30.0, 10, 20) == 1
    assert O.ice_mask_value(20.0, 10, 20) == 0.5
    h = 1e-6
    assert abs(O.ice_mask_value(10 + h, 10, 20) - O.ice_mask_value(10.0, 10, 20)) / h < 1e-6
    assert abs(O.ice_mask_value(30.0, 10, 20) - O.ice_mask_value(30 - h, 10, 20)) / h < 1e-6


# ---- grounded fractions: test/test_grounding.jl:14-48, doctests subgrid_grounded_fraction.jl:207-232,274-291,318-334 --

def test_grounded_fraction_doctest():
    g = grid(4, 1, Lx=4, Ly=1, topology=(B, P))
    m = model(g, grounding=abi.GROUNDING_SUBGRID, ice_density=1.0, ocean_density=1.0)
    set_field(m, "H", 100)
    set_field(m, "BED", lambda x, y: -100 * (x - 0.75))
    m.update_geometry()
    assert list(m.get("PHI")[0]) == [1.0, 0.75, 0.0, 0.0]


@pytest.mark.parametrize("x_GL", [3.3, 7.5, 10.9])
def test_grounded_fraction_sums_to_grounding_line(x_GL):
    g = grid(16, 2, Lx=16, Ly=2, topology=(B, P))
    m = model(g, grounding=abi.GROUNDING_SUBGRID)
    ri, rw = 918.0, 1028.0
    set_field(m, "H", 1000)
    set_field(m, "BED", lambda x, y: ri / rw * (50 * (x_GL - x) - 1000))
    m.update_geometry()
    phi, phiu = m.get("PHI"), m.get("PHIU")
    assert math.isclose(phi[0].sum(), x_GL, rel_tol=1e-12)
    assert np.all((phi >= 0) & (phi <= 1))
    i = math.floor(x_GL + 0.5) + 1
    assert abs(phiu[0, i - 1] - (x_GL + 0.5 - (i - 1))) < 1e-12


def test_grounded_fraction_all_grounded_or_floating():
    g = grid(16, 2, Lx=16, Ly=2, topology=(B, P))
    for bed, expected in ((-800, 1.0), (-1000, 0.0)):
        m = model(g, grounding=abi.GROUNDING_SUBGRID)
        set_field(m, "H", 1000)
        set_field(m, "BED", bed)
        m.update_geometry()
        for name in ("PHI", "PHIU", "PHIV"):
            assert np.all(m.get(name) == expected)


@pytest.mark.parametrize("c,cell,expected", [(9, (5, 5), 0.5), (9, (4, 5), 1.0), (9.3, (5, 5), 1 - 0.7 ** 2 / 2)])
def test_triangle_fractions(c, cell, expected):
    g = grid(8, 8, Lx=8, Ly=8)
    m = model(g, grounding=abi.GROUNDING_SUBGRID, ice_density=1.0, ocean_density=1.0)
    set_field(m, "H", 0)
    set_field(m, "BED", lambda x, y: c - x - y)
    m.update_geometry()
    i, j = cell
    assert math.isclose(m.get("PHI")[j - 1, i - 1], expected, rel_tol=1e-12)


def test_effective_pressure_and_surface():
    # test/test_grounding.jl:41-48
    g = grid(16, 2, Lx=16, Ly=2, topology=(B, P))
    m = model(g)
    set_field(m, "H", 1000)
    set_field(m, "BED", lambda x, y: np.where(x < 4, 100.0, np.where(x < 8, -500.0, -1000.0)))
    m.update_geometry()
    N = m.get("N")[0]
    ri, rw, gr = 918.0, 1028.0, 9.81
    assert np.allclose(N[:4], ri * gr * 1000, rtol=1e-14)
    assert np.allclose(N[4:8], ri * gr * 1000 - rw * gr * 500, rtol=1e-14)
    assert np.all(N[8:] == 0)
    # doctest subgrid_grounded_fraction.jl:149-166: N = [10, 5]
    g2 = grid(2, 1, Lx=2, Ly=1, topology=(B, P))
    m2 = model(g2, ice_density=1.0, ocean_density=1.0, gravitational_acceleration=1.0)
    set_field(m2, "H", 10)
    set_field(m2, "BED", lambda x, y: np.where(x < 1, 5.0, -5.0))
    m2.update_geometry()
    assert list(m2.get("N")[0]) == [10.0, 5.0]
    # doctest subgrid_grounded_fraction.jl:192-209: s = [80, 50]
    m3 = model(g2, grounding=abi.GROUNDING_SUBGRID, ice_density=1.0, ocean_density=2.0)
    set_field(m3, "H", 100)
    set_field(m3, "BED", lambda x, y: np.where(x < 1, -20.0, -500.0))
    m3.update_geometry()
    assert list(m3.get("S")[0]) == [80.0, 50.0]


# ---- driving stress: doctests driving_stress.jl:28-40, test/test_glen_viscosity.jl:57-72 ---------------------------

def test_driving_stress_doctests():
    g = grid(4, 4)
    m = model(g, masking=0, ice_density=1.0, ocean_density=1.0, gravitational_acceleration=1.0)
    set_field(m, "H", 1)
    set_field(m, "S", lambda x, y: x)
    m.compute_driving_stress()
    assert m.get("TAUU")[0, 1] == -0.0625
    m2 = model(g, masking=0, ice_density=1.0, ocean_density=2.0, gravitational_acceleration=1.0,
               grounding=abi.GROUNDING_SUBGRID_DRIVING_STRESS)
    set_field(m2, "BED", lambda x, y: -2 * x)
    set_field(m2, "H", lambda x, y: 2 - 2 * x)
    set_field(m2, "S", lambda x, y: np.maximum(2 - 4 * x, 1 - x))
    m2.compute_driving_stress()
    assert m2.get("TAUU")[0, 1] == 0.33463541666666663


def test_driving_stress_values():
    g = grid(6, 5, Lx=3, Ly=1)
    m = model(g, masking=0, ice_density=2.0, ocean_density=2.0, gravitational_acceleration=5.0)
    set_field(m, "H", lambda x, y: 1 + x)
    set_field(m, "S", lambda x, y: 2 * x + 3 * y)
    m.compute_driving_stress()
    Az = 0.5 * 0.2
    tu, tv = m.get("TAUU"), m.get("TAUV")
    xf = np.arange(0.5, 2.51, 0.5)
    xc = np.arange(0.25, 2.76, 0.5)
    assert np.allclose(tu[:, 1:6], np.tile(-Az * 10 * (1 + xf) * 2, (5, 1)), rtol=1e-12)
    assert np.allclose(tv[1:5, :], np.tile(-Az * 10 * (1 + xc) * 3, (4, 1)), rtol=1e-12)
    assert np.all(tu[:, [0, 6]] == 0) and np.all(tv[[0, 5], :] == 0)


@pytest.mark.parametrize("bed,D", [(-2000, 918 * 500 / 1028), (-300, 300), (100, 0)])
def test_front_force(bed, D):
    # test/test_calving_front.jl:4-21
    g = grid(6, 4, Lx=6e3, Ly=4e3)
    ri, rw, gr, H0 = 918.0, 1028.0, 9.81, 500.0
    m = model(g)
    set_field(m, "H", lambda x, y: np.where(x < 3e3, H0, 0.0))
    m.set("MU", field_values(g, "MU", lambda x, y: np.where(x < 3e3, 1.0, 0.0)))
    set_field(m, "S", lambda x, y: np.where(x < 3e3, max(bed + H0, H0 * (1 - ri / rw)), 0.0))
    m.compute_driving_stress()
    tu = m.get("TAUU")
    assert np.allclose(tu[:, 3], 1e3 * gr * (ri * H0 ** 2 - rw * D ** 2) / 2, rtol=1e-12)
    assert np.all(tu[:, 4:] == 0)
    cu, _ = m.classes()
    assert np.all(cu[:, 3] & abi.CLASS_FRONT) and np.all(cu[:, 5] & abi.CLASS_ICE_FREE)
    assert np.all(cu[:, 0] & abi.CLASS_WALL) and np.all(cu[:, 6] & abi.CLASS_WALL)


@pytest.mark.parametrize("Nx,Ny", [(16, 2), (2, 16)])
def test_grounding_line_split(Nx, Ny):
    # test/test_grounding.jl:69-101
    ri, rw, gr, x_GL = 918.0, 1028.0, 9.81, 7.3e3
    g = grid(Nx, Ny, Lx=Nx * 1e3, Ly=Ny * 1e3)
    along = (lambda x, y: x) if Nx > Ny else (lambda x, y: y)
    mk = lambda gs: model(g, masking=0, grounding=gs)  # noqa: E731
    m0, m = mk(abi.GROUNDING_SUBGRID), mk(abi.GROUNDING_SUBGRID_DRIVING_STRESS)
    for mm in (m0, m):
        set_field(mm, "H", lambda x, y: 1000 - 0.02 * along(x, y))
        set_field(mm, "BED", lambda x, y: ri / rw * (0.05 * (x_GL - along(x, y)) - 1000 + 0.02 * along(x, y)))
        mm.update_geometry()
        mm.compute_driving_stress()
    name = "TAUU" if Nx > Ny else "TAUV"
    face = (lambda a, n: a[0, n - 1]) if Nx > Ny else (lambda a, n: a[n - 1, 0])
    t0, t = m0.get(name), m.get(name)
    n = 8
    H1, H2 = 1000 - 0.02 * 6.5e3, 1000 - 0.02 * 7.5e3
    f1, f2 = 0.05 * (x_GL - 6.5e3), 0.05 * (x_GL - 7.5e3)
    th = f1 / (f1 - f2)
    Hg = H1 + th * (H2 - H1)
    s = m.get("S")
    s1 = face(s, n - 1)
    s2, sg = (1 - ri / rw) * H2, (1 - ri / rw) * Hg
    assert math.isclose(face(t, n), -1e3 * ri * gr * ((H1 + Hg) * (sg - s1) + (Hg + H2) * (s2 - sg)) / 2, rel_tol=1e-12)
    assert not math.isclose(face(t, n), face(t0, n), rel_tol=1e-8)
    others = [k for k in range(2, max(Nx, Ny) + 1) if k != n]
    assert [face(t, k) for k in others] == [face(t0, k) for k in others]


# ---- viscosity and column integrals: test/test_glen_viscosity.jl:3-55, glen_flow_law.jl:78-94, test_sigma_grid -----

def test_viscosity_doctest():
    g = grid(4, 4, topology=(P, P), Nsigma=3)
    m = model(g, glen_exponent=1.0, masking=0)
    m.set("A", 1.0)
    m.update_coefficients()
    assert m.get("ETASIGMA")[0, 0, 0] == 0.5


@pytest.mark.parametrize("topology,axis", [((B, P), "x"), ((P, B), "y")])
def test_uniform_strain_viscosity(topology, axis):
    a = 2.0
    Nx, Ny = (8, 4) if axis == "x" else (4, 8)
    g = grid(Nx, Ny, topology=topology, Nsigma=3)
    m = model(g, masking=0, strain_rate_regularization=0.0)
    m.set("A", 1.0)
    if axis == "x":
        set_field(m, "U", lambda x, y: a * x)
    else:
        set_field(m, "U", lambda x, y: a * y)
    m.update_coefficients()
    es = m.get("ETASIGMA")
    if axis == "x":
        assert np.allclose(es[:, :, 1:7], a ** (-2 / 3) / 2, rtol=1e-14)
    else:
        assert np.allclose(es[:, 1:7, :], (a / 2) ** (-2 / 3) / 2, rtol=1e-14)


def test_column_integrals():
    # test/test_sigma_grid.jl: η̄H = η H, F₂ = H / (3η); Σ layer thickness = 1
    g = grid(4, 3, topology=(P, B), Nsigma=7)
    m = model(g, masking=0, glen_exponent=1.0, strain_rate_regularization=0.0)
    eta = 3e13
    m.set("A", 1 / (2 * eta))   # n = 1: ησ = A^(-1) (ε̇²)^0 / 2 = η
    set_field(m, "H", lambda x, y: 1000 + 500 * x)
    m.update_coefficients()
    zf, zc, dz, w2 = m.sigma_metrics()
    assert math.isclose(dz.sum(), 1.0, rel_tol=1e-15)
    assert np.allclose(zf, (np.arange(8) / 7) ** 2, rtol=0, atol=0)
    H = m.get("H")
    assert np.allclose(m.get("ETASIGMA"), eta, rtol=1e-14)
    assert np.allclose(m.get("ETAH"), eta * H, rtol=1e-14)
    assert np.allclose(m.get("F2"), H / (3 * eta), rtol=1e-14)


def test_depth_varying_rate_factor_integrals():
    # test/test_rate_factor.jl:117-145
    g = grid(8, 8, Lx=1e5, Ly=1e5, topology=(P, P), Nsigma=8)
    u = lambda x, y: 1e-6 * np.sin(2 * np.pi * x / 1e5)  # noqa: E731
    A0 = 1e-24
    m0 = model(g, masking=0)
    m0.set("A", A0)
    set_field(m0, "U", u)
    set_field(m0, "H", 1000)
    m0.update_coefficients()
    m = model(g, masking=0, rate_factor_depth=abi.RATE_FACTOR_3D)
    zf, zc, dz, _ = m.sigma_metrics()
    m.set("A", (A0 * (1 + 7 * zc))[:, None, None] * np.ones((8, 8, 8)))
    set_field(m, "U", u)
    set_field(m, "H", 1000)
    m.update_coefficients()
    scale = (1 + 7 * zc) ** (-1 / 3)
    e0 = m0.get("ETASIGMA")[0, 2, 2]
    assert np.allclose(m.get("ETASIGMA")[:, 2, 2], e0 * scale, rtol=1e-13)
    assert math.isclose(m.get("ETAH")[2, 2], 1000 * e0 * np.sum(dz * scale), rel_tol=1e-13)
    weights = ((1 - zf[:-1]) ** 3 - (1 - zf[1:]) ** 3) / 3
    assert math.isclose(m.get("F2")[2, 2], 1000 * np.sum(weights / scale) / e0, rel_tol=1e-13)


# ---- friction: test/test_friction.jl:59-68, doctest frictions.jl:102-125 --------------------------------------------

def test_drag_doctest_and_limits():
    g = grid(4, 4, topology=(P, P))
    m = model(g, masking=0, friction=abi.FRICTION_WEERTMAN, friction_exponent=1 / 3, velocity_regularization=0.0)
    m.set("C", 2.0)
    m.set("U", 1.0)
    m.set("V", 0.0)
    m.set("PHIU", 1.0)
    m.set("PHIV", 1.0)
    m.set("A", 1.0)
    m.set("H", 1.0)
    m.update_coefficients()
    assert m.get("BETAU")[0, 1] == 2.0
    C, mm, U = 3e4, 1 / 3, 1e-5
    assert math.isclose(O.basal_drag(U, C, 0.0, abi.FRICTION_WEERTMAN, mm, 0.0), C * U ** (mm - 1), rel_tol=1e-14)
    a2, N0 = 0.5, 1e6
    slow = O.basal_drag(1e-12, C, N0, abi.FRICTION_REGULARIZED_COULOMB, mm, a2)
    assert math.isclose(slow / (C * 1e-12 ** (mm - 1)), 1, rel_tol=1e-3)
    fast = O.basal_drag(1e8, C, N0, abi.FRICTION_REGULARIZED_COULOMB, mm, a2)
    assert math.isclose(fast * 1e8 / (a2 * N0), 1, rel_tol=1e-3)


def test_drag_scaled_by_grounded_fraction_and_zero_at_walls():
    g = grid(4, 4)
    m = model(g, masking=0, friction=abi.FRICTION_WEERTMAN, velocity_regularization=0.0)
    C, mm, U = 3e4, 1 / 3, 1e-5
    for name, v in (("C", C), ("U", U), ("V", 0.0), ("PHIU", 0.5), ("PHIV", 0.5), ("A", 1.0), ("H", 1.0)):
        m.set(name, v)
    m.update_coefficients()
    bu = m.get("BETAU")
    assert np.allclose(bu[:, 1:4], C * U ** (mm - 1) / 2, rtol=1e-14)
    assert np.all(bu[:, [0, 4]] == 0)


# ---- Arrhenius: test/test_rate_factor.jl:66-89, doctest arrhenius_rate_factor.jl:96-99 ------------------------------

def _branch(T, prefactor, Q):
    return prefactor * math.exp(-Q / (8.31441 * T))


def test_arrhenius_branches():
    cold = lambda T: _branch(T, 3.61e-13, 6.0e4)  # noqa: E731
    warm = lambda T: _branch(T, 1.73e3, 13.9e4)  # noqa: E731
    for T in (223.15, 243.15, 262.6):
        assert math.isclose(O.arrhenius_rate_factor(T, 0.0), cold(T), rel_tol=1e-13)
    for T in (263.7, 268.15, 273.15):
        assert math.isclose(O.arrhenius_rate_factor(T, 0.0), warm(T), rel_tol=1e-13)
    assert math.isclose(O.arrhenius_rate_factor(273.15, 0.005), warm(273.15) * (1 + 181.25 * 0.005), rel_tol=1e-13)
    assert math.isclose(O.arrhenius_rate_factor(273.15, 0.05), warm(273.15) * (1 + 181.25 * 0.01), rel_tol=1e-13)
    assert math.isclose(O.arrhenius_rate_factor(263.15, 0.0), math.sqrt(cold(263.15) * warm(263.15)), rel_tol=1e-13)
    assert float(f"{O.arrhenius_rate_factor(253.15, 0.0):.2e}") == 1.5e-25
    T = np.linspace(262, 264, 2001)
    A = np.array([O.arrhenius_rate_factor(t, 0.0) for t in T])
    assert np.all(np.diff(A) > 0)
    slope = np.diff(np.log(A)) / (T[1] - T[0])
    assert np.max(np.abs(np.diff(slope))) * 2000 / 2 < 1


def test_rate_factor_from_cold_temperature_is_arrhenius():
    # enthalpy round trip: cold ice keeps T and ω = 0; T′ = T − Tₚₘ(p) + T₀
    g = grid(3, 2, Lx=3e3, Ly=2e3, Nsigma=4)
    m = model(g)
    set_field(m, "H", 1000)
    zf, zc, _, _ = m.sigma_metrics()
    T = np.empty((4, 2, 3))
    for k in range(4):
        T[k] = 250.0 + 5 * zc[k]
    m.set("TEMPERATURE", T)
    m.set_physics(abi.default_physics(rate_factor_depth=abi.RATE_FACTOR_3D))
    m.rate_factor_from_temperature()
    A = m.get("A")
    for k in range(4):
        p = 918 * 9.81 * 1000 * (1 - zc[k])
        Tp = T[k, 0, 0] - (273.15 - 7.9e-8 * p) + 273.15
        assert math.isclose(A[k, 0, 0], O.arrhenius_rate_factor(Tp, 0.0), rel_tol=1e-14)
        assert Tp > T[k, 0, 0]


# ---- update_thermal_diagnostics!: doctest enthalpy_conversions.jl:20-33, column_energy_solve.jl:85-117 ------------

def test_enthalpy_conversions_doctest_and_temperate_branch():
    # doctest: E = enthalpy(263.15, 0, 1e6) → (temperature, water_fraction) = (263.15, 0.0)
    E, T, w = O.enthalpy_conversions(263.15, 0.0, 1e6)
    assert (T, w) == (263.15, 0.0)
    assert math.isclose(E, 2009.0 * (263.15 - 223.15), rel_tol=1e-15)
    # temperate: T above Tₚₘ is capped, ω carries the latent heat; Tₚₘ = T₀ − β p
    p = 5e6
    Tpm = 273.15 - 7.9e-8 * p
    E, T, w = O.enthalpy_conversions(280.0, 0.03, p)
    assert T == Tpm
    assert math.isclose(w, 0.03, rel_tol=1e-12)
    assert math.isclose(E, 2009.0 * (Tpm - 223.15) + 0.03 * 3.34e5, rel_tol=1e-15)
    # cold ice round trip keeps T (to rounding) and ω = 0
    for T0 in (230.0, 250.0, 265.0):
        E, T, w = O.enthalpy_conversions(T0, 0.0, p)
        assert math.isclose(T, T0, rel_tol=1e-15) and w == 0.0


def test_thermal_diagnostics_from_enthalpy():
    """T, ω and A of every layer from an enthalpy field mixing cold and temperate ice: pointwise the conversions and
    arrhenius_rate_factor(T − Tₚₘ + T₀, ω); for cold ice the same A as the temperature path (a cold round trip)."""
    rng = np.random.default_rng(3)
    Ns = 5
    g = grid(4, 3, Lx=4e3, Ly=3e3, Nsigma=Ns)
    m = model(g, masking=0, rate_factor_depth=abi.RATE_FACTOR_3D)
    H = 500 + 1500 * rng.random((3, 4))
    m.set("H", H)
    _, zc, _, _ = m.sigma_metrics()
    Ecold = 2009.0 * (230.0 + 40 * rng.random((Ns, 3, 4)) - 223.15)
    temperate = rng.random((Ns, 3, 4)) < 0.4
    E = np.where(temperate, 2009.0 * (273.15 - 223.15) + 0.02 * 3.34e5 * rng.random((Ns, 3, 4)), Ecold)
    m.set("ENTHALPY", E)
    m.thermal_diagnostics()
    T, W, A = m.get("TEMPERATURE"), m.get("WATER_FRACTION"), m.get("A")
    for k in range(Ns):
        for j in range(3):
            for i in range(4):
                p = 918 * 9.81 * H[j, i] * (1 - zc[k])
                Tpm = 273.15 - 7.9e-8 * p
                assert T[k, j, i] == min(Tpm, 223.15 + E[k, j, i] / 2009.0)
                assert W[k, j, i] == max(0.0, E[k, j, i] - 2009.0 * (Tpm - 223.15)) / 3.34e5
                assert A[k, j, i] == O.arrhenius_rate_factor(T[k, j, i] - Tpm + 273.15, W[k, j, i])
    assert np.any(W > 0) and np.any(W == 0)
    cold_T = T.copy()
    m2 = model(g, masking=0, rate_factor_depth=abi.RATE_FACTOR_3D)
    m2.set("H", H)
    m2.set("TEMPERATURE", np.where(temperate, 250.0, cold_T))
    m2.rate_factor_from_temperature()
    A2 = m2.get("A")
    assert np.allclose(A2[~temperate], A[~temperate], rtol=1e-13)
