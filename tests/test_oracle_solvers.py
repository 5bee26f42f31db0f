"""Pins the CPU oracle's operator, PCG, Picard and multigrid against the reference's own tests.

Ported from test/test_ssa_operator.jl, test/test_friction.jl, test/test_conjugate_gradient.jl,
test/test_glen_viscosity.jl, test/test_calving_front.jl, test/test_ice_mask.jl, test/test_model.jl,
test/test_diva.jl and test/test_multigrid_preconditioner.jl. The reference's stretched-grid cases run here on
uniform grids (the solver supports uniform RectilinearGrids); every other case keeps the reference's sizes,
tolerances and assertions.
"""
import math

import numpy as np
import pytest

import oracle as O
from helpers import B, P, YEAR, abi, apply_operator, field_values, grid, set_field, solved, vdot


def model(g, phys=None, **opts):
    return O.OracleModel(g, phys or abi.default_physics(), abi.default_options(**opts))


def ssa_coefficients(m, eta=lambda x, y: 1 + x ** 2 + np.sin(2 * np.pi * y) ** 2 / 2, beta=lambda x, y: 1 + x * y):
    set_field(m, "ETAH", eta)
    set_field(m, "ETAHFF", eta)
    set_field(m, "BETAU", beta)
    set_field(m, "BETAV", beta)


# ---- operator: doctest ssa_operator.jl:41-44, test/test_ssa_operator.jl:51-111 --------------------------------------

def test_operator_doctest():
    g = grid(4, 4, topology=(P, P))
    m = model(g)
    ssa_coefficients(m, 1.0, 2.0)
    qu, _ = apply_operator(m, 3.0, 0.0)
    assert qu[0, 0] == 0.375


@pytest.mark.parametrize("topology,walls", [((B, P), abi.FREE_SLIP), ((B, B), abi.FREE_SLIP),
                                            ((B, B), abi.NO_SLIP), ((P, P), abi.FREE_SLIP)])
def test_operator_symmetric_positive(topology, walls):
    g = grid(12, 10, topology=topology)
    m = model(g, abi.default_physics(walls=walls))
    ssa_coefficients(m)
    pu = field_values(g, "U", lambda x, y: np.sin(np.pi * x) * (np.cos(2 * np.pi * y) + 2))
    pv = field_values(g, "V", lambda x, y: np.cos(2 * x) + np.sin(4 * np.pi * y) * x)
    yu = field_values(g, "U", lambda x, y: x ** 2 - np.cos(2 * np.pi * y))
    yv = field_values(g, "V", lambda x, y: np.exp(-x) * np.sin(2 * np.pi * y))
    Apu, Apv = apply_operator(m, pu, pv)
    Ayu, Ayv = apply_operator(m, yu, yv)
    assert math.isclose(vdot(g, Apu, Apv, yu, yv), vdot(g, pu, pv, Ayu, Ayv), rel_tol=1e-12)
    assert vdot(g, Apu, Apv, pu, pv) > 0 and vdot(g, Ayu, Ayv, yu, yv) > 0


def test_operator_dense_assembly_is_symmetric_and_spd():
    # brute force: assemble A column by column from unit vectors on a tiny wall-bounded grid with an ice-free hole
    g = grid(5, 4, topology=(B, B))
    m = model(g, abi.default_physics(walls=abi.NO_SLIP))
    eta = lambda x, y: np.where((x - 0.7) ** 2 + (y - 0.4) ** 2 < 0.04, 0.0, 1 + x)  # noqa: E731
    set_field(m, "ETAH", eta)
    set_field(m, "ETAHFF", lambda x, y: 1 + x * y)
    set_field(m, "BETAU", lambda x, y: 1 + x)
    set_field(m, "BETAV", lambda x, y: 2 + y)
    nu, nv = (g.Ny, g.Nx), (g.Ny, g.Nx)
    n = 2 * g.Nx * g.Ny
    A = np.zeros((n, n))
    su, sv = m.shape("U"), m.shape("V")
    for col in range(n):
        pu, pv = np.zeros((su[1], su[0])), np.zeros((sv[1], sv[0]))
        if col < n // 2:
            pu[np.unravel_index(col, nu)] = 1
        else:
            pv[np.unravel_index(col - n // 2, nv)] = 1
        qu, qv = apply_operator(m, pu, pv)
        A[:, col] = np.concatenate([qu[: g.Ny, : g.Nx].ravel(), qv[: g.Ny, : g.Nx].ravel()])
    assert np.allclose(A, A.T, rtol=1e-13, atol=1e-13 * np.abs(A).max())
    assert np.all(np.linalg.eigvalsh((A + A.T) / 2) > 0)
    # the Jacobi diagonal preconditioner divides by exactly this diagonal (test/test_friction.jl:83-94)
    ones_u, ones_v = np.ones((su[1], su[0])), np.ones((sv[1], sv[0]))
    m.set("PU", ones_u)
    m.set("PV", ones_v)
    m.precondition()
    z = np.concatenate([m.get("QU")[: g.Ny, : g.Nx].ravel(), m.get("QV")[: g.Ny, : g.Nx].ravel()])
    d = np.diag(A)
    assert np.allclose(z * d, 1.0, rtol=1e-13)


def test_free_slip_translation_feels_only_drag():
    Nx, Ny, b0, U = 6, 5, 2.0, 3.0
    for topology in ((P, B), (B, P)):
        g = grid(Nx, Ny, topology=topology)
        m = model(g)
        ssa_coefficients(m, beta=b0)
        along_x = topology[0] == P
        pu = field_values(g, "U", U if along_x else 0.0)
        pv = field_values(g, "V", 0.0 if along_x else U)
        qu, qv = apply_operator(m, pu, pv)
        q = qu[: g.Ny, : g.Nx] if along_x else qv[: g.Ny, : g.Nx]
        if along_x:
            assert np.allclose(q, b0 * U / (Nx * Ny), rtol=1e-12)
        else:
            assert np.allclose(q[1:, :], b0 * U / (Nx * Ny), rtol=1e-12)


def test_ice_free_rows_are_identity():
    # test/test_ice_mask.jl:38-61
    Nx, Ny = 8, 6
    g = grid(Nx, Ny, topology=(B, P))
    m = model(g)
    ice = lambda x, y: np.where(x < 0.5, 1.0, 0.0)  # noqa: E731
    ssa_coefficients(m, eta=lambda x, y: ice(x, y) * (1 + x ** 2))
    pu = field_values(g, "U", lambda x, y: np.sin(np.pi * x) + y)
    pv = field_values(g, "V", lambda x, y: x * y ** 2 + 1)
    yu = field_values(g, "U", lambda x, y: np.cos(x) * y + 1)
    yv = field_values(g, "V", lambda x, y: np.sin(3 * y) + x)
    Apu, Apv = apply_operator(m, pu, pv)
    Ayu, Ayv = apply_operator(m, yu, yv)
    assert math.isclose(vdot(g, Apu, Apv, yu, yv), vdot(g, pu, pv, Ayu, Ayv), rel_tol=1e-12)
    mu = m.get("ETAH")
    u_free = np.array([[i == 0 or i == Nx or (mu[j, i - 1] == 0 and mu[j, i] == 0) for i in range(Nx + 1)] for j in range(Ny)])
    v_free = np.array([[mu[(j - 1) % Ny, i] == 0 and mu[j, i] == 0 for i in range(Nx)] for j in range(Ny)])
    assert u_free.sum() > 2 * Ny and v_free.sum() > 0
    u_free[:, Nx] = False
    assert np.all(Apu[u_free] == pu[u_free]) and np.all(Apv[v_free] == pv[v_free])


def manufactured_error(N):
    g = grid(N, N, topology=(P, P))
    m = model(g, linear=abi.convergence(1e-12, 0.0, 10_000), preconditioner=abi.PRECONDITIONER_NONE)
    e0, b0, a, b = 1.0, 1.0, 2 * np.pi, 4 * np.pi
    ssa_coefficients(m, e0, b0)
    A = 1 / N ** 2
    set_field(m, "TAUU", lambda x, y: A * (e0 * (4 * a ** 2 + b ** 2 + 3 * a * b) + b0) * np.sin(a * x) * np.cos(b * y))
    set_field(m, "TAUV", lambda x, y: A * (e0 * (4 * b ** 2 + a ** 2 + 3 * a * b) + b0) * np.cos(a * x) * np.sin(b * y))
    m.set("U", 0.0)
    m.set("V", 0.0)
    m.solve()
    eu = field_values(g, "U", lambda x, y: np.sin(a * x) * np.cos(b * y))
    ev = field_values(g, "V", lambda x, y: np.cos(a * x) * np.sin(b * y))
    return max(np.abs(m.get("U") - eu).max(), np.abs(m.get("V") - ev).max())


def test_manufactured_solution_second_order():
    e = [manufactured_error(N) for N in (32, 64)]
    assert e[1] < 5e-3 and e[0] / e[1] > 3.5


def test_cg_recovers_p():
    g = grid(12, 10, topology=(B, P))
    m = model(g, linear=abi.convergence(1e-12, 0.0, 10_000), preconditioner=abi.PRECONDITIONER_NONE)
    ssa_coefficients(m)
    pu = field_values(g, "U", lambda x, y: np.sin(np.pi * x) * (np.cos(2 * np.pi * y) + 2))
    pv = field_values(g, "V", lambda x, y: np.cos(2 * x) + np.sin(4 * np.pi * y) * x)
    Apu, Apv = apply_operator(m, pu, pv)
    m.set("TAUU", Apu)
    m.set("TAUV", Apv)
    m.set("U", 0.0)
    m.set("V", 0.0)
    m.solve()
    u, v = m.get("U"), m.get("V")
    assert np.allclose(u[:, 1:12], pu[:, 1:12], rtol=1e-8)
    assert np.allclose(v, pv, rtol=1e-8)
    assert np.all(u[:, [0, 12]] == 0)


def channel_error(N, along_x):
    W, e0, f = 1.0, 2.0, 3.0
    topology = (P, B) if along_x else (B, P)
    g = grid(4, N, topology=topology) if along_x else grid(N, 4, topology=topology)
    m = model(g, abi.default_physics(walls=abi.NO_SLIP), linear=abi.convergence(1e-13, 0.0, 10_000))
    ssa_coefficients(m, e0, 0.0)
    Az = g.dx * g.dy
    m.set("TAUU", f * Az if along_x else 0.0)
    m.set("TAUV", 0.0 if along_x else f * Az)
    m.set("U", 0.0)
    m.set("V", 0.0)
    m.solve()
    if along_x:
        U, Ue, across = m.get("U"), field_values(g, "U", lambda x, y: f * y * (W - y) / (2 * e0)), m.get("V")
    else:
        U, Ue, across = m.get("V"), field_values(g, "V", lambda x, y: f * x * (W - x) / (2 * e0)), m.get("U")
    return np.abs(U - Ue).max() / np.abs(Ue).max(), np.abs(across).max()


@pytest.mark.parametrize("along_x", [True, False])
def test_no_slip_channel(along_x):
    (e1, a1), (e2, a2) = channel_error(16, along_x), channel_error(32, along_x)
    assert e2 < 2e-3 and e1 / e2 > 3.5 and max(a1, a2) < 1e-12


def test_no_slip_stiffer_than_free_slip():
    g = grid(6, 5)
    pu = field_values(g, "U", lambda x, y: np.sin(np.pi * x) + y)
    pv = field_values(g, "V", lambda x, y: x * y ** 2)
    out = []
    for walls in (abi.NO_SLIP, abi.FREE_SLIP):
        m = model(g, abi.default_physics(walls=walls))
        ssa_coefficients(m)
        qu, qv = apply_operator(m, pu, pv)
        out.append(vdot(g, qu, qv, pu, pv))
    assert out[0] > out[1]


def test_jacobi_fewer_iterations_same_solution():
    # test/test_friction.jl:96-111
    g = grid(48, 32, topology=(B, P))
    b = dict(TAUU=lambda x, y: np.sin(np.pi * x) * np.cos(2 * np.pi * y), TAUV=lambda x, y: x * np.sin(2 * np.pi * y))
    res = []
    for pre in (abi.PRECONDITIONER_NONE, abi.PRECONDITIONER_DIAGONAL):
        m = model(g, linear=abi.convergence(1e-10, 0.0, 10_000), preconditioner=pre)
        ssa_coefficients(m, eta=lambda x, y: 1 + 1e3 * x ** 4, beta=lambda x, y: 1e4 * np.exp(-20 * x))
        for k, f in b.items():
            set_field(m, k, f)
        m.set("U", 0.0)
        m.set("V", 0.0)
        info = m.solve()
        res.append((m.get("U"), info["linear_iterations"]))
    assert res[1][1] < res[0][1]
    assert np.allclose(res[1][0], res[0][0], rtol=1e-7, atol=1e-7 * np.abs(res[0][0]).max())


# ---- Picard: test/test_glen_viscosity.jl:14-80 -------------------------------------------------------------------

def ice_stream_error(Nx, L=100e3, H0=500.0, delta=20.0, rho=917.0, gr=9.81, A=1e-24):
    g = grid(Nx, 4, Lx=L, Ly=L / Nx * 4, Nsigma=4)
    phys = abi.default_physics(masking=0, ice_density=rho, ocean_density=rho, gravitational_acceleration=gr)
    m = model(g, phys, linear=abi.convergence(1e-12, 0.0, 10_000), picard=abi.convergence(1e-10, 0.0, 300),
              preconditioner=abi.PRECONDITIONER_NONE)
    set_field(m, "H", H0)
    set_field(m, "S", lambda x, y: H0 + delta * np.cos(np.pi * x / L))
    m.set("A", A)
    m.compute_driving_stress()
    m.set("U", 0.0)
    m.set("V", 0.0)
    info = m.picard_solve()
    U = A * (rho * gr * delta / 2) ** 3 * L / np.pi
    ue = field_values(g, "U", lambda x, y: U * (np.sin(np.pi * x / L) - np.sin(np.pi * x / L) ** 3 / 3))
    return np.abs(m.get("U") - ue).max() / (2 * U / 3), info["picard_iterations"], np.abs(m.get("V")).max() / (2 * U / 3)


def test_ice_stream_picard():
    r = [ice_stream_error(N) for N in (32, 64)]
    assert all(x[1] < 300 for x in r) and all(x[2] < 1e-10 for x in r)
    assert r[1][0] < 1e-3 and r[0][0] / r[1][0] > 3.5


# ---- model-level solves: test/test_calving_front.jl:23-66, test/test_ice_mask.jl:63-79, test/test_model.jl -------

def confined_shelf(Nx=32, L=64e3, x_front=40e3, H0=500.0, A=1e-24):
    g = grid(Nx, 4, Lx=L, Ly=4 * L / Nx)
    phys = abi.default_physics(grounding=abi.GROUNDING_SUBGRID, friction=abi.FRICTION_WEERTMAN)
    m = model(g, phys, linear=abi.convergence(1e-12, 0.0, 10_000), picard=abi.convergence(1e-10, 0.0, 500))
    set_field(m, "H", lambda x, y: np.where(x < x_front, H0, 0.0))
    m.set("BED", -2000.0)
    m.set("C", 1e6)
    m.set("A", A)
    set_field(m, "U", lambda x, y: 1e-9 * x)
    m.set("V", 0.0)
    info = m.update_velocities()
    return g, m, info


def test_confined_shelf_spreading_rate():
    L, x_front, H0, A = 64e3, 40e3, 500.0, 1e-24
    g, m, info = confined_shelf()
    ri, rw, gr = 918.0, 1028.0, 9.81
    rp = ri * (1 - ri / rw)
    e = A * (rp * gr * H0 / 4) ** 3
    x = np.arange(g.Nx + 1) * g.dx
    u = m.get("U")[1]
    ice = x <= x_front
    assert np.abs(u[ice] - e * x[ice]).max() / (e * x_front) < 1e-6
    assert np.abs(m.get("V")).max() < 1e-12 * e * x_front


def test_steady_shelf_second_order():
    L, x_front, H0, A, n = 256e3, 192e3, 1e3, 1e-25, 3
    q = H0 * 1e-5
    errs = []
    for Nx in (64, 128):
        g = grid(Nx, 4, Lx=L, Ly=4 * L / Nx)
        phys = abi.default_physics(grounding=abi.GROUNDING_SUBGRID, friction=abi.FRICTION_WEERTMAN)
        m = model(g, phys, linear=abi.convergence(1e-12, 0.0, 20_000), picard=abi.convergence(1e-11, 0.0, 500))
        ri, rw, gr = 918.0, 1028.0, 9.81
        rp = ri * (1 - ri / rw)
        k = (n + 1) * (A / q) * (rp * gr / 4) ** n
        Hx = lambda x: (H0 ** -(n + 1) + k * x) ** (-1 / (n + 1))  # noqa: E731
        set_field(m, "H", lambda x, y: np.where(x < x_front, Hx(x), 0.0))
        m.set("BED", -2000.0)
        m.set("C", 1e6)
        m.set("A", A)
        set_field(m, "U", lambda x, y: 1e-9 * x)
        m.set("V", 0.0)
        m.update_velocities()
        x = np.arange(Nx + 1) * g.dx
        ice = x <= x_front
        u = m.get("U")[1][ice]
        ue = q / Hx(x[ice]) - q / H0
        errs.append(np.abs(u - ue).max() / ue.max())
    assert errs[1] < 5e-3 and errs[0] / errs[1] > 3


def ice_cap(Nx=20, L=200e3, R=60e3, H0=1000.0, H_edge=0.0, hmin=1.0, htr=1.0, pre=abi.PRECONDITIONER_DIAGONAL):
    g = grid(Nx, Nx, Lx=L, Ly=L)
    phys = abi.default_physics(friction=abi.FRICTION_WEERTMAN, mask_minimum_thickness=hmin, mask_transition_thickness=htr)
    m = model(g, phys, linear=abi.convergence(1e-10, 0.0, 10_000), picard=abi.convergence(1e-8, 0.0, 200),
              preconditioner=pre)

    def dome(x, y):
        r = np.hypot(x - L / 2, y - L / 2)
        return np.where(r < R, H0 * np.sqrt(np.maximum(1 - (r / R) ** 2, 0.0)) + H_edge, H_edge)

    set_field(m, "H", dome)
    m.set("BED", 0.0)
    m.set("C", 3e6)
    m.set("A", 1e-24)
    m.set("U", 0.0)
    m.set("V", 0.0)
    info = m.update_velocities()
    return g, m, info


@pytest.mark.parametrize("H_edge", [0.0, 0.5])
def test_ice_cap_with_ice_free_margins(H_edge):
    g, m, info = ice_cap(H_edge=H_edge)
    mu = m.get("MU")
    assert (mu == 0).sum() > 0
    assert 0 < info["picard_iterations"] < 200
    assert np.all(m.get("ETAH")[mu == 0] == 0)
    u, v = m.get("U"), m.get("V")
    Nx = g.Nx
    u_free = np.array([[i == 0 or i == Nx or (mu[j, i - 1] == 0 and mu[j, i] == 0) for i in range(Nx + 1)] for j in range(Nx)])
    v_free = np.array([[j == 0 or j == Nx or (mu[j - 1, i] == 0 and mu[j, i] == 0) for i in range(Nx)] for j in range(Nx + 1)])
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(v))
    assert np.all(u[u_free] == 0) and np.all(v[v_free] == 0)
    assert np.abs(u).max() > 0
    assert np.allclose(u, -u[:, ::-1], rtol=1e-6, atol=1e-6 * np.abs(u).max())


def model_ice_stream(Nx=64, L=100e3, H0=500.0, delta=20.0, A=1e-24, diva=False):
    g = grid(Nx, 4, Lx=L, Ly=L / Nx * 4, Nsigma=4)
    phys = abi.default_physics(ice_density=917.0, stress_balance=abi.DIVA if diva else abi.SSA)
    m = model(g, phys, linear=abi.convergence(1e-12, 0.0, 10_000), picard=abi.convergence(1e-10, 0.0, 300))
    m.set("H", H0)
    set_field(m, "BED", lambda x, y: delta * np.cos(np.pi * x / L))
    m.set("A", A)
    m.set("U", 0.0)
    m.set("V", 0.0)
    info = m.update_velocities()
    rg = 917.0 * 9.81
    U = A * (rg * delta / 2) ** 3 * L / np.pi
    ue = field_values(g, "U", lambda x, y: U * (np.sin(np.pi * x / L) - np.sin(np.pi * x / L) ** 3 / 3))
    return np.abs(m.get("U") - ue).max() / (2 * U / 3), m, info


def test_model_ice_stream_ssa_and_diva():
    err, m, info = model_ice_stream()
    assert err < 1e-3 and 0 < info["picard_iterations"] < 300
    assert np.allclose(m.get("S"), m.get("BED") + m.get("H"), rtol=1e-15)
    assert np.all(m.get("PHI") == 1) and np.all(m.get("PHIU") == 1)
    assert np.array_equal(m.get("UB"), m.get("U"))
    us = m.get("USIGMA")
    assert all(np.array_equal(us[k], m.get("U")) for k in range(us.shape[0]))
    err, m, info = model_ice_stream(diva=True)
    assert err < 1e-3
    assert np.array_equal(m.get("UB"), m.get("U"))
    us = m.get("USIGMA")
    assert all(np.allclose(us[k], m.get("U"), rtol=1e-12) for k in range(us.shape[0]))


def test_diva_faster_than_ssa_with_sliding():
    # test/test_model.jl:95-111
    g = grid(32, 4, Lx=1e5, Ly=1.25e4)
    out = {}
    for sb in (abi.SSA, abi.DIVA):
        phys = abi.default_physics(friction=abi.FRICTION_WEERTMAN, friction_exponent=1.0, stress_balance=sb)
        m = model(g, phys)
        set_field(m, "H", lambda x, y: 1000 + x / 1e2)
        m.set("BED", 0.0)
        m.set("C", 1e12)
        m.set("A", 1e-24)
        m.set("U", 0.0)
        m.set("V", 0.0)
        out[sb] = (m, m.update_velocities())
    diva, info = out[abi.DIVA]
    ub, ubar = diva.get("UB"), diva.get("U")
    us = diva.get("USIGMA")
    _, _, dz, _ = diva.sigma_metrics()
    assert 0 < info["picard_iterations"] < 100
    assert np.abs(ubar).max() > 2 * np.abs(out[abi.SSA][0].get("U")).max()
    assert np.all(np.abs(ub) <= np.abs(ubar))
    assert np.abs(np.tensordot(dz, us, axes=(0, 0)) - ubar).max() < 1e-2 * np.abs(ubar).max()


def slab_problem(make_model, Ns, H0=1000.0, tau0=1e4, A=1e-21, Cc=1e10):
    """slab_profile_error of test/test_diva.jl:5-48: a uniform slab on a 4 × 4 doubly periodic grid under a uniform
    driving stress with linear sliding, Picard to 1e-12, then compute_velocity_profile!. Returns the reference's
    error measures (iterations, profile, mean, basal, v) and the model."""
    g = grid(4, 4, Lx=4e3, Ly=4e3, topology=(P, P), Nsigma=Ns)
    phys = abi.default_physics(masking=0, strain_rate_regularization=1e-16, friction=abi.FRICTION_WEERTMAN,
                               friction_exponent=1.0, stress_balance=abi.DIVA)
    opts = abi.default_options(linear=abi.convergence(1e-12, 0.0, 100), picard=abi.convergence(1e-12, 0.0, 300),
                               preconditioner=abi.PRECONDITIONER_NONE)
    m = make_model(g, phys, opts)
    m.set("H", H0)
    m.set("PHIU", 1.0)
    m.set("PHIV", 1.0)
    m.set("TAUU", tau0 * 1e3 ** 2)
    m.set("TAUV", 0.0)
    m.set("A", A)
    m.set("C", Cc)
    m.set("U", 0.0)
    m.set("V", 0.0)
    info = m.picard_solve()
    m.compute_velocity_profile()
    n = 3.0
    ub = tau0 / Cc
    U = 2 * A * tau0 ** n * H0 / (n + 1)
    zc = (m.sigma_metrics()[1] if hasattr(m, "sigma_metrics") else None)
    if zc is None:
        zf = (np.arange(Ns + 1) / Ns) ** 2
        zc = (zf[:-1] + zf[1:]) / 2
    exact = ub + U * (1 - (1 - zc) ** (n + 1))
    us = m.get("USIGMA")
    ubar_e = ub + 2 * A * tau0 ** n * H0 / (n + 2)
    err = {"iterations": info["picard_iterations"],
           "profile": float(np.max(np.abs(us - exact[:, None, None])) / U),
           "mean": float(np.max(np.abs(m.get("U") - ubar_e)) / (ubar_e - ub)),
           "basal": float(np.max(np.abs(m.get("UB") - ub)) / ub),
           "v": float(np.max(np.abs(m.get("VSIGMA"))))}
    return err, m


def check_slab(coarse, fine):
    """The assertions of test/test_diva.jl:50-57."""
    assert 0 < fine["iterations"] < 300
    assert fine["profile"] < 1e-2
    assert fine["mean"] < 1e-2
    assert fine["basal"] < 1e-6
    assert fine["v"] == 0
    assert fine["profile"] < coarse["profile"] / 4


def test_diva_slab_profile():
    coarse, _ = slab_problem(O.OracleModel, 10)
    fine, _ = slab_problem(O.OracleModel, 40)
    check_slab(coarse, fine)


# ---- the reference's multigrid scaling benchmark (benchmarks/multigrid_scaling.jl, docs/src/model/numerics.md) --

@pytest.mark.parametrize("dx", [8000.0, 4000.0, 2000.0, 1000.0])
def test_mismip_scaling_iterations_equal_published(dx):
    """The published CG iteration counts of one solve from zero to 1e-8 on the MISMIP+ ice stream, after
    FixedIterations(5) Picard with the multigrid PCG at 1e-6 (docs/src/model/numerics.md:44-49): the diagonal
    and the multigrid preconditioner, iteration for iteration. The two solutions agree as the benchmark reports
    (max |Δu| / max |u| of the two solves, both at 1e-8)."""
    from nunatak_hip import workloads as W
    g, p, o, f = W.mismip_scaling(dx)
    m = O.OracleModel(g, p, o)
    W.load_into(m, f)
    r = W.mismip_timed_solves(m)
    assert (r["jacobi"][0], r["multigrid"][0]) == W.MISMIP_SCALING_ITERATIONS[dx]
    uj, um = r["jacobi"][1], r["multigrid"][1]
    assert np.max(np.abs(uj - um)) / np.max(np.abs(uj)) < 1e-6


# ---- multigrid: test/test_multigrid_preconditioner.jl ---------------------------------------------------------------

def mg_model(g, walls=abi.FREE_SLIP, smoother=abi.POINT_JACOBI):
    return model(g, abi.default_physics(walls=walls), preconditioner=abi.PRECONDITIONER_MULTIGRID, smoother=smoother)


def test_multigrid_level_shapes():
    assert mg_model(grid(32, 16, topology=(B, P), halo=3)).multigrid_levels() == [(16, 8), (8, 4), (4, 4)]
    assert mg_model(grid(15, 8, topology=(B, P), halo=3)).multigrid_levels() == [(15, 4)]
    assert mg_model(grid(64, 4, topology=(B, B), halo=4)).multigrid_levels() == [(32, 4), (16, 4), (8, 4), (4, 4)]
    assert mg_model(grid(15, 5, topology=(B, P), halo=3)).multigrid_levels() == []
    assert mg_model(grid(640, 80, halo=4)).multigrid_levels() == [(320, 40), (160, 20), (80, 10), (40, 5), (20, 5),
                                                                    (10, 5), (5, 5)]


def holed(m):
    g = m.grid
    eta = lambda x, y: np.where((x - 0.7) ** 2 + (y - 0.4) ** 2 < 0.04, 0.0, 1 + x)  # noqa: E731
    set_field(m, "ETAH", eta)
    e = field_values(g, "ETAH", eta)
    # η̄Hᶠᶠ = ℑxyᶠᶠ(η̄H) as _interpolate_to_corners! computes it; obtained through a coefficient-free corner pass:
    m.set("ETAHFF", corners_of(g, e))
    m.set("BETAU", 3.0)
    m.set("BETAV", 3.0)


def corners_of(g, e):
    """ℑxyᶠᶠ of a Center field with Oceananigans halo conventions (numpy restatement for test inputs only)."""
    Nx, Ny = g.Nx, g.Ny
    nx = Nx + (1 if g.topology_x == B else 0)
    ny = Ny + (1 if g.topology_y == B else 0)

    def at(i, j):  # 1-based, halo by topology
        i = min(max(i, 1), Nx) if g.topology_x == B else (i - 1) % Nx + 1
        j = min(max(j, 1), Ny) if g.topology_y == B else (j - 1) % Ny + 1
        return e[j - 1, i - 1]

    out = np.zeros((ny, nx))
    for j in range(1, ny + 1):
        for i in range(1, nx + 1):
            out[j - 1, i - 1] = ((at(i - 1, j - 1) + at(i, j - 1)) / 2 + (at(i - 1, j) + at(i, j)) / 2) / 2
    return out


def smooth_vector(g, mm, shape_u=None, shape_v=None):
    u = field_values(g, "U", lambda x, y: np.sin(2 * np.pi * mm * x + 1) * np.cos(2 * np.pi * y + mm))
    v = field_values(g, "V", lambda x, y: np.cos(2 * np.pi * x + mm) * np.sin(2 * np.pi * mm * y + 2))
    return u, v


TOPS = [(B, P), (B, B), (P, P)]
SIZES = [(16, 8), (32, 4), (4, 32)]


@pytest.mark.parametrize("topology", TOPS)
@pytest.mark.parametrize("size", SIZES)
def test_restriction_is_prolongation_transpose(topology, size):
    g = grid(*size, topology=topology, halo=4)
    m = mg_model(g)
    holed(m)
    levels = m.multigrid_levels()
    cNx, cNy = levels[0]
    cg = grid(cNx, cNy, topology=topology, halo=4)
    fu, fv = smooth_vector(g, 3)
    cu, cv = smooth_vector(cg, 2)
    su = (cNy, cNx + (1 if topology[0] == B else 0))
    sv = (cNy + (1 if topology[1] == B else 0), cNx)
    Rfu, Rfv = m.mg_transfer(0, fu, fv, su, sv)
    Pcu, Pcv = m.mg_transfer(1, cu, cv, None, None)
    assert math.isclose(vdot(cg, Rfu, Rfv, cu, cv), vdot(g, fu, fv, Pcu, Pcv), rel_tol=1e-12)


@pytest.mark.parametrize("topology", TOPS)
@pytest.mark.parametrize("size", SIZES)
@pytest.mark.parametrize("smoother", [abi.POINT_JACOBI, abi.LINE_JACOBI])
@pytest.mark.parametrize("walls", [abi.FREE_SLIP, abi.NO_SLIP])
def test_v_cycle_symmetric_positive(topology, size, smoother, walls):
    g = grid(*size, topology=topology, halo=4)
    m = mg_model(g, walls, smoother)
    holed(m)
    r1u, r1v = smooth_vector(g, 1)
    r2u, r2v = smooth_vector(g, 2)
    z = []
    for ru, rv in ((r1u, r1v), (r2u, r2v)):
        m.set("PU", ru)
        m.set("PV", rv)
        m.precondition()
        z.append((m.get("QU"), m.get("QV")))
    assert math.isclose(vdot(g, *z[0], r2u, r2v), vdot(g, r1u, r1v, *z[1]), rel_tol=1e-10)
    assert vdot(g, *z[0], r1u, r1v) > 0


def test_multigrid_beats_jacobi_on_ice_cap():
    # test/test_multigrid_preconditioner.jl:66-82: same coefficients, solve from zero with both preconditioners
    g, m, _ = ice_cap(Nx=128, hmin=10.0, htr=400.0)
    fields = {k: m.get(k) for k in ("ETAH", "ETAHFF", "BETAU", "BETAV", "TAUU", "TAUV")}
    out = []
    for pre in (abi.PRECONDITIONER_DIAGONAL, abi.PRECONDITIONER_MULTIGRID):
        mm = model(g, abi.default_physics(), linear=abi.convergence(1e-10, 0.0, 20_000), preconditioner=pre)
        for k, v in fields.items():
            mm.set(k, v)
        mm.set("U", 0.0)
        mm.set("V", 0.0)
        info = mm.solve()
        out.append((mm.get("U"), mm.get("V"), info["linear_iterations"]))
    (uj, vj, ij), (um, vm, im) = out
    assert im < ij / 3
    assert np.allclose(solved(g, um, "U"), solved(g, uj, "U"), rtol=1e-7, atol=1e-7 * np.abs(uj).max())
    assert np.allclose(solved(g, vm, "V"), solved(g, vj, "V"), rtol=1e-7, atol=1e-7 * np.abs(vj).max())
