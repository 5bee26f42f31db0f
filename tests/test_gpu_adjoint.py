"""Fixed-point adjoint of the converged SSA Picard solve (nk_adjoint) against central finite differences.

Mirrors test/test_enzyme.jl:352-410 of the reference: the ridge model (16 × 4, Bounded × Periodic, Glen A = 1e-24,
Weertman C = 3e6(1 + x/L), J = Σu² + Σv² over the solved cells), ∂J/∂C against a centred difference with
h = 1e-4·C (bar 1e-5 relative), the multigrid and diagonal preconditioners giving the same gradient (bar 1e-8
relative), and the warning of an unconverged adjoint. Added: ∂J/∂A (2D and per σ layer), ∂J/∂τ (= λ), the regularized
Coulomb drag, and the friction-inversion workload's ∂J/∂log C along seeded random directions.

No oracle adjoint exists: the gradient is pinned by finite differences of the forward solve, which is itself pinned
to the oracle by the parity tests.
"""
import sys

import numpy as np
import pytest

from helpers import B, P, abi, field_values, grid
from nunatak_hip import workloads as W
from nunatak_hip.model import HIPModel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


L, LY = 64e3, 16e3


def ridge(preconditioner=abi.PRECONDITIONER_DIAGONAL, friction=abi.FRICTION_WEERTMAN, Nsigma=1, A3d=False,
          adjoint_maxiter=None):
    """ridge_model of test/test_enzyme.jl:92-101 (bed 0: grounded everywhere)."""
    g = grid(16, 4, L, LY, topology=(B, P), halo=4, Nsigma=Nsigma)
    kw = {}
    if friction == abi.FRICTION_REGULARIZED_COULOMB:
        kw = dict(coulomb_coefficient=0.5, velocity_regularization=300.0 / W.YEAR)
    phys = abi.default_physics(friction=friction, friction_exponent=1.0 / 3.0,
                               rate_factor_depth=abi.RATE_FACTOR_3D if A3d else abi.RATE_FACTOR_2D, **kw)
    opts = abi.default_options(linear=abi.convergence(1e-14, 0.0, 10_000), picard=abi.convergence(1e-12, 0.0, 300),
                               preconditioner=preconditioner)
    m = HIPModel(g, phys, opts)
    H = field_values(g, "H", lambda x, y: 500 + 300 * np.exp(-((x - L / 2) / 12e3) ** 2) * (1 + np.sin(2 * np.pi * y / LY) / 5))
    m.set("H", H)
    m.set("BED", 0.0)
    if A3d:
        A = 1e-24 * (1 + 0.3 * np.arange(Nsigma)[:, None, None] / Nsigma) * np.ones((Nsigma, 4, 16))
        m.set("A", A)
    else:
        m.set("A", 1e-24)
    m.set("C", field_values(g, "C", lambda x, y: 3e6 * (1 + x / L)))
    return m, g


def cost(m, g):
    """sum_of_squares(u) + sum_of_squares(v) over the solved cells, after a solve from u = v = 0."""
    m.set("U", 0.0)
    m.set("V", 0.0)
    m.update_velocities(geometry=True)
    u, v = m.get("U"), m.get("V")
    return float(np.sum(u[: g.Ny, : g.Nx] ** 2) + np.sum(v[: g.Ny, : g.Nx] ** 2))


def adjoint_of_cost(m, g, control=None):
    J = cost(m, g)
    u, v = m.get("U"), m.get("V")
    ub, vb = np.zeros_like(u), np.zeros_like(v)
    ub[: g.Ny, : g.Nx] = 2 * u[: g.Ny, : g.Nx]
    vb[: g.Ny, : g.Nx] = 2 * v[: g.Ny, : g.Nx]
    m.set("UBAR", ub)
    m.set("VBAR", vb)
    info = m.adjoint(control)
    return J, info


def centred_difference(m, g, name, index, h):
    a = m.get(name)
    a0 = a.copy()
    a[index] += h
    m.set(name, a)
    Jp = cost(m, g)
    a[index] -= 2 * h
    m.set(name, a)
    Jm = cost(m, g)
    m.set(name, a0)
    return (Jp - Jm) / (2 * h)


@pytest.mark.parametrize("preconditioner", [abi.PRECONDITIONER_DIAGONAL, abi.PRECONDITIONER_MULTIGRID])
def test_friction_and_rate_factor_gradients_match_finite_differences(preconditioner):
    say(f"[adjoint] ridge 16x4, preconditioner {preconditioner}")
    m, g = ridge(preconditioner)
    J, info = adjoint_of_cost(m, g)
    say(f"[adjoint] J = {J:.6e}, adjoint iterations {info['picard_iterations']}, linear {info['total_linear_iterations']}")
    assert info["picard_iterations"] >= 1
    Cb, Ab = m.get("CBAR").copy(), m.get("ABAR").copy()
    C, A = m.get("C"), m.get("A")
    for name, grad, field, index in (("C", Cb, C, (1, 6)), ("C", Cb, C, (3, 11)), ("C", Cb, C, (0, 2)),
                                     ("A", Ab, A, (1, 2)), ("A", Ab, A, (2, 9))):
        fd = centred_difference(m, g, name, index, 1e-4 * field[index])
        ad = grad[index]
        say(f"[adjoint] dJ/d{name}{index}: adjoint {ad:.10e}, fd {fd:.10e}")
        assert abs(ad - fd) < 1e-5 * abs(fd)


def test_mixed_galerkin_adjoint_matches_fp64():
    """The adjoint's linear solves through the mixed form (FP64 PCG, FP32 Galerkin V-cycle): same gradient as the FP64
    diagonal-preconditioned adjoint. Bar: 1e-8 relative (max norm)."""
    m, g = ridge(abi.PRECONDITIONER_DIAGONAL)
    adjoint_of_cost(m, g)
    ref = m.get("CBAR").copy()
    m2, g2 = ridge(abi.PRECONDITIONER_MULTIGRID)
    o = m2.options
    o.precision, o.coarse_operator = abi.PRECISION_MIXED, abi.COARSE_GALERKIN
    m2 = HIPModel(g2, m2.physics, o)
    H = field_values(g2, "H", lambda x, y: 500 + 300 * np.exp(-((x - L / 2) / 12e3) ** 2) * (1 + np.sin(2 * np.pi * y / LY) / 5))
    m2.set("H", H)
    m2.set("BED", 0.0)
    m2.set("A", 1e-24)
    m2.set("C", field_values(g2, "C", lambda x, y: 3e6 * (1 + x / L)))
    _, info = adjoint_of_cost(m2, g2)
    err = np.max(np.abs(m2.get("CBAR") - ref)) / np.max(np.abs(ref))
    per_solve = info["total_linear_iterations"] / max(info["picard_iterations"], 1)
    say(f"[adjoint] mixed Galerkin vs FP64 diagonal: {err:.3e}, {per_solve:.1f} CG iterations per adjoint solve")
    assert err < 1e-8
    # the warm-started adjoint solves reach residuals near 1e-27; the FP32 V-cycle works on a power-of-two rescaled
    # residual, so they converge like any other solve instead of stalling in FP32 denormals
    assert per_solve < 100


def test_multigrid_and_diagonal_adjoints_agree():
    grads = []
    for pre in (abi.PRECONDITIONER_DIAGONAL, abi.PRECONDITIONER_MULTIGRID):
        m, g = ridge(pre)
        adjoint_of_cost(m, g)
        grads.append((m.get("CBAR").copy(), m.get("ABAR").copy(), m.get("LAMBDAU").copy()))
    for d, mg in zip(*grads):
        err = np.max(np.abs(mg - d)) / np.max(np.abs(d))
        say(f"[adjoint] multigrid vs diagonal: {err:.3e}")
        assert err < 1e-8


def test_driving_stress_gradient_is_lambda():
    """b̄ = λ: ∂J/∂τ through picard_solve! with τ held fixed."""
    m, g = ridge()
    m.update_geometry()
    m.compute_driving_stress()
    tu = m.get("TAUU").copy()

    def J():
        m.set("U", 0.0)
        m.set("V", 0.0)
        m.picard_solve()
        u, v = m.get("U"), m.get("V")
        return float(np.sum(u[: g.Ny, : g.Nx] ** 2) + np.sum(v[: g.Ny, : g.Nx] ** 2))

    J()
    u, v = m.get("U"), m.get("V")
    ub, vb = np.zeros_like(u), np.zeros_like(v)
    ub[: g.Ny, : g.Nx] = 2 * u[: g.Ny, : g.Nx]
    vb[: g.Ny, : g.Nx] = 2 * v[: g.Ny, : g.Nx]
    m.set("UBAR", ub)
    m.set("VBAR", vb)
    m.adjoint()
    lam = m.get("LAMBDAU").copy()
    for index in ((1, 5), (2, 10)):
        h = 1e-4 * abs(tu[index])
        t = tu.copy()
        t[index] += h
        m.set("TAUU", t)
        Jp = J()
        t[index] -= 2 * h
        m.set("TAUU", t)
        Jm = J()
        m.set("TAUU", tu)
        fd = (Jp - Jm) / (2 * h)
        say(f"[adjoint] dJ/dtau_u{index}: lambda {lam[index]:.10e}, fd {fd:.10e}")
        assert abs(lam[index] - fd) < 1e-5 * abs(fd)


def test_regularized_coulomb_gradient_matches_finite_differences():
    m, g = ridge(friction=abi.FRICTION_REGULARIZED_COULOMB)
    adjoint_of_cost(m, g)
    Cb = m.get("CBAR").copy()
    C = m.get("C")
    for index in ((1, 6), (2, 12)):
        fd = centred_difference(m, g, "C", index, 1e-4 * C[index])
        say(f"[adjoint] regularized Coulomb dJ/dC{index}: adjoint {Cb[index]:.10e}, fd {fd:.10e}")
        assert abs(Cb[index] - fd) < 1e-5 * abs(fd)


def test_layered_rate_factor_gradient_matches_finite_differences():
    m, g = ridge(Nsigma=4, A3d=True)
    adjoint_of_cost(m, g)
    Ab = m.get("ABAR").copy()
    A = m.get("A")
    assert Ab.shape == A.shape == (4, 4, 16)
    for index in ((0, 1, 6), (3, 2, 9)):
        fd = centred_difference(m, g, "A", index, 1e-4 * A[index])
        say(f"[adjoint] dJ/dA{index}: adjoint {Ab[index]:.10e}, fd {fd:.10e}")
        assert abs(Ab[index] - fd) < 1e-5 * abs(fd)


def test_unconverged_adjoint_warns(capfd):
    m, g = ridge()
    info = adjoint_of_cost(m, g, control=abi.convergence(1e-12, 0.0, 1))[1]
    assert info["picard_iterations"] == 1
    err = capfd.readouterr().err
    assert "Picard adjoint did not converge" in err


def diva_problem(picard=None):
    """diva_problem of test/test_enzyme.jl:63-90: 8 × 6 (Bounded, Periodic), Nσ = 6, τ given, φ = 1, no ice mask,
    A = 1e-21(1 + cos(2πy/Ly)/5), Weertman C = 1e6(1 + x/Lx), PCG 1e-14, Picard 1e-12 (or the given control)."""
    Lx, Ly, tau0 = 8e4, 6e4, 1e4
    g = grid(8, 6, Lx, Ly, topology=(B, P), halo=4, Nsigma=6)
    phys = abi.default_physics(stress_balance=abi.DIVA, masking=0, friction=abi.FRICTION_WEERTMAN,
                               strain_rate_regularization=1e-16)
    opts = abi.default_options(linear=abi.convergence(1e-14, 0.0, 10_000),
                               picard=picard if picard is not None else abi.convergence(1e-12, 0.0, 300))
    m = HIPModel(g, phys, opts)
    Az = (Lx / 8) * (Ly / 6)
    m.set("H", field_values(g, "H", lambda x, y: 1000 + x / 1e2 + 100 * np.sin(2 * np.pi * y / Ly)))
    m.set("PHIU", 1.0)
    m.set("PHIV", 1.0)
    m.set("TAUU", field_values(g, "TAUU", lambda x, y: tau0 * Az * np.sin(np.pi * x / Lx) * (1 + np.sin(2 * np.pi * y / Ly) / 5)))
    m.set("TAUV", field_values(g, "TAUV", lambda x, y: tau0 * Az * np.cos(np.pi * x / Lx) / 5))
    m.set("A", field_values(g, "A", lambda x, y: 1e-21 * (1 + np.cos(2 * np.pi * y / Ly) / 5)))
    m.set("C", field_values(g, "C", lambda x, y: 1e6 * (1 + x / Lx)))
    return m, g


def diva_cost(m, g):
    for name in ("U", "V", "ETASIGMA", "F2", "BETABU", "BETABV"):
        m.fill(name, 0.0)
    m.picard_solve()
    u, v = m.get("U"), m.get("V")
    return float(np.sum(u[: g.Ny, : g.Nx] ** 2) + np.sum(v[: g.Ny, : g.Nx] ** 2))


def test_diva_fixed_point_adjoint_matches_finite_differences():
    """Fixed-point adjoint of converged DIVA Picard iterations (test/test_enzyme.jl:319-350) with the cost on the
    depth-averaged velocities: ∂J/∂A and ∂J/∂C against centred differences (h = 1e-4 of the value), bar 1e-5."""
    m, g = diva_problem()
    J = diva_cost(m, g)
    u, v = m.get("U"), m.get("V")
    ub, vb = np.zeros_like(u), np.zeros_like(v)
    ub[: g.Ny, : g.Nx] = 2 * u[: g.Ny, : g.Nx]
    vb[: g.Ny, : g.Nx] = 2 * v[: g.Ny, : g.Nx]
    m.set("UBAR", ub)
    m.set("VBAR", vb)
    info = m.adjoint()
    say(f"[adjoint] DIVA J = {J:.6e}; adjoint iterations {info['picard_iterations']}")
    Cb, Ab, lam = m.get("CBAR").copy(), m.get("ABAR").copy(), m.get("LAMBDAU").copy()
    C, A, tu = m.get("C").copy(), m.get("A").copy(), m.get("TAUU").copy()

    def fd(name, arr, index, h):
        a = arr.copy()
        a[index] += h
        m.set(name, a)
        Jp = diva_cost(m, g)
        a[index] -= 2 * h
        m.set(name, a)
        Jm = diva_cost(m, g)
        m.set(name, arr)
        return (Jp - Jm) / (2 * h)

    for name, grad, arr, index in (("A", Ab, A, (1, 2)), ("A", Ab, A, (4, 5)), ("C", Cb, C, (2, 3)), ("C", Cb, C, (5, 6)),
                                   ("TAUU", lam, tu, (2, 4))):
        d = fd(name, arr, index, 1e-4 * abs(arr[index]))
        say(f"[adjoint] DIVA dJ/d{name}{index}: adjoint {grad[index]:.10e}, fd {d:.10e}")
        assert abs(grad[index] - d) < 1e-5 * abs(d)


def diva_layer_cost(m, g):
    """diva_velocity_cost of test/test_enzyme.jl:44-61: Picard from zero, then compute_velocity_profile!, and
    J = Σu² + Σv² + Σuσ² + Σvσ² over every stored face (wall faces hold zero)."""
    for name in ("U", "V", "ETASIGMA", "F2", "BETABU", "BETABV"):
        m.fill(name, 0.0)
    m.picard_solve()
    m.compute_velocity_profile()
    u, v, us, vs = m.get("U"), m.get("V"), m.get("USIGMA"), m.get("VSIGMA")
    return float(np.sum(u ** 2) + np.sum(v ** 2) + np.sum(us ** 2) + np.sum(vs ** 2)), (u, v, us, vs)


@pytest.mark.gpu
def test_diva_adjoint_with_layer_velocity_cost_matches_finite_differences():
    """Fixed-point adjoint of converged DIVA Picard iterations (test/test_enzyme.jl:307-326) with the reference's full cost,
    which also reads the layer velocities uσ, vσ of compute_velocity_profile!: their shadows (USIGMABAR, VSIGMABAR)
    enter x̄ and the state's shadows of every fixed-point step. ∂J/∂A and ∂J/∂C at the reference's four points against
    centred differences (h = 1e-4 of the value); bar 1e-6 relative, the reference's."""
    m, g = diva_problem()
    J, (u, v, us, vs) = diva_layer_cost(m, g)
    assert np.abs(us).max() > 1.05 * np.abs(u).max() or np.abs(vs).max() > 1.05 * np.abs(v).max()   # shear is resolved
    m.set("UBAR", 2 * u)
    m.set("VBAR", 2 * v)
    m.set("USIGMABAR", 2 * us)
    m.set("VSIGMABAR", 2 * vs)
    info = m.adjoint()
    say(f"[adjoint] DIVA layer cost J = {J:.6e}; adjoint iterations {info['picard_iterations']}")
    Cb, Ab = m.get("CBAR").copy(), m.get("ABAR").copy()
    C, A = m.get("C").copy(), m.get("A").copy()

    def fd(name, arr, index, h):
        a = arr.copy()
        a[index] += h
        m.set(name, a)
        Jp = diva_layer_cost(m, g)[0]
        a[index] -= 2 * h
        m.set(name, a)
        Jm = diva_layer_cost(m, g)[0]
        m.set(name, arr)
        return (Jp - Jm) / (2 * h)

    # Julia (i, j) 1-based → packed [j − 1, i − 1]: A (3, 2), (6, 5); C (4, 3), (7, 6)
    for name, grad, arr, index in (("A", Ab, A, (1, 2)), ("A", Ab, A, (4, 5)), ("C", Cb, C, (2, 3)), ("C", Cb, C, (5, 6))):
        d = fd(name, arr, index, 1e-4 * abs(arr[index]))
        say(f"[adjoint] DIVA layer cost dJ/d{name}{index}: adjoint {grad[index]:.10e}, fd {d:.10e}")
        assert abs(grad[index] - d) < 1e-6 * abs(d)
    # the layer term matters: dropping it from the adjoint changes the gradient well beyond the bar
    m.fill("USIGMABAR", 0.0)
    m.fill("VSIGMABAR", 0.0)
    diva_layer_cost(m, g)
    m.adjoint()
    assert abs(m.get("CBAR")[2, 3] - Cb[2, 3]) > 1e-3 * abs(Cb[2, 3])


@pytest.mark.parametrize("N,chunk", [(128, None), (512, 0), (512, 1), (512, 2), (512, 3)])
def test_friction_inversion_gradient_along_random_directions(N, chunk):
    """Friction-inversion workload: ∂J/∂log C · d against centred differences along 16 seeded smooth directions d
    (GRF, ℓ = 20 km, unit maximum), Richardson-extrapolated from steps h = 1e-3 and h/2 so the O(h²) truncation term
    cancels: fd = (4 D(h/2) − D(h)) / 3, D(h) = (J(C e^{hd}) − J(C e^{−hd})) / 2h. Bar: 1e-8 relative. At 512² the 16
    directions run as four cases of four (each draws all 16, so direction k is the same in every case)."""
    directions = 16
    g, phys, opts, fields, logC_true = W.friction_inversion(N=N, dx=512e3 / N)
    say(f"[adjoint] friction inversion {N}x{N}")
    m = HIPModel(g, phys, opts)
    W.load_into(m, fields)
    m.set("C", np.exp(logC_true))
    m.update_velocities(geometry=True)
    rng = np.random.Generator(np.random.PCG64(44))
    s_true, _, _ = W.centre_speed(m.get("U"), m.get("V"))
    s_obs = s_true + rng.normal(0.0, 5.0 / W.YEAR, size=s_true.shape)
    mask = fields["H"] > 0
    C0 = fields["C"]

    def J_of(C):
        m.set("C", C)
        m.set("U", 0.0)
        m.set("V", 0.0)
        info = m.update_velocities(geometry=True)
        return W.speed_misfit(m.get("U"), m.get("V"), s_obs, mask)[0], info

    J0, info = J_of(C0)
    _, gu, gv = W.speed_misfit(m.get("U"), m.get("V"), s_obs, mask)
    m.set("UBAR", gu)
    m.set("VBAR", gv)
    ainfo = m.adjoint()
    say(f"[adjoint] J0 = {J0:.6e}; forward picard {info['picard_iterations']}, adjoint iterations "
        f"{ainfo['picard_iterations']}, {ainfo['elapsed_ms']:.1f} ms")
    glogC = m.get("CBAR") * C0
    h = 1e-3

    def D(d, h):
        return (J_of(C0 * np.exp(h * d))[0] - J_of(C0 * np.exp(-h * d))[0]) / (2 * h)

    worst = 0.0
    dirs = [W.gaussian_random_field(rng, (N, N), (g.dx, g.dy), 20e3, 1.0) for _ in range(directions)]
    todo = range(directions) if chunk is None else range(4 * chunk, 4 * chunk + 4)
    for k in todo:
        d = dirs[k] / np.max(np.abs(dirs[k]))
        fd = (4 * D(d, h / 2) - D(d, h)) / 3
        ad = float(np.sum(glogC * d))
        rel = abs(ad - fd) / abs(fd)
        worst = max(worst, rel)
        say(f"[adjoint] direction {k}: adjoint {ad:.10e}, fd {fd:.10e}, rel {rel:.2e}")
    assert worst < 1e-8


# ---- geometry gradients: ∂J/∂H, ∂J/∂b -------------------------------------------------------------------------------

@pytest.mark.parametrize("preconditioner", [abi.PRECONDITIONER_DIAGONAL, abi.PRECONDITIONER_MULTIGRID])
def test_thickness_gradient_matches_finite_differences(preconditioner):
    """∂J/∂H₀ of the reference's ridge model through the converged SSA Picard solve with the geometry on the device
    (test/test_enzyme.jl:341-356: cells (6, 2) and (10, 3), centred differences with h = 1e-4 H, bar 1e-5)."""
    say(f"[adjoint] ridge 16x4 thickness gradient, preconditioner {preconditioner}")
    m, g = ridge(preconditioner)
    adjoint_of_cost(m, g)
    Hb = m.get("HBAR").copy()
    H = m.get("H")
    for index in ((1, 5), (2, 9)):   # the reference's cells, its step and bar
        fd = centred_difference(m, g, "H", index, 1e-4 * H[index])
        say(f"[adjoint] dJ/dH{index}: adjoint {Hb[index]:.10e}, fd {fd:.10e}")
        assert abs(Hb[index] - fd) < 1e-5 * abs(fd)
    # cells at the Bounded walls too, Richardson-extrapolated; bar 1e-7 of the gradient's largest entry (a cell whose
    # own gradient is 10³ smaller sits at the cost's rounding level in the differences)
    scale = np.max(np.abs(Hb))
    for index in ((0, 0), (3, 15), (1, 5)):
        h = 1e-3 * H[index]
        fd = (4 * centred_difference(m, g, "H", index, h / 2) - centred_difference(m, g, "H", index, h)) / 3
        say(f"[adjoint] dJ/dH{index} (Richardson): adjoint {Hb[index]:.10e}, fd {fd:.10e}")
        assert abs(Hb[index] - fd) < 1e-7 * scale


def test_thickness_gradient_multigrid_equals_diagonal():
    out = []
    for pre in (abi.PRECONDITIONER_DIAGONAL, abi.PRECONDITIONER_MULTIGRID):
        m, g = ridge(pre)
        adjoint_of_cost(m, g)
        out.append(m.get("HBAR").copy())
    rel = np.max(np.abs(out[1] - out[0])) / np.max(np.abs(out[0]))
    say(f"[adjoint] H̄ multigrid vs diagonal: {rel:.2e}")
    assert rel < 1e-8


def marine_geometry(N, grounding, friction, diva=False):
    """The friction-inversion geometry (grounded upstream, subgrid grounding line, floating shelf, calving front at
    the last column) at N × N, with the given grounding treatment and friction law."""
    g, phys, opts, fields, _ = W.friction_inversion(N=N, dx=512e3 / N)
    phys.grounding = grounding
    phys.friction = friction
    if friction == abi.FRICTION_REGULARIZED_COULOMB:
        phys.coulomb_coefficient, phys.velocity_regularization = 0.5, 300.0 / W.YEAR
        fields = dict(fields, C=np.full_like(fields["C"], 3e4))
    if diva:
        phys.stress_balance = abi.DIVA
    opts.picard = abi.convergence(1e-12, 0.0, 500)
    opts.linear = abi.convergence(1e-13, 0.0, 20_000)
    m = HIPModel(g, phys, opts)
    W.load_into(m, fields)
    return m, g, fields


def speed_cost(m, g, fields, H, B):
    m.set("H", H)
    m.set("BED", B)
    for name in ("U", "V", "ETASIGMA", "F2", "BETABU", "BETABV", "UB", "VB"):
        m.fill(name, 0.0)
    m.update_velocities(geometry=True)
    u, v = m.get("U"), m.get("V")
    return float(np.sum(u[: g.Ny, : g.Nx] ** 2) + np.sum(v[: g.Ny, : g.Nx] ** 2)) * W.YEAR ** 2


@pytest.mark.parametrize("case", ["subgrid-weertman", "gl-split-coulomb", "gl-split-weertman-diva"])
def test_geometry_gradients_along_random_directions(case):
    """∂J/∂H · d and ∂J/∂b · d on the marine geometry (masks, calving-front force, subgrid grounding line with and
    without the split driving stress, Weertman and regularized Coulomb drag with φ and N), SSA and DIVA, against
    Richardson-extrapolated centred differences along 3 seeded smooth directions each; d is zero on the open-ocean
    column so the ice mask does not move. Bar: 1e-6 relative."""
    grounding = abi.GROUNDING_SUBGRID if case.startswith("subgrid") else abi.GROUNDING_SUBGRID_DRIVING_STRESS
    friction = abi.FRICTION_REGULARIZED_COULOMB if "coulomb" in case else abi.FRICTION_WEERTMAN
    m, g, fields = marine_geometry(48, grounding, friction, diva=case.endswith("diva"))
    H0, B0 = fields["H"].copy(), fields["BED"].copy()
    say(f"[adjoint] marine geometry 48x48 {case}")
    J0 = speed_cost(m, g, fields, H0, B0)
    u, v = m.get("U"), m.get("V")
    ub, vb = np.zeros_like(u), np.zeros_like(v)
    ub[: g.Ny, : g.Nx] = 2 * u[: g.Ny, : g.Nx] * W.YEAR ** 2
    vb[: g.Ny, : g.Nx] = 2 * v[: g.Ny, : g.Nx] * W.YEAR ** 2
    m.set("UBAR", ub)
    m.set("VBAR", vb)
    info = m.adjoint()
    Hb, Bb = m.get("HBAR").copy(), m.get("BEDBAR").copy()
    say(f"[adjoint] J0 = {J0:.8e}, adjoint iterations {info['picard_iterations']}")
    rng = np.random.Generator(np.random.PCG64(61))
    ice = H0 > 0
    worst = 0.0
    for name, grad in (("H", Hb), ("BED", Bb)):
        for k in range(3):
            d = W.gaussian_random_field(rng, H0.shape, (g.dx, g.dy), 40e3, 1.0)
            d = np.where(ice, d / np.max(np.abs(d)), 0.0)

            def D(h):
                Hp, Bp = (H0 + h * d, B0) if name == "H" else (H0, B0 + h * d)
                Hm, Bm = (H0 - h * d, B0) if name == "H" else (H0, B0 - h * d)
                return (speed_cost(m, g, fields, Hp, Bp) - speed_cost(m, g, fields, Hm, Bm)) / (2 * h)

            h = 1e-2
            fd = (4 * D(h / 2) - D(h)) / 3
            ad = float(np.sum(grad * d))
            rel = abs(ad - fd) / abs(fd)
            worst = max(worst, rel)
            say(f"[adjoint] {case} d{name} direction {k}: adjoint {ad:.10e}, fd {fd:.10e}, rel {rel:.2e}")
    assert worst < 1e-6


# ---- reproducibility -------------------------------------------------------------------------------------------------

def test_adjoint_is_bit_reproducible():
    """The transposed coefficient kernels add into their neighbours in coloured passes (no atomics), so two adjoints
    of the same solve on fresh handles give the same gradients bit for bit and the same iteration counts: the ridge
    model (Bounded × Periodic with Ny = 4, so a Periodic direction has a point outside the stride-3 classes) through
    the mixed Galerkin form, and the marine geometry (50², DIVA, split driving stress) with its geometry gradients."""
    def ridge_mixed():
        m, g = ridge(abi.PRECONDITIONER_MULTIGRID)
        o = m.options
        o.precision, o.coarse_operator = abi.PRECISION_MIXED, abi.COARSE_GALERKIN
        m2 = HIPModel(g, m.physics, o)
        for name in ("H", "BED", "A", "C"):
            m2.set(name, m.get(name))
        _, info = adjoint_of_cost(m2, g)
        return m2, info

    def marine_diva():
        m, g, fields = marine_geometry(50, abi.GROUNDING_SUBGRID_DRIVING_STRESS, abi.FRICTION_WEERTMAN, diva=True)
        speed_cost(m, g, fields, fields["H"], fields["BED"])
        u, v = m.get("U"), m.get("V")
        m.set("UBAR", 2 * u * W.YEAR ** 2)
        m.set("VBAR", 2 * v * W.YEAR ** 2)
        return m, m.adjoint()

    for name, make in (("ridge mixed", ridge_mixed), ("marine DIVA 50x50", marine_diva)):
        runs = []
        for _ in range(2):
            m, info = make()
            runs.append(({k: m.get(k).copy() for k in ("CBAR", "ABAR", "HBAR", "BEDBAR", "LAMBDAU", "LAMBDAV")},
                         (info["picard_iterations"], info["total_linear_iterations"])))
        (a, ia), (b, ib) = runs
        say(f"[adjoint] {name}: iterations {ia} and {ib}")
        assert ia == ib
        for k in a:
            assert np.array_equal(a[k], b[k]), k
        assert np.max(np.abs(a["CBAR"])) > 0 and np.max(np.abs(a["HBAR"])) > 0


# ---- unrolled adjoint of FixedIterations Picard ------------------------------------------------------------------------

def ridge_fixed(n, preconditioner=abi.PRECONDITIONER_DIAGONAL):
    """The ridge model with Picard FixedIterations(n): nk_adjoint then differentiates the n unrolled steps."""
    m, g = ridge(preconditioner)
    o = m.options
    o.picard = abi.fixed_iterations(n)
    m2 = HIPModel(g, m.physics, o)
    for name in ("H", "BED", "A", "C"):
        m2.set(name, m.get(name))
    return m2, g


def close_rtol(a, b, rtol):
    """Julia's isapprox on arrays: ‖a − b‖ ≤ rtol · max(‖a‖, ‖b‖)."""
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(a), np.linalg.norm(b))


def test_fixed_point_adjoint_equals_unrolled_adjoint_of_60_iterations():
    """test/test_enzyme.jl:341-365: the fixed-point adjoint of the converged ridge solve (Convergence(reltol = 1e-12))
    against the taped adjoint of FixedIterations(60) — here the unrolled reverse of the 60 Picard steps — for ∂J/∂H₀
    and ∂J/∂C, rtol 1e-8 (the reference's bar)."""
    m, g = ridge(abi.PRECONDITIONER_DIAGONAL)
    adjoint_of_cost(m, g)
    fp = {k: m.get(k).copy() for k in ("HBAR", "CBAR", "LAMBDAU")}
    t, gt = ridge_fixed(60)
    _, info = adjoint_of_cost(t, gt)
    say(f"[adjoint] unrolled FixedIterations(60): {info['picard_iterations']} steps, {info['total_linear_iterations']} CG")
    assert info["picard_iterations"] == 60
    for k in ("HBAR", "CBAR", "LAMBDAU"):
        a, b = fp[k], t.get(k)
        say(f"[adjoint] fixed point vs unrolled {k}: {np.max(np.abs(a - b)) / np.max(np.abs(b)):.2e}")
        assert close_rtol(a, b, 1e-8), k


def test_unrolled_adjoint_of_three_iterations_matches_finite_differences():
    """Three Picard steps from zero are far from the fixed point, so their derivative differs from the fixed-point
    adjoint's: the unrolled reverse (each step linearized at its own input and multiplying its own result) against
    centred differences of J after FixedIterations(3), for C, A and H; bar 1e-5 (the reference's for the ridge)."""
    m, g = ridge_fixed(3)
    J, info = adjoint_of_cost(m, g)
    assert info["picard_iterations"] == 3
    Cb, Ab, Hb = m.get("CBAR").copy(), m.get("ABAR").copy(), m.get("HBAR").copy()
    C, A, H = m.get("C"), m.get("A"), m.get("H")
    for name, grad, field, index in (("C", Cb, C, (1, 6)), ("C", Cb, C, (3, 11)), ("A", Ab, A, (2, 9)),
                                     ("H", Hb, H, (1, 5)), ("H", Hb, H, (2, 9))):
        fd = centred_difference(m, g, name, index, 1e-4 * field[index])
        say(f"[adjoint] FixedIterations(3) dJ/d{name}{index}: unrolled {grad[index]:.10e}, fd {fd:.10e}")
        assert abs(grad[index] - fd) < 1e-5 * abs(fd)
    # the converged solve's gradient is a different quantity here
    mc, gc = ridge(abi.PRECONDITIONER_DIAGONAL)
    adjoint_of_cost(mc, gc)
    assert np.max(np.abs(mc.get("CBAR") - Cb)) > 1e-3 * np.max(np.abs(Cb))


def test_unrolled_adjoint_multigrid_equals_diagonal_and_restores_the_forward_state():
    """The unrolled reverse through the multigrid preconditioner gives the diagonal one's gradients (1e-8), and leaves
    the handle's (U, V) as the forward solve did."""
    out = []
    for pre in (abi.PRECONDITIONER_DIAGONAL, abi.PRECONDITIONER_MULTIGRID):
        m, g = ridge_fixed(8, pre)
        adjoint_of_cost(m, g)
        u_after = m.get("U").copy()
        cost(m, g)
        assert np.array_equal(u_after, m.get("U"))
        out.append((m.get("CBAR").copy(), m.get("HBAR").copy()))
    for d, mg in zip(*out):
        assert close_rtol(mg, d, 1e-8)


def diva_layer_adjoint(m, g):
    """diva_layer_cost and its adjoint; returns J and copies of (CBAR, ABAR, LAMBDAU)."""
    J, (u, v, us, vs) = diva_layer_cost(m, g)
    m.set("UBAR", 2 * u)
    m.set("VBAR", 2 * v)
    m.set("USIGMABAR", 2 * us)
    m.set("VSIGMABAR", 2 * vs)
    info = m.adjoint()
    return J, info, {k: m.get(k).copy() for k in ("CBAR", "ABAR", "LAMBDAU")}


def test_diva_fixed_point_adjoint_equals_unrolled_adjoint_of_60_iterations():
    """test/test_enzyme.jl:306-337: with the reference's full DIVA cost (u, v and the layer velocities), the fixed-point
    adjoint of Convergence(reltol = 1e-12) against the taped adjoint of FixedIterations(60) — the unrolled reverse
    carrying the lagged state (βᵇ, F₂, ησ) of each step — for ∂J/∂A and ∂J/∂C, rtol 1e-8 (the reference's bar)."""
    m, g = diva_problem()
    _, _, fp = diva_layer_adjoint(m, g)
    t, gt = diva_problem(abi.fixed_iterations(60))
    _, info, un = diva_layer_adjoint(t, gt)
    assert info["picard_iterations"] == 60
    for k in ("ABAR", "CBAR", "LAMBDAU"):
        say(f"[adjoint] DIVA fixed point vs unrolled {k}: {np.max(np.abs(fp[k] - un[k])) / np.max(np.abs(un[k])):.2e}")
    # the reference checks the parameter gradients; τ̄ = Σ λ_k of the 60 steps converges more slowly through the lagged
    # state (observed 5e-8), so it gets a looser bar
    for k in ("ABAR", "CBAR"):
        assert close_rtol(fp[k], un[k], 1e-8), k
    assert close_rtol(fp["LAMBDAU"], un["LAMBDAU"], 1e-6)


def test_diva_unrolled_adjoint_of_three_iterations_matches_finite_differences():
    """Three DIVA Picard steps from zero with the full cost: the unrolled reverse (lagged state from the tape, each step
    linearized at its input) against centred differences of J after FixedIterations(3) for A and C (h = 1e-4 of the
    value); bar 1e-5."""
    m, g = diva_problem(abi.fixed_iterations(3))
    J, info, gr = diva_layer_adjoint(m, g)
    assert info["picard_iterations"] == 3
    C, A = m.get("C").copy(), m.get("A").copy()

    def fd(name, arr, index, h):
        a = arr.copy()
        a[index] += h
        m.set(name, a)
        Jp = diva_layer_cost(m, g)[0]
        a[index] -= 2 * h
        m.set(name, a)
        Jm = diva_layer_cost(m, g)[0]
        m.set(name, arr)
        return (Jp - Jm) / (2 * h)

    for name, grad, arr, index in (("A", gr["ABAR"], A, (1, 2)), ("A", gr["ABAR"], A, (4, 5)), ("C", gr["CBAR"], C, (2, 3)),
                                   ("C", gr["CBAR"], C, (5, 6))):
        d = fd(name, arr, index, 1e-4 * abs(arr[index]))
        say(f"[adjoint] DIVA FixedIterations(3) dJ/d{name}{index}: unrolled {grad[index]:.10e}, fd {d:.10e}")
        assert abs(grad[index] - d) < 1e-5 * abs(d)
