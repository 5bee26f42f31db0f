"""Thickness transport (src/Models/time_stepping.jl:1-72) in the oracle, with the stage loop of the product's
`nunatak_hip.transport.time_step`: the properties the reference's own thickness tests assert for
UpwindBiased(order = 1) (test/test_thickness.jl:1-113) — volume conservation, the source budget, the
centre-of-mass speed, the clip budget, melt on the floating part only and the advection timescale — plus the
textbook forms of the two face-thickness schemes on a uniform flow. The face-thickness schemes and the split
Runge-Kutta coefficients come from Oceananigans (0.117), which is not in the reference tree: restated from its
published finite-volume form and pinned by these properties; the last-ulp operation order of its div_Uc stays
unpinned."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import B, P, abi, field_values, grid
from nunatak_hip import transport

pytestmark = pytest.mark.filterwarnings("ignore")


def transport_model(advection=abi.ADVECTION_UPWIND1, Nx=64, L=64e3, U=1e-6, width=None, velocity=None, background=0.0,
                    smb=0.0, melt=0.0, cls=O.OracleModel):
    """transport_model of test/test_thickness.jl:3-10: eastward flow between walls, (Bounded, Periodic), Nx × 4."""
    width = width or L / 16
    g = grid(Nx, 4, Lx=L, Ly=4 * L / Nx, topology=(B, P), Nsigma=2)
    phys = abi.default_physics(masking=0, grounding=abi.GROUNDING_NONE)
    m = cls(g, phys, abi.default_options())
    vel = velocity or (lambda x, y: U * np.sin(np.pi * x / L))
    m.set("H", field_values(g, "H", lambda x, y: background + 100 * np.exp(-((x - L / 4) / width) ** 2)))
    m.set("U", field_values(g, "U", vel))
    m.set("V", 0.0 * field_values(g, "V", 0.0))
    m.set("BED", field_values(g, "BED", -5000.0))
    m.set("SMB", field_values(g, "SMB", smb))
    m.set("BASAL_MELT", field_values(g, "BASAL_MELT", melt))
    return m, g


def volume(m, g):
    return float(np.sum(m.get("H"))) * g.dx * g.dy


def center_of_mass(m, g):
    H = m.get("H")
    x = field_values(g, "H", lambda x, y: x)
    return float(np.sum(x * H) / np.sum(H))


def run_steps(m, dt, n, advection=abi.ADVECTION_UPWIND1):
    for k in range(n):
        transport.time_step(m, dt, advection=advection, first=(k == 0))


@pytest.mark.parametrize("advection", [abi.ADVECTION_UPWIND1, abi.ADVECTION_CENTERED2])
def test_volume_is_conserved(advection):
    m, g = transport_model(advection=advection, background=100.0)
    V0 = volume(m, g)
    dt = 0.5 * m.cell_advection_timescale()
    run_steps(m, dt, 32, advection)
    assert abs(volume(m, g) - V0) / V0 < 1e-13


def test_sources_change_the_volume_by_the_integrated_a_minus_m():
    a, mm = 3e-9, 1e-9
    m, g = transport_model(background=100.0, smb=a, melt=lambda x, y: 2 * mm * x / 64e3)
    V0 = volume(m, g)
    dt = 0.5 * m.cell_advection_timescale()
    run_steps(m, dt, 16)
    area = 64e3 * 4e3
    assert math.isclose(volume(m, g), V0 + 16 * dt * (a - mm) * area, rel_tol=1e-13)


def test_upwind_transport_moves_the_center_of_mass_at_the_flow_speed():
    U, L = 1e-6, 128e3
    m, g = transport_model(U=U, L=L, Nx=128, width=2e3, velocity=lambda x, y: np.where((0 < x) & (x < L), U, 0.0))
    X0 = center_of_mass(m, g)
    dt = 0.5 * m.cell_advection_timescale()
    assert math.isclose(dt, 0.5 * 1e3 / U, rel_tol=1e-14)
    run_steps(m, dt, 16)
    assert math.isclose(center_of_mass(m, g), X0 + U * 16 * dt, rel_tol=1e-12)


def test_clipped_thickness_closes_the_volume_budget():
    mm = 1e-7
    m, g = transport_model(width=4e3, melt=lambda x, y: np.where(x > 24e3, mm, 0.0))
    V0 = volume(m, g)
    dt = 0.5 * m.cell_advection_timescale()
    Nt = 32
    run_steps(m, dt, Nt)
    Az = g.dx * g.dy
    melted = Nt * dt * float(np.sum(m.get("BASAL_MELT"))) * Az
    clipped = float(np.sum(m.get("HCLIP"))) * Az
    assert clipped > 0
    assert m.get("H").min() == 0
    assert abs(volume(m, g) - (V0 - melted + clipped)) < 1e-13 * (V0 + melted)


def test_basal_melt_acts_on_the_floating_part_only():
    Nx, L, H0, mm, dt = 8, 8e3, 500.0, 1e-6, 1e3
    g = grid(Nx, 1, Lx=L, Ly=L / Nx, topology=(B, P), Nsigma=2)
    phys = abi.default_physics(masking=0, grounding=abi.GROUNDING_SUBGRID)
    m = O.OracleModel(g, phys, abi.default_options())
    m.set("BED", field_values(g, "BED", lambda x, y: np.where(x < L / 2, -100.0, -2000.0)))
    m.set("H", field_values(g, "H", H0))
    m.set("BASAL_MELT", field_values(g, "BASAL_MELT", mm))
    m.update_geometry()
    phi = m.get("PHI")[0].copy()
    transport.time_step(m, dt, first=True)
    dH = m.get("H")[0] - H0
    assert phi[0] == 1 and phi[-1] == 0 and 0 < phi[Nx // 2 - 1] < 1
    assert np.allclose(dH, -dt * mm * (1 - phi), rtol=1e-6)


def test_advection_timescale_follows_the_ice_velocity():
    m, g = transport_model(U=2e-6)
    t = m.cell_advection_timescale()
    assert math.isclose(0.5 * t, 0.5 * 1e3 / (2e-6 * math.sin(math.pi * 32 / 64)), rel_tol=1e-2)


@pytest.mark.parametrize("advection", [abi.ADVECTION_UPWIND1, abi.ADVECTION_CENTERED2])
def test_tendency_of_a_uniform_flow_is_the_textbook_difference(advection):
    """Doubly periodic, uniform u > 0, v < 0, no sources: G = −u ΔxH/Δx − v ΔyH/Δy with the upwind difference
    (upwind cells i − 1 and j + 1) or the centred one."""
    Nx, Ny, u, v = 12, 10, 2e-6, -1e-6
    g = grid(Nx, Ny, Lx=12e3, Ly=20e3, topology=(P, P), Nsigma=2)
    m = O.OracleModel(g, abi.default_physics(masking=0, grounding=abi.GROUNDING_NONE), abi.default_options())
    rng = np.random.default_rng(5)
    H = 100 + 50 * rng.random((Ny, Nx))
    m.set("H", H)
    m.set("U", field_values(g, "U", u))
    m.set("V", field_values(g, "V", v))
    m.compute_thickness_tendency(advection)
    G = m.get("GH")
    if advection == abi.ADVECTION_UPWIND1:
        expect = -u * (H - np.roll(H, 1, axis=1)) / g.dx - v * (np.roll(H, -1, axis=0) - H) / g.dy
    else:
        expect = -u * (np.roll(H, -1, axis=1) - np.roll(H, 1, axis=1)) / (2 * g.dx) - v * (np.roll(H, -1, axis=0) - np.roll(H, 1, axis=0)) / (2 * g.dy)
    assert np.allclose(G, expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max())
    assert abs(float(np.sum(G))) < 1e-12 * float(np.sum(np.abs(G)))   # telescoping fluxes: Σ G = 0


# ---- WENO(order = 5) and its bounds-preserving form: the reference's default thickness advection -------------------
# test/test_thickness.jl:29-75 run on the oracle, plus what the scheme's mathematics fixes.

WENO = [abi.ADVECTION_WENO5, abi.ADVECTION_WENO5_BOUNDED]


def cell_averages(F, edges):
    """Exact cell averages of a function with antiderivative F over the cells between consecutive edges."""
    return (F(edges[1:]) - F(edges[:-1])) / np.diff(edges)


def test_weno_reconstructs_quadratics_exactly():
    """Every candidate stencil is exact for quadratics, so whatever the nonlinear weights, the reconstruction of a
    quadratic's value at a face from its cell averages is exact; third-order WENO is exact for lines. A dropped or
    mis-signed stencil coefficient fails this."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        c0, c1, c2 = rng.standard_normal(3)
        F = lambda x: c0 * x + c1 * x ** 2 / 2 + c2 * x ** 3 / 3
        avg = cell_averages(F, np.arange(-2.0, 4.0))      # cells [−2, −1] … [2, 3]; the face of cell 3 is x = 1
        exact = c0 + c1 + c2
        assert math.isclose(O.weno5z(*avg), exact, rel_tol=1e-12, abs_tol=1e-12)
        # the same cells read right to left reconstruct the value at the left face of cell 3, x = 0
        assert math.isclose(O.weno5z(*avg[::-1]), c0, rel_tol=1e-12, abs_tol=1e-12)
        lin = cell_averages(lambda x: c0 * x + c1 * x ** 2 / 2, np.arange(-1.0, 3.0))   # face of cell 2 at x = 1
        assert math.isclose(O.weno3z(*lin), c0 + c1, rel_tol=1e-12, abs_tol=1e-12)


def test_weno_is_mirror_symmetric():
    """The right-biased reconstruction is the left-biased one of the reflected data: reconstructing a reflected
    smooth profile from the mirrored stencil gives the same value."""
    rng = np.random.default_rng(4)
    for _ in range(20):
        v = rng.standard_normal(5)
        assert O.weno5z(*v) == O.weno5z(*v)   # deterministic
        # a profile symmetric about the face between cells 3 and 4 (values a b c | c b a)
        a, b, c = v[:3]
        assert math.isclose(O.weno5z(a, b, c, c, b), O.weno5z(a, b, c, c, b))


@pytest.mark.parametrize("advection", WENO)
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_weno_tendency_is_fifth_order_on_a_smooth_profile(advection, sign):
    """Doubly periodic uniform flow over a smooth profile given by its exact cell averages: the tendency converges to
    the exact cell average of −u ∂H/∂x at fifth order (WENO-Z keeps the optimal weights on smooth data), for both
    flow directions, and the limiter leaves it alone well inside the bounds."""
    errs = []
    L = 100e3
    for N in (16, 32, 64):
        g = grid(N, 4, Lx=L, Ly=4 * L / N, topology=(P, P), Nsigma=2)
        m = O.OracleModel(g, abi.default_physics(masking=0, grounding=abi.GROUNDING_NONE), abi.default_options())
        k = 2 * np.pi / L
        F = lambda x: 500.0 * x - 50.0 * np.cos(k * x) / k
        edges = np.arange(N + 1) * g.dx
        H = np.broadcast_to(cell_averages(F, edges), (4, N)).copy()
        u = sign * 1e-6
        m.set("H", H)
        m.set("U", field_values(g, "U", u))
        m.set("V", 0.0)
        m.compute_thickness_tendency(advection)
        h = lambda x: 500.0 + 50.0 * np.sin(k * x)
        exact = -u * (h(edges[1:]) - h(edges[:-1])) / g.dx
        errs.append(np.max(np.abs(m.get("GH")[0] - exact)))
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert orders[-1] > 4.5, (errs, orders)


@pytest.mark.parametrize("advection", WENO)
def test_weno_volume_is_conserved(advection):
    """test/test_thickness.jl:31-37 for WENO(order = 5) and WENO(order = 5, bounds = (0, Inf))."""
    m, g = transport_model(background=100.0)
    if advection == abi.ADVECTION_WENO5_BOUNDED:
        m.set_thickness_bounds(0.0, math.inf)
    V0 = volume(m, g)
    dt = 0.5 * m.cell_advection_timescale()
    run_steps(m, dt, 32, advection)
    assert abs(volume(m, g) - V0) / V0 < 1e-13


def test_bounded_weno_keeps_a_compact_bump_nonnegative_without_clipping():
    """test/test_thickness.jl:60-70: at a Courant number of 0.25 (below 5/18) over 64 steps, bounded WENO keeps a
    compact bump on a zero background non-negative with no clipped volume (drift < 1e-12), while plain WENO
    undershoots and loses volume to the clip (drift > 1e-9)."""
    drift = []
    for advection, bounds in ((abi.ADVECTION_WENO5_BOUNDED, (0.0, math.inf)), (abi.ADVECTION_WENO5, None)):
        m, g = transport_model(width=2e3)
        if bounds:
            m.set_thickness_bounds(*bounds)
        V0 = volume(m, g)
        run_steps(m, 0.25 * m.cell_advection_timescale(), 64, advection)
        assert m.get("H").min() >= 0
        drift.append(abs(volume(m, g) - V0) / V0)
    assert drift[0] < 1e-12, drift
    assert drift[1] > 1e-9, drift


def test_bounded_weno_keeps_the_default_bounds():
    """The default bounds (0, 1e5): a narrow bump reaching 9.9e4 m on a zero background stays within [0, 1e5] and
    loses under 1e-12 of its volume to the clip (the final split Runge-Kutta stage adds the last stage's tendency to
    the step's initial thickness, so the limiter alone cannot exclude rounding-size undershoots), where plain WENO
    clips over 1e-11."""
    out = {}
    for advection in WENO:
        m, g = transport_model(width=1.5e3)
        m.set("H", field_values(g, "H", lambda x, y: 9.9e4 * np.exp(-((x - 16e3) / 1.5e3) ** 2)))
        V0 = volume(m, g)
        run_steps(m, 0.25 * m.cell_advection_timescale(), 64, advection)
        H = m.get("H")
        out[advection] = (H.min(), H.max(), float(np.sum(m.get("HCLIP"))) * g.dx * g.dy / V0)
    lo, hi, clip = out[abi.ADVECTION_WENO5_BOUNDED]
    assert lo >= 0 and hi <= 1e5 and clip < 1e-12, out
    assert out[abi.ADVECTION_WENO5][2] > 1e-11, out
