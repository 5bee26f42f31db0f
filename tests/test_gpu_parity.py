"""GPU parity: libnunatak_hip.so (through its C ABI) against the CPU oracle on the same seeded inputs.

Bars (written in each test): bit-exact for the FP64 operator, masks, grounded fractions, effective pressure,
surface, driving stress, face classes, Jacobi diagonal and multigrid transfers/V-cycle (integer-exact results of
identical IEEE operation sequences); ≤ 1e-13 relative for coefficients that call pow (device libm vs glibc);
≤ 1e-9 relative L2 in (u, v) for complete solves (reduction order differs); mixed precision ≤ 1e-8 relative L2
against the FP64 GPU solve.
"""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import B, P, abi, field_values, grid, set_field
from nunatak_hip import workloads as W
from nunatak_hip.model import HIPModel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(g, phys, opts):
    return HIPModel(g, phys, opts), O.OracleModel(g, phys, opts)


def rel_l2(a, b):
    a, b = np.concatenate([x.ravel() for x in a]), np.concatenate([x.ravel() for x in b])
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def ulps(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    ia, ib = a.view(np.int64), b.view(np.int64)
    return int(np.max(np.abs(ia - ib))) if a.size else 0


def random_coefficients(rng, g, hole=True):
    x, y = field_values(g, "ETAH", lambda x, y: x), field_values(g, "ETAH", lambda x, y: y)
    Lx, Ly = g.Nx * g.dx, g.Ny * g.dy
    eta = 1e15 * (1 + rng.random(x.shape))
    if hole:
        eta[((x / Lx - 0.7) ** 2 + (y / Ly - 0.4) ** 2) < 0.02] = 0.0
    etaff_shape = (g.Ny + (g.topology_y == B), g.Nx + (g.topology_x == B))
    return {"ETAH": eta, "ETAHFF": 1e15 * (1 + rng.random(etaff_shape)),
            "BETAU": 1e9 * rng.random(field_values(g, "U", 0).shape),
            "BETAV": 1e9 * rng.random(field_values(g, "V", 0).shape)}


TOPS = [(B, B), (B, P), (P, B), (P, P)]


@pytest.mark.parametrize("topology", TOPS)
@pytest.mark.parametrize("walls", [abi.FREE_SLIP, abi.NO_SLIP])
@pytest.mark.parametrize("size", [(37, 29), (64, 48)])
def test_operator_bit_exact(topology, walls, size):
    say(f"operator topology={topology} walls={walls} size={size}")
    rng = np.random.default_rng(11)
    g = grid(*size, Lx=size[0] * 1e3, Ly=size[1] * 1e3, topology=topology)
    h, o = pair(g, abi.default_physics(walls=walls), abi.default_options())
    for k, v in random_coefficients(rng, g).items():
        h.set(k, v)
        o.set(k, v)
    pu, pv = rng.standard_normal(h.shape("PU")[1::-1]), rng.standard_normal(h.shape("PV")[1::-1])
    for m in (h, o):
        m.set("PU", pu)
        m.set("PV", pv)
        m.apply_operator()
    Ny, Nx = g.Ny, g.Nx
    for name in ("QU", "QV"):
        a, b = h.get(name)[:Ny, :Nx], o.get(name)[:Ny, :Nx]
        assert np.array_equal(a, b), f"{name}: max ulps {ulps(a, b)}"


def test_operator_bit_exact_1024():
    say("operator 1024x1024")
    rng = np.random.default_rng(12)
    g = grid(1024, 1024, Lx=512e3, Ly=512e3)
    h, o = pair(g, abi.default_physics(walls=abi.NO_SLIP), abi.default_options())
    for k, v in random_coefficients(rng, g).items():
        h.set(k, v)
        o.set(k, v)
    pu, pv = rng.standard_normal(h.shape("PU")[1::-1]), rng.standard_normal(h.shape("PV")[1::-1])
    for m in (h, o):
        m.set("PU", pu)
        m.set("PV", pv)
        m.apply_operator()
    assert np.array_equal(h.get("QU")[:1024, :1024], o.get("QU")[:1024, :1024])
    assert np.array_equal(h.get("QV")[:1024, :1024], o.get("QV")[:1024, :1024])


GEOMETRY_FIELDS = ["S", "PHI", "PHIU", "PHIV", "N", "MU", "TAUU", "TAUV"]


@pytest.mark.parametrize("case", ["channel", "marine", "random", "periodic"])
@pytest.mark.parametrize("grounding", [abi.GROUNDING_SUBGRID, abi.GROUNDING_SUBGRID_DRIVING_STRESS])
def test_geometry_masks_driving_stress_bit_exact(case, grounding):
    say(f"geometry case={case} grounding={grounding}")
    if case == "channel":
        g, phys, opts, f = W.floating_channel()
    elif case == "marine":
        g, phys, opts, f = W.marine_sheet(Nx=160, Ny=20, dx=4e3)
    else:
        rng = np.random.default_rng(5)
        topo = (B, B) if case == "random" else (P, P)
        g = grid(50, 40, Lx=50e3, Ly=40e3, topology=topo)
        phys, opts = abi.default_physics(mask_minimum_thickness=5.0, mask_transition_thickness=20.0), abi.default_options()
        H = np.maximum(0.0, 600 * rng.random((40, 50)) - 150)
        f = {"H": H, "BED": -400 + 600 * rng.random((40, 50)) - 300 * np.linspace(0, 1, 50)[None, :]}
    phys.grounding = grounding
    h, o = pair(g, phys, opts)
    for m in (h, o):
        for k, v in f.items():
            if k in ("H", "BED"):
                m.set(k, v)
        m.update_geometry()
        m.compute_driving_stress()
    for name in GEOMETRY_FIELDS:
        a, b = h.get(name), o.get(name)
        assert np.array_equal(a, b), f"{name}: max ulps {ulps(a, b)}"
    cu_h, cv_h = h.classes()
    cu_o, cv_o = o.classes()
    assert np.array_equal(cu_h, cu_o) and np.array_equal(cv_h, cv_o)
    if case == "marine":
        assert (cu_h & abi.CLASS_FRONT).any() and (cu_h & abi.CLASS_ICE_FREE).any()
        if grounding == abi.GROUNDING_SUBGRID_DRIVING_STRESS:
            assert (cu_h & abi.CLASS_GROUNDING).any()


@pytest.mark.parametrize("kind", ["ssa2d", "ssa3d", "diva"])
def test_coefficient_update(kind):
    say(f"coefficients {kind}")
    rng = np.random.default_rng(7)
    g = grid(48, 40, Lx=48e3, Ly=40e3, Nsigma=6)
    phys = abi.default_physics(friction=abi.FRICTION_REGULARIZED_COULOMB, coulomb_coefficient=0.5,
                               walls=abi.NO_SLIP, stress_balance=abi.DIVA if kind == "diva" else abi.SSA,
                               rate_factor_depth=abi.RATE_FACTOR_3D if kind != "ssa2d" else abi.RATE_FACTOR_2D)
    h, o = pair(g, phys, abi.default_options())
    H = 500 + 400 * rng.random((40, 48))
    A = 1e-24 * (1 + rng.random((6, 40, 48)) if kind != "ssa2d" else 1 + rng.random((40, 48)))
    u = 1e-5 * rng.standard_normal(h.shape("U")[1::-1])
    v = 1e-5 * rng.standard_normal(h.shape("V")[1::-1])
    for m in (h, o):
        m.set("H", H)
        m.set("BED", -200 + 100 * rng.random((40, 48)) * 0 + np.linspace(0, 400, 48)[None, :])
        m.set("A", A)
        m.set("C", 3e6)
        m.set("U", u)
        m.set("V", v)
        m.update_geometry()
    if kind == "diva":
        es0 = 1e13 * (1 + rng.random((6, 40, 48)))
        for m in (h, o):
            m.set("ETASIGMA", es0)
            m.set("F2", 1e-10 * rng.random((40, 48)) * 0 + 1e-10)
            m.set("BETABU", 1e8 * np.ones(h.shape("U")[1::-1]))
            m.set("BETABV", 1e8 * np.ones(h.shape("V")[1::-1]))
    for m in (h, o):
        m.update_coefficients()
    names = ["ETAH", "F2", "ETASIGMA", "ETAHFF", "BETAU", "BETAV"] + (["UB", "VB", "BETABU", "BETABV"] if kind == "diva" else [])
    for name in names:
        a, b = h.get(name), o.get(name)
        scale = np.abs(b).max()
        assert np.allclose(a, b, rtol=1e-13, atol=1e-13 * scale), f"{name}: max rel {np.max(np.abs(a - b)) / scale:.2e}"


@pytest.mark.parametrize("topology", [(B, B), (B, P), (P, P)])
@pytest.mark.parametrize("smoother", [abi.POINT_JACOBI, abi.LINE_JACOBI])
@pytest.mark.parametrize("precond", [abi.PRECONDITIONER_DIAGONAL, abi.PRECONDITIONER_MULTIGRID])
def test_preconditioner_bit_exact(topology, smoother, precond):
    say(f"precondition topology={topology} smoother={smoother} precond={precond}")
    rng = np.random.default_rng(9)
    g = grid(64, 32, Lx=64e3, Ly=32e3, topology=topology, halo=4)
    opts = abi.default_options(preconditioner=precond, smoother=smoother)
    h, o = pair(g, abi.default_physics(walls=abi.NO_SLIP), opts)
    c = random_coefficients(rng, g)
    # multigrid corners follow the reference's interpolation of η̄H; set η̄H and compute corners on both sides
    ru, rv = rng.standard_normal(h.shape("PU")[1::-1]), rng.standard_normal(h.shape("PV")[1::-1])
    for m in (h, o):
        for k, v in c.items():
            m.set(k, v)
        m.set("PU", ru)
        m.set("PV", rv)
        m.precondition()
    Ny, Nx = g.Ny, g.Nx
    for name in ("QU", "QV"):
        a, b = h.get(name)[:Ny, :Nx], o.get(name)[:Ny, :Nx]
        assert np.array_equal(a, b), f"{name}: max ulps {ulps(a, b)} rel {np.max(np.abs(a - b)) / np.abs(b).max():.2e}"


def run_pair(g, phys, opts, fields):
    h, o = pair(g, phys, opts)
    for m in (h, o):
        W.load_into(m, fields)
    ih = h.update_velocities()
    io = o.update_velocities()
    return h, o, ih, io


def test_floating_channel_65x32_parity():
    say("floating channel 65x32 SSA Jacobi")
    h, o, ih, io = run_pair(*W.floating_channel())
    err = rel_l2([h.get("U"), h.get("V")], [o.get("U"), o.get("V")])
    say(f"  picard {ih['picard_iterations']} vs {io['picard_iterations']}, cg {ih['total_linear_iterations']} vs "
        f"{io['total_linear_iterations']}, rel L2 {err:.2e}")
    assert ih["picard_iterations"] == io["picard_iterations"]
    assert err <= 1e-9
    for name in ("MU", "PHI", "PHIU", "PHIV", "S", "N", "TAUU", "TAUV"):
        assert np.array_equal(h.get(name), o.get(name)), name
    assert np.array_equal(h.classes()[0], o.classes()[0]) and np.array_equal(h.classes()[1], o.classes()[1])


@pytest.mark.parametrize("precond", [abi.PRECONDITIONER_DIAGONAL, abi.PRECONDITIONER_MULTIGRID])
def test_marine_sheet_160x20_parity(precond):
    say(f"marine sheet 160x20 at 4 km, precond={precond}")
    g, phys, opts, f = W.marine_sheet(Nx=160, Ny=20, dx=4e3, preconditioner=precond)
    h, o, ih, io = run_pair(g, phys, opts, f)
    err = rel_l2([h.get("U"), h.get("V")], [o.get("U"), o.get("V")])
    say(f"  picard {ih['picard_iterations']} vs {io['picard_iterations']}, cg {ih['total_linear_iterations']} vs "
        f"{io['total_linear_iterations']}, rel L2 {err:.2e}")
    assert ih["picard_iterations"] == io["picard_iterations"]
    assert err <= 1e-9


def test_diva_floating_channel_64x64_parity():
    say("DIVA floating channel 64x64, Arrhenius A, MG")
    g, phys, opts, f = W.diva_shelf(N=64, dx=8e3)
    h, o, ih, io = run_pair(g, phys, opts, f)
    assert np.allclose(h.get("A"), o.get("A"), rtol=1e-14, atol=0)
    err = rel_l2([h.get("U"), h.get("V")], [o.get("U"), o.get("V")])
    say(f"  picard {ih['picard_iterations']} vs {io['picard_iterations']}, rel L2 {err:.2e}")
    assert ih["picard_iterations"] == io["picard_iterations"]
    assert err <= 1e-9
    errs = rel_l2([h.get("USIGMA")], [o.get("USIGMA")])
    assert errs <= 1e-9


def test_mixed_precision_matches_fp64():
    say("mixed vs FP64 on the DIVA shelf at 128x128")
    out = {}
    for prec in (abi.PRECISION_FP64, abi.PRECISION_MIXED):
        g, phys, opts, f = W.diva_shelf(N=128, dx=4e3, precision=prec)
        m = HIPModel(g, phys, opts)
        W.load_into(m, f)
        info = m.update_velocities()
        out[prec] = (m.get("U"), m.get("V"), info)
    err = rel_l2(out[1][:2], out[0][:2])
    say(f"  fp64 picard {out[0][2]['picard_iterations']} cg {out[0][2]['total_linear_iterations']}; mixed picard "
        f"{out[1][2]['picard_iterations']} inner {out[1][2]['total_linear_iterations']}; rel L2 {err:.2e}")
    assert err <= 1e-8


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64, abi.PRECISION_MIXED])
def test_graph_replays_match_eager_launches(precision):
    # PCG chunks replayed from a captured hipGraph must give the eager result bit for bit, solve after solve
    say(f"graph vs eager precision={precision}")
    res = {}
    for graphs in (0, 1):
        g, phys, opts, f = W.diva_shelf(N=96, dx=512e3 / 96, precision=precision)
        opts.graphs = graphs
        opts.check_interval = 3
        m = HIPModel(g, phys, opts)
        W.load_into(m, f)
        runs = []
        for rep in range(3):
            for name in ("U", "V", "ETASIGMA", "F2", "BETABU", "BETABV"):
                m.fill(name, 0.0)
            info = m.update_velocities()
            runs.append((m.get("U"), m.get("V"), info["total_linear_iterations"]))
            m.synchronize()
        res[graphs] = runs
    for k in range(3):
        assert res[0][k][2] == res[1][k][2]
        assert np.array_equal(res[0][k][0], res[1][k][0]) and np.array_equal(res[0][k][1], res[1][k][1])
        assert np.array_equal(res[1][k][0], res[1][0][0])


def test_thermal_diagnostics_parity():
    """nk_thermal_diagnostics against the oracle on an enthalpy field with cold and temperate layers: T and ω bit for
    bit (arithmetic, min and max only), A within 1e-13 (device exp/log)."""
    say("thermal diagnostics")
    rng = np.random.default_rng(11)
    Ns = 10
    g = grid(64, 48, Lx=64e3, Ly=48e3, Nsigma=Ns)
    phys = abi.default_physics(rate_factor_depth=abi.RATE_FACTOR_3D)
    h, o = pair(g, phys, abi.default_options())
    H = 3000 * rng.random((48, 64))
    E = 2009.0 * (220.0 + 55 * rng.random((Ns, 48, 64)) - 223.15)
    E[rng.random(E.shape) < 0.3] += 0.03 * 3.34e5 * rng.random()
    for m in (h, o):
        m.set("H", H)
        m.set("ENTHALPY", E)
        m.thermal_diagnostics()
    for name in ("TEMPERATURE", "WATER_FRACTION"):
        assert ulps(h.get(name), o.get(name)) == 0, name
    a, b = h.get("A"), o.get("A")
    assert np.allclose(a, b, rtol=1e-13, atol=0)
    assert np.any(o.get("WATER_FRACTION") > 0)


# ---- full-size configurations and the reference's published benchmark ----------------------------------------------

def test_marine_sheet_640x80_full_size_parity():
    """The grounded + floating marine sheet at its stated size (640 × 80 at 1 km, FP64 MG-PCG with the reference
    defaults): ≤ 1e-9 relative L2 in (u, v), equal Picard counts, masks and classes bit for bit."""
    say("marine sheet 640x80 at 1 km, FP64 MG")
    h, o, ih, io = run_pair(*W.marine_sheet())
    err = rel_l2([h.get("U"), h.get("V")], [o.get("U"), o.get("V")])
    say(f"  picard {ih['picard_iterations']} vs {io['picard_iterations']}, cg {ih['total_linear_iterations']} vs "
        f"{io['total_linear_iterations']}, rel L2 {err:.2e}")
    assert ih["picard_iterations"] == io["picard_iterations"]
    assert err <= 1e-9
    for name in ("MU", "PHI", "PHIU", "PHIV", "S", "N", "TAUU", "TAUV"):
        assert np.array_equal(h.get(name), o.get(name)), name
    assert np.array_equal(h.classes()[0], o.classes()[0]) and np.array_equal(h.classes()[1], o.classes()[1])


def test_diva_shelf_256x256_parity():
    """The DIVA shelf (Arrhenius A on 10 σ levels, regularized Coulomb, no-slip walls) at 256² (2 km), FP64 MG-PCG:
    ≤ 1e-9 relative L2 in (u, v) and in the layer velocities, equal Picard counts."""
    say("DIVA shelf 256x256 at 2 km, FP64 MG")
    g, phys, opts, f = W.diva_shelf(N=256, dx=2e3)
    h, o, ih, io = run_pair(g, phys, opts, f)
    err = rel_l2([h.get("U"), h.get("V")], [o.get("U"), o.get("V")])
    say(f"  picard {ih['picard_iterations']} vs {io['picard_iterations']}, cg {ih['total_linear_iterations']} vs "
        f"{io['total_linear_iterations']}, rel L2 {err:.2e}")
    assert ih["picard_iterations"] == io["picard_iterations"]
    assert err <= 1e-9
    assert rel_l2([h.get("USIGMA"), h.get("VSIGMA")], [o.get("USIGMA"), o.get("VSIGMA")]) <= 1e-9


def test_diva_slab_profile_on_device():
    """test/test_diva.jl:48-58 on the device (Picard, then nk_compute_velocity_profile): profile < 1e-2, mean
    < 1e-2, basal < 1e-6, v = 0, and second order in Nσ; the layer velocities also match the oracle's (≤ 1e-9)."""
    from test_oracle_solvers import check_slab, slab_problem
    say("DIVA slab, Nσ = 10 and 40")
    coarse, _ = slab_problem(HIPModel, 10)
    fine, mh = slab_problem(HIPModel, 40)
    _, mo = slab_problem(O.OracleModel, 40)
    say(f"  fine {fine}")
    check_slab(coarse, fine)
    assert rel_l2([mh.get("USIGMA")], [mo.get("USIGMA")]) <= 1e-9


@pytest.mark.parametrize("dx", [8000.0, 4000.0, 2000.0, 1000.0, 500.0, 250.0])
def test_mismip_scaling_iterations_on_device(dx):
    """The reference's multigrid scaling benchmark (benchmarks/multigrid_scaling.jl) through the C ABI: FP64 Jacobi-
    diagonal and multigrid PCG counts of the timed solve equal the published ones (docs/src/model/numerics.md:44-49)
    at every resolution from 80 × 10 to 2560 × 320; up to 1 km the two solutions also match the oracle's ≤ 1e-9."""
    g, p, o, f = W.mismip_scaling(dx)
    say(f"MISMIP+ scaling {g.Nx}x{g.Ny}")
    h = HIPModel(g, p, o)
    W.load_into(h, f)
    r = W.mismip_timed_solves(h)
    got = (r["jacobi"][0], r["multigrid"][0])
    say(f"  iterations {got} published {W.MISMIP_SCALING_ITERATIONS[dx]}; device ms jacobi "
        f"{r['jacobi'][3]['elapsed_ms']:.2f} mg {r['multigrid'][3]['elapsed_ms']:.2f}")
    assert got == W.MISMIP_SCALING_ITERATIONS[dx]
    if dx >= 1000.0:
        m = O.OracleModel(g, p, o)
        W.load_into(m, f)
        ro = W.mismip_timed_solves(m)
        for name in ("jacobi", "multigrid"):
            assert rel_l2(r[name][1:3], ro[name][1:3]) <= 1e-9, name
