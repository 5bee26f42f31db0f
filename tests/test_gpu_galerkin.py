"""Galerkin coarse operators (NK_COARSE_GALERKIN) of the multigrid preconditioner.

No reference counterpart: the reference rediscretizes coarse levels. Checks: the Galerkin-preconditioned PCG
converges to the same solution as the reference's rediscretized MG-PCG (bar 1e-7 relative L2: both stop at a
relative residual of 1e-10), in fewer iterations on the calving-front channel (the case the Galerkin product is
for; the CPU prototype of the same construction needs 29 / 39 iterations at 64² / 128² against the reference's
48 / 80), with every coupling inside the 31-point coarse stencil on all four topologies and on semi-coarsened
hierarchies; and the mixed-precision form with Galerkin operators stays within 1e-8 of the FP64 solve.
"""
import sys

import numpy as np
import pytest

from helpers import B, P, abi
from nunatak_hip import workloads as W
from nunatak_hip.model import HIPModel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def channel(N, coarse, precision=abi.PRECISION_FP64, picard=1, linear=1e-10, topology=(B, B), Ny=None, walls=None,
            smoother=None, preconditioner=None):
    g, phys, opts, fields = W.diva_shelf(N=N, dx=512e3 / N)
    if walls is not None:
        phys.walls = walls
    Ny = Ny or N
    if Ny != N or topology != (B, B):
        g = abi.default_grid(N, Ny, g.dx, g.dy, topology=topology, halo=4, Nsigma=g.Nsigma)
        fields = {k: (v[:Ny] if v.ndim == 2 else v[:, :Ny]) for k, v in fields.items()}
    phys.stress_balance = abi.SSA
    phys.rate_factor_depth = abi.RATE_FACTOR_2D
    fields = dict(fields)
    fields.pop("TEMPERATURE")
    fields["A"] = np.full((Ny, N), 1e-25)
    opts.linear = abi.convergence(linear, 0.0, 20_000)
    opts.picard = abi.fixed_iterations(picard)
    opts.coarse_operator = coarse
    opts.precision = precision
    if smoother is not None:
        opts.smoother = smoother
    if preconditioner is not None:
        opts.preconditioner = preconditioner
    m = HIPModel(g, phys, opts)
    W.load_into(m, fields)
    info = m.update_velocities(geometry=True)
    return m, info


def residual(m):
    """‖A x − τ‖ / ‖τ‖ of the handle's solution through nk_apply_operator (coefficients at the last Picard iterate)."""
    m.set("PU", m.get("U"))
    m.set("PV", m.get("V"))
    m.apply_operator()
    qu, qv, tu, tv = m.get("QU"), m.get("QV"), m.get("TAUU"), m.get("TAUV")
    g = m.grid
    r = np.concatenate([(qu - tu)[: g.Ny, : g.Nx].ravel(), (qv - tv)[: g.Ny, : g.Nx].ravel()])
    b = np.concatenate([tu[: g.Ny, : g.Nx].ravel(), tv[: g.Ny, : g.Nx].ravel()])
    return float(np.linalg.norm(r) / np.linalg.norm(b))


def rel(m1, m2):
    a = np.concatenate([m1.get("U").ravel(), m1.get("V").ravel()])
    b = np.concatenate([m2.get("U").ravel(), m2.get("V").ravel()])
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("N,bound", [(64, 40), (128, 50)])
def test_galerkin_converges_to_the_same_solution_in_fewer_iterations(N, bound):
    mr, ir = channel(N, abi.COARSE_REDISCRETIZED)
    mg, ig = channel(N, abi.COARSE_GALERKIN)
    e = rel(mg, mr)
    say(f"[galerkin] {N}²: rediscretized {ir['total_linear_iterations']} iterations, Galerkin "
        f"{ig['total_linear_iterations']}, rel diff {e:.2e}")
    assert ig["total_linear_iterations"] <= bound
    assert ig["total_linear_iterations"] < ir["total_linear_iterations"]
    assert e < 1e-7


@pytest.mark.parametrize("topology", [(B, B), (B, P), (P, B), (P, P)])
@pytest.mark.parametrize("shape", [(96, 96), (128, 40)])
def test_galerkin_pattern_closes_on_every_topology(topology, shape):
    N, Ny = shape
    mr, ir = channel(N, abi.COARSE_REDISCRETIZED, topology=topology, Ny=Ny)
    mg, ig = channel(N, abi.COARSE_GALERKIN, topology=topology, Ny=Ny)
    e = rel(mg, mr)
    res = residual(mg)
    say(f"[galerkin] {shape} {topology}: levels {mg.multigrid_levels()}, iterations {ir['total_linear_iterations']} → "
        f"{ig['total_linear_iterations']}, rel diff {e:.2e}, residual {res:.2e}")
    assert res < 1e-9
    if topology != (P, P):   # doubly periodic and afloat: rigid translations are a null space, solutions may differ by one
        assert e < 1e-7


def test_mixed_galerkin_matches_fp64():
    m64, i64 = channel(128, abi.COARSE_REDISCRETIZED, picard=4, linear=1e-8)
    mmx, imx = channel(128, abi.COARSE_GALERKIN, precision=abi.PRECISION_MIXED, picard=4, linear=1e-8)
    e = rel(mmx, m64)
    say(f"[galerkin] mixed+Galerkin vs FP64: {e:.2e}; CG {i64['total_linear_iterations']} → {imx['total_linear_iterations']}")
    assert e < 1e-8


def test_mixed_galerkin_on_the_marine_sheet_matches_fp64():
    """Config 2 (640 × 80, grounded + floating, Weertman, semi-coarsened hierarchy): the mixed form with Galerkin
    coarse operators against the reference algorithm in FP64, full Picard solve. Bar: 1e-8 relative L2."""
    out = {}
    for label, prec, coarse in (("fp64", abi.PRECISION_FP64, abi.COARSE_REDISCRETIZED),
                                ("mixed", abi.PRECISION_MIXED, abi.COARSE_GALERKIN)):
        g, phys, opts, fields = W.marine_sheet()
        opts.precision, opts.coarse_operator = prec, coarse
        m = HIPModel(g, phys, opts)
        W.load_into(m, fields)
        info = m.update_velocities(geometry=True)
        out[label] = (m, info)
    e = rel(out["mixed"][0], out["fp64"][0])
    say(f"[galerkin] marine sheet: Picard {out['fp64'][1]['picard_iterations']} / {out['mixed'][1]['picard_iterations']}, "
        f"CG {out['fp64'][1]['total_linear_iterations']} → {out['mixed'][1]['total_linear_iterations']}, rel {e:.2e}")
    assert out["mixed"][1]["picard_iterations"] == out["fp64"][1]["picard_iterations"]
    assert e < 1e-8


@pytest.mark.parametrize("precision,coarse", [(abi.PRECISION_MIXED, abi.COARSE_GALERKIN), (abi.PRECISION_FP64, abi.COARSE_GALERKIN),
                                              (abi.PRECISION_MIXED, abi.COARSE_REDISCRETIZED)])
def test_fused_first_sweeps_equal_separate_sweeps(monkeypatch, precision, coarse):
    """The first two smoothing sweeps from zero run as one kernel (the LDS-tiled k_sweep_pre_t at the finest FP32 level,
    k_g_sweep_pre on Galerkin levels); with NK_NO_PRE_FUSE=1 at creation they run as two. The solutions and iteration
    counts must be identical bit for bit."""
    out = []
    for fuse in (True, False):
        if fuse:
            monkeypatch.delenv("NK_NO_PRE_FUSE", raising=False)
        else:
            monkeypatch.setenv("NK_NO_PRE_FUSE", "1")
        m, info = channel(96, coarse, precision=precision, picard=2, linear=1e-8)
        out.append((m.get("U"), m.get("V"), info["total_linear_iterations"]))
        m.close()
    say(f"[galerkin] fused vs separate first sweeps (precision {precision}, coarse {coarse}): CG {out[0][2]} / {out[1][2]}")
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


MIX, F64 = abi.PRECISION_MIXED, abi.PRECISION_FP64


@pytest.mark.parametrize("topology,coarse,walls,precision", [((B, B), abi.COARSE_GALERKIN, abi.NO_SLIP, MIX),
                                                             ((P, B), abi.COARSE_GALERKIN, abi.NO_SLIP, MIX),
                                                             ((B, P), abi.COARSE_GALERKIN, abi.NO_SLIP, MIX),
                                                             ((P, P), abi.COARSE_GALERKIN, abi.NO_SLIP, MIX),
                                                             ((B, B), abi.COARSE_GALERKIN, abi.FREE_SLIP, MIX),
                                                             ((B, B), abi.COARSE_REDISCRETIZED, abi.NO_SLIP, MIX),
                                                             ((B, B), abi.COARSE_REDISCRETIZED, abi.NO_SLIP, F64),
                                                             ((P, B), abi.COARSE_REDISCRETIZED, abi.FREE_SLIP, F64)])
def test_masked_finest_sweeps_equal_plain_sweeps(monkeypatch, topology, coarse, walls, precision):
    """The finest-level operator applies and sweeps read a 16-bit neighbour mask per point (k_apply_m, k_sweep_m,
    k_sweep_m_dot, k_sweep_m_z64); with NK_NO_NMASK=1 at creation they test the activity flag bytes in place (k_apply,
    k_sweep, k_sweep_dot, k_sweep_z64). Solutions and iteration counts must be identical bit for bit, on every topology,
    with both wall conditions (the peripheral corner bits) and in both precisions."""
    out = []
    for use in (True, False):
        if use:
            monkeypatch.delenv("NK_NO_NMASK", raising=False)
        else:
            monkeypatch.setenv("NK_NO_NMASK", "1")
        m, info = channel(96, coarse, precision=precision, picard=2, linear=1e-8, topology=topology, walls=walls)
        out.append((m.get("U"), m.get("V"), info["total_linear_iterations"]))
        m.close()
    say(f"[galerkin] masked vs plain finest kernels ({topology}, coarse {coarse}, walls {walls}, precision {precision}): "
        f"CG {out[0][2]} / {out[1][2]}")
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("topology,precision", [((B, B), MIX), ((P, B), MIX), ((B, P), MIX), ((P, P), MIX), ((B, B), F64)])
def test_packed_stencil_sweeps_match_full_stencil(monkeypatch, topology, precision):
    """The launched Galerkin sweeps and residuals read a symmetric packed copy of each level's stencil (32 of the 62
    coupling planes; a left-out coupling is read from the column point's row); with NK_NO_GPACK=1 at creation they read
    the full stencil. The packed operator is the exact symmetrization of the full one (the two rows of a pair of the
    Pᵀ A P product can differ in the last bit), so the two preconditioned solves reach the same solution to the
    linear tolerance (bar 1e-9 relative L2 at a relative residual of 1e-10) with nearly the same iteration count. On the
    doubly periodic floating shelf the operator's null space (the rigid translations) leaves the solution defined up to
    a translation, so there both solves are held to the residual bar and the comparison is made without the mean."""
    out = []
    for pack in (True, False):
        if pack:
            monkeypatch.delenv("NK_NO_GPACK", raising=False)
        else:
            monkeypatch.setenv("NK_NO_GPACK", "1")
        m, info = channel(128, abi.COARSE_GALERKIN, precision=precision, picard=2, linear=1e-10, topology=topology)
        out.append((m, info["total_linear_iterations"]))
    if topology == (P, P):
        a = [np.concatenate([(f - f[: m.grid.Ny, : m.grid.Nx].mean()).ravel() for f in (m.get("U"), m.get("V"))]) for m, _ in out]
        e = float(np.linalg.norm(a[0] - a[1]) / np.linalg.norm(a[1]))
    else:
        e = rel(out[0][0], out[1][0])
    r0, r1 = residual(out[0][0]), residual(out[1][0])
    say(f"[galerkin] packed vs full stencil ({topology}, precision {precision}): CG {out[0][1]} / {out[1][1]}, rel {e:.2e}, "
        f"residuals {r0:.1e} / {r1:.1e}")
    for m, _ in out:
        m.close()
    assert e < 1e-9
    assert abs(out[0][1] - out[1][1]) <= max(2, out[1][1] // 20)
    assert r0 < 1e-9 and r1 < 1e-9


@pytest.mark.parametrize("precision,coarse,topology", [(MIX, abi.COARSE_GALERKIN, (B, B)), (MIX, abi.COARSE_GALERKIN, (P, P)),
                                                       (F64, abi.COARSE_GALERKIN, (B, P)), (F64, abi.COARSE_REDISCRETIZED, (P, B))])
def test_lds_tiled_kernels_equal_global_reads(monkeypatch, precision, coarse, topology):
    """The FP64 finest operator applies and sweeps and the packed Galerkin sweeps and residuals stage the iterate
    through LDS tiles (k_apply_mt, k_sweep_mt, k_sweep_mt_dot, k_gp_sweep_t, k_gp_sweep_pre_t, k_gp_residual_t);
    with NK_NO_ATILE=1 at creation they read it from global memory. The same values in the same order, so solutions and iteration counts
    must be identical bit for bit."""
    out = []
    for tile in (True, False):
        if tile:
            monkeypatch.delenv("NK_NO_ATILE", raising=False)
        else:
            monkeypatch.setenv("NK_NO_ATILE", "1")
        m, info = channel(128, coarse, precision=precision, picard=2, linear=1e-8, topology=topology)
        out.append((m.get("U"), m.get("V"), info["total_linear_iterations"]))
        m.close()
    say(f"[galerkin] LDS tiles vs global reads (precision {precision}, coarse {coarse}, {topology}): CG {out[0][2]} / {out[1][2]}")
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("coarse,topology", [(abi.COARSE_GALERKIN, (B, B)), (abi.COARSE_GALERKIN, (P, B)), (abi.COARSE_REDISCRETIZED, (B, B))])
def test_direction_update_from_fp32_z_equals_fp64_copy(monkeypatch, coarse, topology):
    """In the mixed form the last finest sweep no longer stores z64 = (double) z32 and the direction update reads the
    FP32 z (k_update_p32); with NK_NO_Z32=1 at creation the sweep stores z64 and k_update_p reads it. The same values,
    so solutions and iteration counts must be identical bit for bit."""
    out = []
    for z32 in (True, False):
        if z32:
            monkeypatch.delenv("NK_NO_Z32", raising=False)
        else:
            monkeypatch.setenv("NK_NO_Z32", "1")
        m, info = channel(96, coarse, precision=MIX, picard=2, linear=1e-8, topology=topology)
        out.append((m.get("U"), m.get("V"), info["total_linear_iterations"]))
        m.close()
    say(f"[galerkin] FP32 z vs FP64 copy in the direction update (coarse {coarse}, {topology}): CG {out[0][2]} / {out[1][2]}")
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("precision,topology,N,Ny", [(MIX, (B, B), 128, None), (MIX, (P, P), 96, None), (MIX, (B, P), 160, 96),
                                                     (F64, (P, B), 128, None)])
def test_prolongation_inside_first_post_sweep_equals_separate_kernels(monkeypatch, precision, topology, N, Ny):
    """On packed Galerkin levels the first post-smoothing sweep forms x + P e while loading its LDS tile
    (k_gp_sweep_pt); with NK_NO_PRO_FUSE=1 at creation k_prolongate writes x first and k_gp_sweep_t reads it. The same
    values in the same order, so solutions and iteration counts must be identical bit for bit."""
    out = []
    for fused in (True, False):
        if fused:
            monkeypatch.delenv("NK_NO_PRO_FUSE", raising=False)
        else:
            monkeypatch.setenv("NK_NO_PRO_FUSE", "1")
        m, info = channel(N, abi.COARSE_GALERKIN, precision=precision, picard=2, linear=1e-8, topology=topology, Ny=Ny)
        out.append((m.get("U"), m.get("V"), info["total_linear_iterations"]))
        m.close()
    say(f"[galerkin] prolongation in the first post-sweep vs k_prolongate (precision {precision}, {topology}, {N}x{Ny or N}): "
        f"CG {out[0][2]} / {out[1][2]}")
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("coarse", [abi.COARSE_GALERKIN, abi.COARSE_REDISCRETIZED])
def test_mixed_linear_solve_scales_exactly_with_a_tiny_right_hand_side(coarse):
    """The mixed PCG hands the FP32 V-cycle r·zs with zs = 2^−⌊log₂‖r‖⌋ (‖r₀‖ for the first preconditioning, the
    pre-update ‖r‖ inside the loop) and scales its z back, so a right-hand side multiplied by 2^−200, far below FP32's
    range, gives the same iteration count and a solution multiplied by 2^−200 bit for bit (every FP64 and FP32
    operation commutes with a power of two while nothing underflows). Without the scale the FP32 V-cycle would see
    zeros here; at ordinary magnitudes its z falls into FP32 denormals near 1e-12 relative and the PCG stalls: the
    warm-started adjoint solves of the ridge model need ~3600 CG iterations each instead of 13."""
    m, _ = channel(96, coarse, precision=MIX, picard=1, linear=1e-12)
    tu, tv = m.get("TAUU"), m.get("TAUV")
    sols = []
    for scale in (1.0, 2.0 ** -200):
        m.set("TAUU", tu * scale)
        m.set("TAUV", tv * scale)
        m.set("U", 0.0)
        m.set("V", 0.0)
        info = m.solve()
        sols.append((m.get("U"), m.get("V"), info["linear_iterations"]))
    m.close()
    (u1, v1, n1), (u2, v2, n2) = sols
    say(f"[galerkin] mixed solve with rhs x 2^-200 (coarse {coarse}): CG {n1} / {n2}")
    assert n1 == n2 and n1 > 0
    assert np.array_equal(u2, u1 * 2.0 ** -200) and np.array_equal(v2, v1 * 2.0 ** -200)


@pytest.mark.parametrize("N", [128, 256])
def test_refined_mixed_form_matches_fp64(N):
    """NK_PRECISION_MIXED_REFINED — FP64 iterative refinement (defect, norm and stopping rule in FP64) around an FP32
    PCG with the FP32 Galerkin V-cycle — on the DIVA shelf: ≤ 1e-8 relative L2 in (u, v) against the FP64 solve, at
    least one refinement step per linear solve, and the last linear solve stopped on the FP64 defect."""
    out = {}
    for prec, coarse in ((abi.PRECISION_FP64, abi.COARSE_REDISCRETIZED), (abi.PRECISION_MIXED_REFINED, abi.COARSE_GALERKIN)):
        g, phys, opts, f = W.diva_shelf(N=N, dx=512e3 / N, precision=prec)
        opts.coarse_operator = coarse
        m = HIPModel(g, phys, opts)
        W.load_into(m, f)
        info = m.update_velocities()
        out[prec] = (m.get("U"), m.get("V"), info)
    a = np.concatenate([x.ravel() for x in out[abi.PRECISION_MIXED_REFINED][:2]])
    b = np.concatenate([x.ravel() for x in out[abi.PRECISION_FP64][:2]])
    err = np.linalg.norm(a - b) / np.linalg.norm(b)
    info = out[abi.PRECISION_MIXED_REFINED][2]
    say(f"refined {N}^2: picard {info['picard_iterations']} inner CG {info['total_linear_iterations']} refinement steps "
        f"(last solve) {info['refinement_steps']}; FP64 picard {out[abi.PRECISION_FP64][2]['picard_iterations']}; rel L2 {err:.2e}")
    assert err <= 1e-8
    assert info["refinement_steps"] >= 1


@pytest.mark.parametrize("precision,topology,N,Ny", [(MIX, (B, B), 128, None), (MIX, (P, P), 96, None), (MIX, (B, P), 160, 96),
                                                     (F64, (P, B), 128, None)])
def test_latency_bound_forms_equal_bandwidth_forms(monkeypatch, precision, topology, N, Ny):
    """On small packed Galerkin levels the sweeps and residuals issue every load of a thread (tile elements and the
    row's 31 couplings) before one wait (k_gp_sweep_tl, k_gp_residual_tl, k_gp_sweep_pre_tl); larger levels load the
    tile, then the couplings after the block's barrier. NK_LATENCY_POINTS at creation sets the largest level that
    takes the first forms: 0 (none) and 10⁸ (all) must give the same solutions and iteration counts bit for bit."""
    out = []
    for points in ("0", "100000000"):
        monkeypatch.setenv("NK_LATENCY_POINTS", points)
        m, info = channel(N, abi.COARSE_GALERKIN, precision=precision, picard=2, linear=1e-8, topology=topology, Ny=Ny)
        out.append((m.get("U"), m.get("V"), info["total_linear_iterations"]))
        m.close()
    monkeypatch.delenv("NK_LATENCY_POINTS", raising=False)
    say(f"[galerkin] latency-bound vs bandwidth-bound level kernels (precision {precision}, {topology}, {N}x{Ny or N}): "
        f"CG {out[0][2]} / {out[1][2]}")
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("precision,topology,N,Ny", [(MIX, (B, B), 128, None), (MIX, (P, P), 96, None), (F64, (B, P), 128, 64),
                                                     (MIX, (P, B), 160, 96), (F64, (P, P), 64, None)])
def test_register_inverse_equals_lds_inverse(monkeypatch, precision, topology, N, Ny):
    """The coarsest Galerkin level's dense generalized inverse is eliminated with the matrix in registers (a column of
    16 rows per thread; the pivot row, its reciprocal pivot and the pivot column shared through LDS, one barrier per
    step, the pivot column formed as (−0) − f·(1/p)); with NK_NO_GINV_REGS=1 at creation the same
    Gauss-Jordan steps run element by element on the matrix in LDS. Each element sees the same operations in the same
    order, so solutions and iteration counts must be identical bit for bit — also on doubly periodic floating shelves,
    whose rigid translations are null directions that the elimination drops."""
    out = []
    for regs in (True, False):
        if regs:
            monkeypatch.delenv("NK_NO_GINV_REGS", raising=False)
        else:
            monkeypatch.setenv("NK_NO_GINV_REGS", "1")
        m, info = channel(N, abi.COARSE_GALERKIN, precision=precision, picard=2, linear=1e-8, topology=topology, Ny=Ny)
        out.append((m.get("U"), m.get("V"), info["total_linear_iterations"]))
        m.close()
    monkeypatch.delenv("NK_NO_GINV_REGS", raising=False)
    say(f"[galerkin] register vs LDS coarsest inverse (precision {precision}, {topology}, {N}x{Ny or N}): "
        f"CG {out[0][2]} / {out[1][2]}")
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("precision,topology,N,Ny,latency", [(MIX, (B, B), 128, None, None), (MIX, (P, P), 96, None, None),
                                                             (MIX, (B, P), 160, 96, None), (F64, (P, B), 128, None, None),
                                                             (MIX, (B, B), 256, None, "100000000"),
                                                             (MIX, (P, B), 200, 120, "100000000"),
                                                             (F64, (B, P), 144, 48, "100000000")])
def test_residual_and_restriction_in_one_launch_equal_separate_kernels(monkeypatch, precision, topology, N, Ny, latency):
    """On latency-bound packed Galerkin levels the residual and the restriction to the level below run as one kernel
    (k_gp_residual_restrict: the residuals the restriction reads formed in LDS at the fine indices k_restrict reads);
    with NK_NO_RR_FUSE=1 at creation k_gp_residual_tl stores r and k_restrict reads it. The same values in the same order,
    so solutions and iteration counts must be identical bit for bit — on all four topologies, on semi-coarsened
    hierarchies and with every level below the finest taking the fused form (NK_LATENCY_POINTS = 10⁸)."""
    out = []
    if latency:
        monkeypatch.setenv("NK_LATENCY_POINTS", latency)
    for fused in (True, False):
        if fused:
            monkeypatch.delenv("NK_NO_RR_FUSE", raising=False)
        else:
            monkeypatch.setenv("NK_NO_RR_FUSE", "1")
        m, info = channel(N, abi.COARSE_GALERKIN, precision=precision, picard=2, linear=1e-8, topology=topology, Ny=Ny)
        out.append((m.get("U"), m.get("V"), info["total_linear_iterations"]))
        m.close()
    monkeypatch.delenv("NK_NO_RR_FUSE", raising=False)
    monkeypatch.delenv("NK_LATENCY_POINTS", raising=False)
    say(f"[galerkin] residual + restriction fused vs separate (precision {precision}, {topology}, {N}x{Ny or N}, latency "
        f"points {latency}): CG {out[0][2]} / {out[1][2]}")
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("precision,coarse,topology,N,Ny,smoother,blocks",
                         [(F64, "diagonal", (B, B), 96, None, None, None),
                          (MIX, abi.COARSE_GALERKIN, (B, B), 128, None, None, None),
                          (MIX, abi.COARSE_GALERKIN, (P, P), 96, None, None, None),
                          (F64, abi.COARSE_REDISCRETIZED, (P, B), 128, None, None, None),
                          (MIX, abi.COARSE_REDISCRETIZED, (B, P), 160, 96, abi.LINE_JACOBI, None),
                          (F64, abi.COARSE_GALERKIN, (B, B), 256, None, None, None),
                          (MIX, abi.COARSE_GALERKIN, (B, B), 1024, 520, None, "1000000000")])
def test_folded_finishers_equal_k_finish(monkeypatch, precision, coarse, topology, N, Ny, smoother, blocks):
    """On grids of at most NK_FOLD_BLOCKS workgroups the multigrid PCG's α and β finishers run inside k_update_xr and
    the direction update (every block sums the reduction slots itself, in k_finish's order) instead of as k_finish
    launches; NK_NO_FOLD=1 at creation keeps the launches. Solutions and iteration counts must be identical bit for bit
    — with the FP32 V-cycle's z read directly (k_update_p32), through the FP64 z (line smoother: k_promote_dot's r·z
    with r·r beside it), in the FP64 form (k_update_p), and on a 1024 × 520 grid with folding forced (2,080 slots per
    value, more than one load batch per thread). With the diagonal preconditioner only α is folded (into k_update_diag)."""
    out = []
    if blocks:
        monkeypatch.setenv("NK_FOLD_BLOCKS", blocks)
    for folded in (True, False):
        if folded:
            monkeypatch.delenv("NK_NO_FOLD", raising=False)
        else:
            monkeypatch.setenv("NK_NO_FOLD", "1")
        if coarse == "diagonal":
            m, info = channel(N, abi.COARSE_REDISCRETIZED, precision=precision, picard=2, linear=1e-8, topology=topology,
                              Ny=Ny, preconditioner=abi.PRECONDITIONER_DIAGONAL)
        else:
            m, info = channel(N, coarse, precision=precision, picard=2, linear=1e-8, topology=topology, Ny=Ny, smoother=smoother)
        out.append((m.get("U"), m.get("V"), info["total_linear_iterations"]))
        m.close()
    monkeypatch.delenv("NK_NO_FOLD", raising=False)
    monkeypatch.delenv("NK_FOLD_BLOCKS", raising=False)
    say(f"[pcg] folded vs launched finishers (precision {precision}, coarse {coarse}, {topology}, {N}x{Ny or N}, "
        f"smoother {smoother}): CG {out[0][2]} / {out[1][2]}")
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("precision,topology,N,Ny,walls", [(MIX, (B, B), 128, None, None), (MIX, (P, P), 96, None, None),
                                                           (F64, (P, B), 128, 64, None), (MIX, (B, P), 160, 96, None),
                                                           (MIX, (B, B), 96, None, abi.FREE_SLIP)])
def test_first_product_over_the_finest_pattern_equals_full_pattern(monkeypatch, precision, topology, N, Ny, walls):
    """The first Galerkin product P^T A P reads, of the finest level's assembled stencil, only the 9 couplings per row
    the rediscretized operator fills (k_g_assemble writes only those); with NK_NO_FIN_PATTERN=1 at creation it reads all
    31 and skips the zeros as it meets them. The same terms in the same order, so solutions and iteration counts must
    be identical bit for bit."""
    out = []
    for fin in (True, False):
        if fin:
            monkeypatch.delenv("NK_NO_FIN_PATTERN", raising=False)
        else:
            monkeypatch.setenv("NK_NO_FIN_PATTERN", "1")
        m, info = channel(N, abi.COARSE_GALERKIN, precision=precision, picard=2, linear=1e-8, topology=topology, Ny=Ny,
                          walls=walls)
        out.append((m.get("U"), m.get("V"), info["total_linear_iterations"]))
        m.close()
    monkeypatch.delenv("NK_NO_FIN_PATTERN", raising=False)
    say(f"[galerkin] first product over the finest pattern vs all 31 couplings (precision {precision}, {topology}, "
        f"{N}x{Ny or N}, walls {walls}): CG {out[0][2]} / {out[1][2]}")
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("precision,topology,N,Ny,walls", [(MIX, (B, B), 128, None, None), (MIX, (P, P), 96, None, None),
                                                           (F64, (P, B), 128, 64, None), (MIX, (B, P), 160, 96, None),
                                                           (MIX, (B, B), 96, None, abi.FREE_SLIP), (F64, (B, B), 144, 48, None)])
@pytest.mark.parametrize("switch", ["NK_NO_RAP_SPLIT", "NK_NO_RAP_COMPACT"])
def test_split_triple_product_equals_one_lane_per_support_point(monkeypatch, switch, precision, topology, N, Ny, walls):
    """The Galerkin products divide each restriction support point's fine couplings between two lanes by column
    component (their coarse targets are disjoint slots); with NK_NO_RAP_SPLIT=1 at creation one lane takes them all.
    The products also write the LDS tail's compact stencil copies; with NK_NO_RAP_COMPACT=1 k_g_compact copies them
    afterwards. Every coarse coupling receives the same terms in the same order and the tail reads the same values, so
    solutions and iteration counts must be identical bit for bit — on all four topologies, with free-slip walls and on
    a semi-coarsened hierarchy."""
    out = []
    for split in (True, False):
        if split:
            monkeypatch.delenv(switch, raising=False)
        else:
            monkeypatch.setenv(switch, "1")
        m, info = channel(N, abi.COARSE_GALERKIN, precision=precision, picard=2, linear=1e-8, topology=topology, Ny=Ny,
                          walls=walls)
        out.append((m.get("U"), m.get("V"), info["total_linear_iterations"]))
        m.close()
    monkeypatch.delenv(switch, raising=False)
    say(f"[galerkin] {switch} off vs on (precision {precision}, {topology}, {N}x{Ny or N}, walls {walls}): "
        f"CG {out[0][2]} / {out[1][2]}")
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
