"""GPU checks of the handle's life cycle at the C ABI: what nk_set_options and nk_set_physics may change on a live
handle, and how failures surface.

Bars: a handle after an allowed change returns (u, v) bit for bit what a fresh handle with the new settings returns;
refused changes return NK_UNSUPPORTED and leave the handle working; a NaN in an input returns NK_NONFINITE (the
reference has no such check: its PCG stops on a NaN residual as if converged, src/Solvers/conjugate_gradient.jl:193)."""
import sys

import numpy as np
import pytest

from helpers import abi
from nunatak_hip import workloads as W
from nunatak_hip.model import HIPModel, NunatakError
from nunatak_hip.server import Server

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def solve_fresh(g, phys, opts, fields):
    h = HIPModel(g, phys, opts)
    W.load_into(h, fields)
    info = h.update_velocities()
    out = (h.get("U").copy(), h.get("V").copy(), info)
    h.close()
    return out


def channel(N=48, **kw):
    g, phys, opts, fields = W.floating_channel(Nx=N - 1, Ny=N // 2, dx=2e3, **kw)
    return g, phys, opts, fields


@pytest.mark.parametrize("mixed", [False, True])
def test_wall_change_on_a_live_handle_equals_a_fresh_handle(mixed):
    """The captured PCG chunk carries the wall factor: after nk_set_physics changes the walls, the next solve must
    be the fresh handle's bit for bit (graphs on, the default)."""
    g, phys, opts, fields = channel(preconditioner=abi.PRECONDITIONER_MULTIGRID)
    if mixed:
        opts.precision, opts.coarse_operator = abi.PRECISION_MIXED, abi.COARSE_GALERKIN
    free = abi.default_physics(**{k: getattr(phys, k) for k, _ in phys._fields_ if not k.startswith("pad")})
    free.walls = abi.FREE_SLIP
    say(f"wall change, mixed={mixed}")
    h = HIPModel(g, free, opts)
    W.load_into(h, fields)
    h.update_velocities()
    h.set_physics(phys)                 # no-slip
    W.load_into(h, fields)
    info = h.update_velocities()
    u, v = h.get("U"), h.get("V")
    uf, vf, infof = solve_fresh(g, phys, opts, fields)
    assert info["total_linear_iterations"] == infof["total_linear_iterations"]
    assert np.array_equal(u, uf) and np.array_equal(v, vf)


def test_multigrid_handle_switches_to_diagonal_and_back():
    g, phys, opts, fields = channel(preconditioner=abi.PRECONDITIONER_MULTIGRID)
    diag = abi.default_options(linear=opts.linear, picard=opts.picard, preconditioner=abi.PRECONDITIONER_DIAGONAL)
    h = HIPModel(g, phys, opts)
    W.load_into(h, fields)
    h.update_velocities()
    h.set_options(diag)
    W.load_into(h, fields)
    info = h.update_velocities()
    ud, vd, infod = solve_fresh(g, phys, diag, fields)
    assert info["total_linear_iterations"] == infod["total_linear_iterations"]
    assert np.array_equal(h.get("U"), ud) and np.array_equal(h.get("V"), vd)
    h.set_options(opts)
    W.load_into(h, fields)
    info = h.update_velocities()
    um, vm, infom = solve_fresh(g, phys, opts, fields)
    assert info["total_linear_iterations"] == infom["total_linear_iterations"]
    assert np.array_equal(h.get("U"), um) and np.array_equal(h.get("V"), vm)


def test_changes_fixed_at_creation_are_refused():
    g, phys, opts, fields = channel(preconditioner=abi.PRECONDITIONER_MULTIGRID)
    opts.precision, opts.coarse_operator = abi.PRECISION_MIXED, abi.COARSE_GALERKIN
    h = HIPModel(g, phys, opts)
    W.load_into(h, fields)
    refused = []
    for name, value in [("coarse_operator", abi.COARSE_REDISCRETIZED), ("precision", abi.PRECISION_FP64),
                        ("minimum_coarse_size", 4), ("smoother", abi.LINE_JACOBI)]:
        o = abi.default_options(**{k: getattr(opts, k) for k, _ in opts._fields_})
        setattr(o, name, value)
        if name == "smoother":
            o.coarse_operator = abi.COARSE_REDISCRETIZED
        with pytest.raises(NunatakError) as e:
            h.set_options(o)
        refused.append((name, e.value.status))
    assert all(st in (abi.NK_UNSUPPORTED, abi.NK_INVALID_ARGUMENT) for _, st in refused), refused
    # a rediscretized FP64 multigrid handle refuses Galerkin operators and line smoothing
    g2, phys2, opts2, fields2 = channel(preconditioner=abi.PRECONDITIONER_MULTIGRID)
    h2 = HIPModel(g2, phys2, opts2)
    for name, value in [("coarse_operator", abi.COARSE_GALERKIN), ("smoother", abi.LINE_JACOBI)]:
        o = abi.default_options(**{k: getattr(opts2, k) for k, _ in opts2._fields_})
        setattr(o, name, value)
        with pytest.raises(NunatakError) as e:
            h2.set_options(o)
        assert e.value.status == abi.NK_UNSUPPORTED
    # a diagonal handle has no hierarchy to switch to
    g3, phys3, opts3, fields3 = channel()
    h3 = HIPModel(g3, phys3, opts3)
    o = abi.default_options(**{k: getattr(opts3, k) for k, _ in opts3._fields_})
    o.preconditioner = abi.PRECONDITIONER_MULTIGRID
    with pytest.raises(NunatakError) as e:
        h3.set_options(o)
    assert e.value.status == abi.NK_UNSUPPORTED
    # the refused handles still solve as fresh ones
    info = h.update_velocities()
    u, v, infof = solve_fresh(g, phys, opts, fields)
    assert info["total_linear_iterations"] == infof["total_linear_iterations"]
    assert np.array_equal(h.get("U"), u)


@pytest.mark.parametrize("form", ["diagonal", "multigrid", "mixed"])
def test_nan_input_returns_nonfinite(form):
    pre = abi.PRECONDITIONER_DIAGONAL if form == "diagonal" else abi.PRECONDITIONER_MULTIGRID
    g, phys, opts, fields = channel(preconditioner=pre)
    if form == "mixed":
        opts.precision, opts.coarse_operator = abi.PRECISION_MIXED, abi.COARSE_GALERKIN
    A = fields["A"].copy()
    A[5, 7] = np.nan
    say(f"NaN in A, {form}")
    h = HIPModel(g, phys, opts)
    W.load_into(h, dict(fields, A=A))
    with pytest.raises(NunatakError) as e:
        h.update_velocities()
    assert e.value.status == abi.NK_NONFINITE, str(e.value)
    # a NaN thickness reaches the driving stress, hence the right-hand side
    H = fields["H"].copy()
    H[3, 4] = np.nan
    W.load_into(h, dict(fields, H=H))
    with pytest.raises(NunatakError) as e:
        h.update_velocities()
    assert e.value.status == abi.NK_NONFINITE
    # the handle recovers with finite inputs
    W.load_into(h, fields)
    info = h.update_velocities()
    u, v, infof = solve_fresh(g, phys, opts, fields)
    assert info["total_linear_iterations"] == infof["total_linear_iterations"]
    assert np.array_equal(h.get("U"), u) and np.array_equal(h.get("V"), v)


def test_nan_request_fails_alone_in_the_server():
    reqs = []
    for k, N in enumerate([64, 80, 64]):
        g, p, o, f = W.serving_request(N, 900 + k)
        if k == 1:
            f = dict(f, A=np.where(np.arange(f["A"].size).reshape(f["A"].shape) == 99, np.nan, f["A"]))
        reqs.append((g, p, o, f))
    with Server(slots=2) as srv:
        ids = [srv.submit(*r) for r in reqs]
        out = []
        for rid in ids:
            try:
                out.append(srv.wait(rid))
            except NunatakError as e:
                out.append(e)
    assert isinstance(out[1], NunatakError) and out[1].status == abi.NK_NONFINITE
    assert not isinstance(out[0], NunatakError) and not isinstance(out[2], NunatakError)
    assert np.all(np.isfinite(out[0][0])) and np.all(np.isfinite(out[2][0]))


def test_not_converged_is_reported_only_when_asked():
    g, phys, opts, fields = channel()
    opts.picard = abi.convergence(1e-12, 0.0, 3)
    h = HIPModel(g, phys, opts)
    W.load_into(h, fields)
    info = h.update_velocities()            # silent, as the reference
    assert info["picard_iterations"] == 3
    opts.report_not_converged = 1
    h.set_options(opts)
    W.load_into(h, fields)
    with pytest.raises(NunatakError) as e:
        h.update_velocities()
    assert e.value.status == abi.NK_NOT_CONVERGED
    opts.linear = abi.convergence(1e-14, 0.0, 5)
    h.set_options(opts)
    with pytest.raises(NunatakError) as e:
        h.solve()
    assert e.value.status == abi.NK_NOT_CONVERGED
