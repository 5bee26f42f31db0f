"""Serving form (nk_server_*): continuous batching of velocity solves of uneven sizes.

Bars: a request's (u, v) from the server — alone, or mixed with others in flight and on a reused handle — equals the
same request solved on a fresh handle bit for bit (the kernels are deterministic: no atomics, fixed reduction
order); HBM admission never lets the handles of overlapping requests exceed the budget; a request that can never
fit fails with NK_OUT_OF_MEMORY without blocking the queue; nk_estimate_device_bytes equals what nk_create allocates.
"""
import sys

import numpy as np
import pytest

from helpers import abi
from nunatak_hip import workloads as W
from nunatak_hip.model import HIPModel, NunatakError, estimate_device_bytes
from nunatak_hip.server import Server, serve

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def alone(g, p, o, f):
    m = HIPModel(g, p, o)
    W.load_into(m, f)
    m.update_velocities(geometry=True)
    u, v = m.get("U"), m.get("V")
    m.close()
    return u, v


@pytest.mark.parametrize("case", ["fp64-diag", "fp64-mg", "mixed-mg-galerkin", "diva-3d"])
def test_estimate_matches_allocation(case):
    g, p, o, f = W.serving_request(96, 1)
    if case == "fp64-diag":
        o.preconditioner, o.precision = abi.PRECONDITIONER_DIAGONAL, abi.PRECISION_FP64
        o.coarse_operator = abi.COARSE_REDISCRETIZED
    elif case == "fp64-mg":
        o.precision, o.coarse_operator = abi.PRECISION_FP64, abi.COARSE_REDISCRETIZED
    elif case == "diva-3d":
        g, p, o, f = W.diva_shelf(N=80, dx=1e3)
    m = HIPModel(g, p, o)
    assert m.device_bytes() == estimate_device_bytes(g, p, o)
    say(f"[serve] {case}: {m.device_bytes()} bytes")


def test_results_same_alone_and_in_the_mix():
    sides = [64, 96, 128, 64, 160, 96, 64, 128, 80, 64]
    reqs = [W.serving_request(n, 1000 + k) for k, n in enumerate(sides)]
    say("[serve] solving each request alone")
    ref = [alone(*r) for r in reqs]
    say("[serve] solving the mix on 4 slots")
    out = serve([(0.0, *r) for r in reqs], slots=4, keep_results=True)
    reused = sum(r["reused_handle"] for r in out["records"])
    overlap = sum(1 for a in out["records"] for b in out["records"]
                  if a is not b and a["start_ms"] < b["finish_ms"] and b["start_ms"] < a["finish_ms"])
    say(f"[serve] reused handles {reused}, overlapping pairs {overlap}, p50 {out['p50_ms']:.1f} ms")
    assert reused >= 1 and overlap >= 1
    for k, ((u, v), (ur, vr)) in enumerate(zip(out["results"], ref)):
        assert np.array_equal(u, ur) and np.array_equal(v, vr), f"request {k} (side {sides[k]}) differs"
    # the same requests again on the warm server pool, in reverse order
    out2 = serve([(0.0, *r) for r in reqs[::-1]], slots=3, keep_results=True)
    for k, (u, v) in enumerate(out2["results"][::-1]):
        assert np.array_equal(u, ref[k][0]) and np.array_equal(v, ref[k][1])


def test_admission_keeps_handles_within_budget():
    sides = [64, 128, 96, 128, 64, 96]
    reqs = [W.serving_request(n, 2000 + k) for k, n in enumerate(sides)]
    need = [estimate_device_bytes(g, p, o) for g, p, o, f in reqs]
    budget = int(1.5 * max(need))
    out = serve([(0.0, *r) for r in reqs], slots=4, hbm_budget_bytes=budget)
    recs = out["records"]
    for a in recs:
        live = sum(b["device_bytes"] for b in recs if b["admit_ms"] <= a["admit_ms"] < b["finish_ms"])
        assert live <= budget, (live, budget)
    say(f"[serve] budget {budget}: all {len(recs)} served, p99 {out['p99_ms']:.1f} ms")


def test_request_larger_than_budget_fails_cleanly():
    g, p, o, f = W.serving_request(64, 7)
    with Server(slots=2, hbm_budget_bytes=1 << 20) as srv:
        rid = srv.submit(g, p, o, f)
        with pytest.raises(NunatakError, match="OUT_OF_MEMORY"):
            srv.wait(rid)


@pytest.mark.parametrize("kind", ["diva", "fp64"])
def test_other_request_kinds_same_alone_and_in_the_mix(kind):
    """DIVA requests (the DIVA shelf's generator, 3D Arrhenius A from temperature) and FP64 requests through the server."""
    reqs = []
    for k, n in enumerate([64, 80, 64, 96]):
        if kind == "diva":
            g, p, o, f = W.diva_shelf(N=n, dx=4e3, seed=300 + k)
        else:
            g, p, o, f = W.serving_request(n, 3000 + k, precision=abi.PRECISION_FP64)
        reqs.append((g, p, o, f))
    ref = []
    for g, p, o, f in reqs:
        m = HIPModel(g, p, o)
        W.load_into(m, f)
        m.update_velocities(geometry=True)
        ref.append((m.get("U"), m.get("V")))
        m.close()
    out = serve([(0.0, *r) for r in reqs], slots=3, keep_results=True)
    for k, (u, v) in enumerate(out["results"]):
        assert np.array_equal(u, ref[k][0]) and np.array_equal(v, ref[k][1]), f"request {k} differs"


def test_stream_requests_with_large_grids_same_alone_and_in_the_mix():
    """Requests of the serving stream itself (seed 5): its first 12 — sides 592, 608, 272, 144, 80, 192, 192, 80, 80,
    1024, 384, 128 — and request 174 (side 512), all arriving at once on 4 slots, so the 1024² and 512² solves overlap
    the small ones. Each result equals the same request solved alone on a fresh handle, bit for bit."""
    arrivals, sides = W.serving_schedule(2000)
    picks = list(range(12)) + [174]
    assert 1024 in sides[picks] and 512 in sides[picks]
    reqs = [W.serving_stream_request(k, sides) for k in picks]
    say(f"[serve] stream requests {picks}, sides {[int(sides[k]) for k in picks]}: alone")
    ref = [alone(*r) for r in reqs]
    say("[serve] the same in the mix on 4 slots")
    out = serve([(0.0, *r) for r in reqs], slots=4, keep_results=True)
    overlap = sum(1 for a in out["records"] for b in out["records"]
                  if a is not b and a["start_ms"] < b["finish_ms"] and b["start_ms"] < a["finish_ms"])
    say(f"[serve] overlapping pairs {overlap}, p50 {out['p50_ms']:.1f} ms, p99 {out['p99_ms']:.1f} ms")
    assert overlap >= 1
    for k, ((u, v), (ur, vr)) in enumerate(zip(out["results"], ref)):
        assert np.array_equal(u, ur) and np.array_equal(v, vr), f"stream request {picks[k]} (side {sides[picks[k]]}) differs"
