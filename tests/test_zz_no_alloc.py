"""nk_apply_operator_noalloc (the allocation-free variant of nk_apply_operator) through the C ABI against
nk_apply_operator on the same device data: bit for bit, on two calls in a row, the second on a smaller grid (the fixed
scratch array keeps the first call's flags, which must not leak into the second)."""
import ctypes as ct
import sys

import numpy as np
import pytest

from helpers import B, P, abi, grid
from nunatak_hip.model import HIPModel, load

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def coefficients(rng, g):
    from test_gpu_parity import random_coefficients
    return random_coefficients(rng, g)


def run_case(rng, g, walls):
    L = load()
    h = HIPModel(g, abi.default_physics(walls=walls), abi.default_options())
    for name, a in coefficients(rng, g).items():
        h.set(name, a)
    pu = rng.standard_normal(h.shape("PU")[1::-1])
    pv = rng.standard_normal(h.shape("PV")[1::-1])
    h.set("PU", pu)
    h.set("PV", pv)
    h.set("QU", 0.0)
    h.set("QV", 0.0)
    h.apply_operator()
    qu, qv = h.get("QU"), h.get("QV")
    h.set("LAMBDAU", 0.0)
    h.set("LAMBDAV", 0.0)
    ptr = {n: h.device_pointer(n)[0] for n in ("ETAH", "ETAHFF", "BETAU", "BETAV", "PU", "PV", "LAMBDAU", "LAMBDAV")}
    pitch = h.device_pointer("ETAH")[1]
    st = L.nk_apply_operator_noalloc(ct.byref(g), walls, ptr["ETAH"], ptr["ETAHFF"], ptr["BETAU"], ptr["BETAV"], ptr["PU"],
                                     ptr["PV"], ptr["LAMBDAU"], ptr["LAMBDAV"], pitch, L.nk_stream(h._h))
    assert st == abi.NK_OK, L.nk_last_error(None).decode()
    h.synchronize()
    au, av = h.get("LAMBDAU"), h.get("LAMBDAV")
    assert np.array_equal(au, qu) and np.array_equal(av, qv)
    assert np.abs(qu).max() > 0 and np.abs(qv).max() > 0
    h.close()


def test_noalloc_operator_equals_nk_apply_operator_on_two_calls():
    rng = np.random.default_rng(17)
    g1 = grid(200, 150, Lx=200e3, Ly=150e3, topology=(B, B))
    g2 = grid(96, 64, Lx=96e3, Ly=64e3, topology=(P, B))
    say("noalloc operator: 200x150 (Bounded, Bounded) no-slip, then 96x64 (Periodic, Bounded) free-slip")
    run_case(rng, g1, abi.NO_SLIP)
    run_case(rng, g2, abi.FREE_SLIP)


def test_noalloc_rejects_oversized_grids_and_wrong_pitch():
    L = load()
    big = grid(3000, 3000, Lx=3e6, Ly=3e6)
    fake = ct.c_void_p(16)   # never dereferenced: the size check returns before any device work
    args = [fake] * 8
    st = L.nk_apply_operator_noalloc(ct.byref(big), abi.FREE_SLIP, *args, 3008, None)
    assert st == abi.NK_INVALID_ARGUMENT
    assert b"too large" in L.nk_last_error(None)
    small = grid(100, 50, Lx=1e5, Ly=5e4)
    st = L.nk_apply_operator_noalloc(ct.byref(small), abi.FREE_SLIP, *args, 101, None)
    assert st == abi.NK_INVALID_ARGUMENT
    assert b"pitch" in L.nk_last_error(None)
