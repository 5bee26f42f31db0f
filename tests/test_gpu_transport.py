"""Thickness transport on the device (nk_compute_thickness_tendency, nk_thickness_substep, nk_cache_thickness,
nk_cell_advection_timescale) against the oracle: bit for bit on random fields over the four topologies and both
face-thickness schemes (the kernels do the oracle's arithmetic in the same order), whole split Runge-Kutta steps bit
for bit, and the reference's transport properties (test/test_thickness.jl) on the device at a larger size."""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import B, P, abi, field_values, grid
from nunatak_hip import transport
from nunatak_hip.model import HIPModel
from test_oracle_transport import transport_model, volume

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ulps(a, b):
    ia, ib = np.asarray(a, np.float64).view(np.int64), np.asarray(b, np.float64).view(np.int64)
    return int(np.max(np.abs(ia - ib)))


ALL_ADVECTION = [abi.ADVECTION_UPWIND1, abi.ADVECTION_CENTERED2, abi.ADVECTION_WENO5, abi.ADVECTION_WENO5_BOUNDED]


@pytest.mark.parametrize("topology", [(B, B), (B, P), (P, B), (P, P)])
@pytest.mark.parametrize("advection", ALL_ADVECTION)
@pytest.mark.parametrize("grounding", [abi.GROUNDING_NONE, abi.GROUNDING_SUBGRID])
def test_tendency_and_substep_bit_exact(topology, advection, grounding):
    say(f"transport {topology} advection {advection} grounding {grounding}")
    rng = np.random.default_rng(21)
    g = grid(37, 29, Lx=37e3, Ly=29e3, topology=topology, Nsigma=2)
    phys = abi.default_physics(masking=0, grounding=grounding)
    h, o = HIPModel(g, phys, abi.default_options()), O.OracleModel(g, phys, abi.default_options())
    f = {"H": 2000 * rng.random((29, 37)) * (rng.random((29, 37)) > 0.2), "BED": -1000 + 1500 * rng.random((29, 37)),
         "U": 1e-5 * rng.standard_normal(h.shape("U")[1::-1]), "V": 1e-5 * rng.standard_normal(h.shape("V")[1::-1]),
         "SMB": 1e-8 * rng.standard_normal((29, 37)), "BASAL_MELT": 1e-8 * rng.random((29, 37)),
         "GROUNDED_MELT": 1e-9 * rng.random((29, 37)), "HCLIP": rng.random((29, 37))}
    for m in (h, o):
        for k, v in f.items():
            m.set(k, v)
        m.update_geometry()
        m.cache_thickness()
        m.compute_thickness_tendency(advection, grounding == abi.GROUNDING_SUBGRID)
    assert ulps(h.get("GH"), o.get("GH")) == 0
    for final in (False, True):
        for m in (h, o):
            m.thickness_substep(3.2e7, final)
        assert ulps(h.get("H"), o.get("H")) == 0 and ulps(h.get("HCLIP"), o.get("HCLIP")) == 0
    assert h.get("H").min() == 0.0 and h.cell_advection_timescale() == o.cell_advection_timescale()


@pytest.mark.parametrize("advection", ALL_ADVECTION)
def test_split_rk3_steps_bit_exact(advection):
    say(f"transport steps, advection {advection}")
    models = [transport_model(background=0.0, width=3e3, melt=lambda x, y: np.where(x > 30e3, 1e-7, 0.0), cls=c)
              for c in (HIPModel, O.OracleModel)]
    dt = 0.5 * models[1][0].cell_advection_timescale()
    for k in range(12):
        for m, _ in models:
            transport.time_step(m, dt, advection=advection, first=(k == 0))
    (h, _), (o, _) = models
    for name in ("H", "HCLIP", "HPREV", "GH"):
        assert ulps(h.get(name), o.get(name)) == 0, name
    assert h.get("HCLIP").max() > 0


@pytest.mark.parametrize("advection", ALL_ADVECTION)
def test_volume_is_conserved_at_1024(advection):
    say(f"transport 1024 advection {advection}")
    m, g = transport_model(advection=advection, Nx=1024, L=1024e3, background=100.0, cls=HIPModel)
    V0 = volume(m, g)
    dt = 0.5 * m.cell_advection_timescale()
    for k in range(32):
        transport.time_step(m, dt, advection=advection, first=(k == 0))
    assert abs(volume(m, g) - V0) / V0 < 1e-13


def test_bounded_weno_keeps_a_compact_bump_nonnegative_on_the_device():
    """test/test_thickness.jl:60-70 on the device: bounded WENO loses < 1e-12 of the volume to the clip, plain WENO
    > 1e-9, and the device's bounded steps equal the oracle's bit for bit."""
    drift, H = [], {}
    for advection, bounds in ((abi.ADVECTION_WENO5_BOUNDED, (0.0, np.inf)), (abi.ADVECTION_WENO5, None)):
        for cls in (HIPModel, O.OracleModel):
            m, g = transport_model(width=2e3, cls=cls)
            if bounds:
                m.set_thickness_bounds(*bounds)
            V0 = volume(m, g)
            dt = 0.25 * m.cell_advection_timescale()
            for k in range(64):
                transport.time_step(m, dt, advection=advection, first=(k == 0))
            H[(advection, cls)] = m.get("H")
            if cls is HIPModel:
                assert m.get("H").min() >= 0
                drift.append(abs(volume(m, g) - V0) / V0)
    say(f"bounded / plain WENO volume lost to the clip: {drift}")
    assert drift[0] < 1e-12 and drift[1] > 1e-9
    for adv in (abi.ADVECTION_WENO5_BOUNDED, abi.ADVECTION_WENO5):
        assert ulps(H[(adv, HIPModel)], H[(adv, O.OracleModel)]) == 0
