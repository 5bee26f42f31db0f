"""CPU checks of the drop-in boundary: libnunatak_hip.so loads, exports every symbol include/nunatak_hip.h
declares, and its defaults are the reference constructors' defaults. No compute call runs here (no GPU)."""
import ctypes as ct
import os
import re

import pytest

from helpers import ROOT, abi
from nunatak_hip import model as M

HEADER = os.path.join(ROOT, "include", "nunatak_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(nk_[a-z_]+)\s*\(", text)))


def test_library_is_built_and_loads():
    assert os.path.exists(M.LIB_PATH), "run __graft_entry__.build() first"
    L = M.load()
    assert L.nk_abi_version() == 1


def test_every_declared_symbol_is_exported():
    L = M.load()
    decl = declared_symbols()
    assert len(decl) >= 25
    missing = [name for name in decl if not hasattr(L, name)]
    assert missing == []
    assert sorted(M.EXPORTS) == decl


def test_defaults_match_reference_constructors():
    L = M.load()
    p, o, g = abi.Physics(), abi.SolverOptions(), abi.Grid()
    L.nk_default_physics(ct.byref(p))
    L.nk_default_options(ct.byref(o))
    L.nk_default_grid(ct.byref(g))
    # IceSheetConstants (src/Models/ice_sheet_constants.jl:33), GlenFlowLaw (src/Rheologies/glen_flow_law.jl:114),
    # WeertmanFriction (src/Frictions/frictions.jl:34), SmoothIceMask (src/Models/smooth_ice_mask.jl:41)
    assert (p.ice_density, p.ocean_density, p.gravitational_acceleration) == (918.0, 1028.0, 9.81)
    assert (p.glen_exponent, p.strain_rate_regularization) == (3.0, 1e-13)
    assert (p.friction_exponent, p.velocity_regularization) == (1 / 3, 1e-8)
    assert (p.mask_minimum_thickness, p.mask_transition_thickness) == (1.0, 1.0)
    # StressBalanceSolver (src/StressBalances/stress_balance_solver.jl:43-52), multigrid (multigrid_preconditioner.jl:57)
    assert (o.linear.reltol, o.linear.maxiter, o.picard.reltol, o.picard.maxiter) == (1e-8, 10_000, 1e-6, 100)
    assert o.preconditioner == abi.PRECONDITIONER_DIAGONAL
    assert (o.smoother, o.smoothing_sweeps, o.coarsest_sweeps, o.minimum_coarse_size) == (abi.POINT_JACOBI, 2, 20, 2)
    # SigmaGrid / IceSheetModel (src/Models/ice_sheet_model.jl:119-120)
    assert (g.Nsigma, g.bed_refinement) == (10, 2.0)
    # the Python mirrors agree with the C defaults
    pp, oo = abi.default_physics(), abi.default_options()
    for name, _ in abi.Physics._fields_:
        assert getattr(pp, name) == getattr(p, name), name
    for name in ("preconditioner", "smoother", "smoothing_sweeps", "coarsest_sweeps", "minimum_coarse_size", "precision"):
        assert getattr(oo, name) == getattr(o, name), name


def test_struct_sizes_match_header_layout():
    # the ctypes mirrors must have the C layout: all-double/int32 structs with natural alignment
    assert ct.sizeof(abi.Grid) == 6 * 4 + 3 * 8
    assert ct.sizeof(abi.Control) == 4 * 4 + 2 * 8
    assert ct.sizeof(abi.Info) == 4 + 4 + 8 + 4 + 4 + 4 * 8 + 8 + 8 + 8 + 4 + 4
    assert ct.sizeof(abi.ServerConfig) == 4 + 4 + 8
    assert ct.sizeof(abi.RequestInfo) == 8 + 4 * 4 + 4 * 8 + 8


def test_create_rejects_bad_arguments_without_touching_a_device():
    # validation runs before any device call: a non-positive transition thickness is an ArgumentError in the
    # reference (src/Models/smooth_ice_mask.jl:42)
    L = M.load()
    g = abi.default_grid(8, 8, 1.0, 1.0)
    p = abi.default_physics(mask_transition_thickness=0.0)
    h = ct.c_void_p()
    st = L.nk_create(ct.byref(g), ct.byref(p), ct.byref(abi.default_options()), ct.byref(h))
    assert st == abi.NK_INVALID_ARGUMENT and not h.value
    p = abi.default_physics(walls=7)
    st = L.nk_create(ct.byref(g), ct.byref(p), ct.byref(abi.default_options()), ct.byref(h))
    assert st == abi.NK_INVALID_ARGUMENT


def test_enum_mirrors_match_header():
    """Every NK_* enumerator include/nunatak_hip.h gives a value has the same value in the Python ABI mirror (named with
    or without the NK_ prefix), and every mirror of a header enum family is present."""
    text = open(HEADER).read()
    pairs = re.findall(r"\bNK_([A-Z0-9_]+)\s*=\s*(-?[0-9]+)", text)
    assert len(pairs) >= 40
    checked = 0
    for name, value in pairs:
        for attr in (name, "NK_" + name):
            if hasattr(abi, attr):
                assert getattr(abi, attr) == int(value), name
                checked += 1
                break
    assert checked >= 35
    for family in ("PROFILE_", "PRECONDITIONER_", "FRICTION_", "COARSE_", "CLASS_"):
        for name, _ in pairs:
            if name.startswith(family):
                assert hasattr(abi, name), name


def test_julia_binding_field_ids_match_header():
    """The Julia binding numbers the fields as consecutive constants; they must follow nk_field_id's order in
    include/nunatak_hip.h exactly (a shifted id would copy one model field into another)."""
    text = open(HEADER).read()
    body = text[text.index("typedef enum nk_field_id {"):text.index("} nk_field_id;")]
    header_ids = re.findall(r"^\s*NK_FIELD_([A-Z0-9_]+)\s*[,=]?", body, re.M)   # enumerators, not mentions in comments
    assert header_ids[-1] == "COUNT"
    header_ids = header_ids[:-1]
    jl = open(os.path.join(ROOT, "nunatak.jl_amd", "julia", "NunatakHIP.jl")).read()
    m = re.search(r"# field ids \(nk_field_id\)\nconst (.*?) = 0:(\d+)", jl, re.S)
    assert m, "field id block not found in NunatakHIP.jl"
    names = [n.strip() for n in m.group(1).replace("\n", " ").split(",")]
    assert names == header_ids
    assert int(m.group(2)) == len(header_ids) - 1
