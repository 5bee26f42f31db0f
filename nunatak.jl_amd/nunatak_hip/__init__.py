"""nunatak_hip — host side of the MI355X velocity solver for Nunatak.jl.

``abi`` mirrors the C ABI of libnunatak_hip.so; ``model`` binds it; ``api`` mirrors the reference's Julia
interface (IceSheetModel, StressBalanceSolver, solve!, picard_solve!, ssa_operator!, ...) over that binding.
"""
from . import abi  # noqa: F401
