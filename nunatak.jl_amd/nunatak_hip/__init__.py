"""nunatak_hip — host side of the MI355X velocity solver for Nunatak.jl.

``abi`` mirrors the C ABI of libnunatak_hip.so (structs, enumerators, defaults of the reference's constructors);
``model`` binds it (``HIPModel``: one solver handle with the reference's operations — update_geometry!,
compute_driving_stress!, update_velocities!, solve!, picard_solve!, ssa_operator!, the adjoint); ``server`` binds the
serving entry points; ``transport`` runs time_step!'s thickness and enthalpy stages on the device; ``workloads``
generates the benchmark inputs.
"""
from . import abi  # noqa: F401
