"""ctypes mirror of include/nunatak_hip.h: structs, enums and field locations.

The structs are shared by the HIP library binding (``nunatak_hip.model``) and by any other C library that
speaks the same field-id protocol.
"""
import ctypes as ct

NK_OK, NK_INVALID_ARGUMENT, NK_NOT_CONVERGED, NK_NONFINITE, NK_OUT_OF_MEMORY, NK_DEVICE_ERROR, NK_UNSUPPORTED = range(7)
STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "NOT_CONVERGED", 3: "NONFINITE", 4: "OUT_OF_MEMORY",
                5: "DEVICE_ERROR", 6: "UNSUPPORTED"}

PERIODIC, BOUNDED = 0, 1
GROUNDING_NONE, GROUNDING_SUBGRID, GROUNDING_SUBGRID_DRIVING_STRESS = 0, 1, 2
FRICTION_NONE, FRICTION_WEERTMAN, FRICTION_REGULARIZED_COULOMB = 0, 1, 2
FREE_SLIP, NO_SLIP = 0, 1
SSA, DIVA = 0, 1
RATE_FACTOR_2D, RATE_FACTOR_3D = 0, 1
CONVERGENCE, FIXED_ITERATIONS = 0, 1
PRECONDITIONER_NONE, PRECONDITIONER_DIAGONAL, PRECONDITIONER_MULTIGRID = 0, 1, 2
POINT_JACOBI, LINE_JACOBI = 0, 1
COARSE_REDISCRETIZED, COARSE_GALERKIN = 0, 1
PRECISION_FP64, PRECISION_MIXED, PRECISION_MIXED_REFINED = 0, 1, 2

CLASS_WALL, CLASS_ICE_FREE, CLASS_FRONT, CLASS_GROUNDING = 1, 2, 4, 8
PROFILE_NONE, PROFILE_OPERATOR, PROFILE_SWEEP, PROFILE_UPDATE, PROFILE_GALERKIN_SWEEP = 0, 1, 2, 3, 4

FIELDS = ["U", "V", "H", "BED", "S", "PHI", "PHIU", "PHIV", "N", "MU", "TAUU", "TAUV", "ETAH", "ETAHFF",
          "BETAU", "BETAV", "F2", "ETASIGMA", "BETABU", "BETABV", "UB", "VB", "USIGMA", "VSIGMA", "A", "C",
          "TEMPERATURE", "PU", "PV", "QU", "QV", "UBAR", "VBAR", "LAMBDAU", "LAMBDAV", "CBAR", "ABAR", "USIGMABAR",
          "VSIGMABAR", "ENTHALPY", "WATER_FRACTION", "SMB", "BASAL_MELT", "GROUNDED_MELT", "HPREV", "GH", "HCLIP",
          "HW", "HN", "GEOTHERMAL", "TSURF", "TOCEAN", "STRAIN_HEATING", "HBAR", "BEDBAR"]
FIELD_ID = {name: k for k, name in enumerate(FIELDS)}

# (x location, y location, on the σ grid); "f" = Face, "c" = Center.
LOCATIONS = {
    "U": ("f", "c", False), "V": ("c", "f", False), "H": ("c", "c", False), "BED": ("c", "c", False),
    "S": ("c", "c", False), "PHI": ("c", "c", False), "PHIU": ("f", "c", False), "PHIV": ("c", "f", False),
    "N": ("c", "c", False), "MU": ("c", "c", False), "TAUU": ("f", "c", False), "TAUV": ("c", "f", False),
    "ETAH": ("c", "c", False), "ETAHFF": ("f", "f", False), "BETAU": ("f", "c", False),
    "BETAV": ("c", "f", False), "F2": ("c", "c", False), "ETASIGMA": ("c", "c", True),
    "BETABU": ("f", "c", False), "BETABV": ("c", "f", False), "UB": ("f", "c", False), "VB": ("c", "f", False),
    "USIGMA": ("f", "c", True), "VSIGMA": ("c", "f", True), "A": ("c", "c", None), "C": ("c", "c", False),
    "TEMPERATURE": ("c", "c", True), "PU": ("f", "c", False), "PV": ("c", "f", False),
    "QU": ("f", "c", False), "QV": ("c", "f", False), "UBAR": ("f", "c", False), "VBAR": ("c", "f", False),
    "LAMBDAU": ("f", "c", False), "LAMBDAV": ("c", "f", False), "CBAR": ("c", "c", False), "ABAR": ("c", "c", None),
    "USIGMABAR": ("f", "c", True), "VSIGMABAR": ("c", "f", True), "ENTHALPY": ("c", "c", True),
    "WATER_FRACTION": ("c", "c", True), "SMB": ("c", "c", False), "BASAL_MELT": ("c", "c", False),
    "GROUNDED_MELT": ("c", "c", False), "HPREV": ("c", "c", False), "GH": ("c", "c", False), "HCLIP": ("c", "c", False),
    "HW": ("c", "c", False), "HN": ("c", "c", False), "GEOTHERMAL": ("c", "c", False), "TSURF": ("c", "c", False),
    "TOCEAN": ("c", "c", False), "STRAIN_HEATING": ("c", "c", True),
    "HBAR": ("c", "c", False), "BEDBAR": ("c", "c", False),
}
ADVECTION_UPWIND1, ADVECTION_CENTERED2, ADVECTION_WENO5, ADVECTION_WENO5_BOUNDED = 0, 1, 2, 3


class Grid(ct.Structure):
    _fields_ = [("Nx", ct.c_int32), ("Ny", ct.c_int32), ("topology_x", ct.c_int32), ("topology_y", ct.c_int32),
                ("halo", ct.c_int32), ("Nsigma", ct.c_int32), ("dx", ct.c_double), ("dy", ct.c_double),
                ("bed_refinement", ct.c_double)]


class Physics(ct.Structure):
    _fields_ = [("ice_density", ct.c_double), ("ocean_density", ct.c_double),
                ("gravitational_acceleration", ct.c_double), ("grounding", ct.c_int32), ("masking", ct.c_int32),
                ("mask_minimum_thickness", ct.c_double), ("mask_transition_thickness", ct.c_double),
                ("glen_exponent", ct.c_double), ("strain_rate_regularization", ct.c_double),
                ("rate_factor_depth", ct.c_int32), ("friction", ct.c_int32), ("friction_exponent", ct.c_double),
                ("coulomb_coefficient", ct.c_double), ("velocity_regularization", ct.c_double),
                ("walls", ct.c_int32), ("stress_balance", ct.c_int32)]


class Control(ct.Structure):
    _fields_ = [("kind", ct.c_int32), ("maxiter", ct.c_int32), ("iterations", ct.c_int32), ("pad", ct.c_int32),
                ("reltol", ct.c_double), ("abstol", ct.c_double)]


class SolverOptions(ct.Structure):
    _fields_ = [("linear", Control), ("picard", Control), ("preconditioner", ct.c_int32),
                ("smoother", ct.c_int32), ("smoothing_sweeps", ct.c_int32), ("coarsest_sweeps", ct.c_int32),
                ("minimum_coarse_size", ct.c_int32), ("precision", ct.c_int32), ("coarse_operator", ct.c_int32),
                ("check_interval", ct.c_int32), ("profile_kernel", ct.c_int32), ("graphs", ct.c_int32),
                ("report_not_converged", ct.c_int32)]


class Info(ct.Structure):
    _fields_ = [("picard_iterations", ct.c_int32), ("linear_iterations", ct.c_int32),
                ("total_linear_iterations", ct.c_int64), ("refinement_steps", ct.c_int32), ("pad", ct.c_int32),
                ("last_picard_change", ct.c_double), ("last_picard_norm", ct.c_double),
                ("last_residual_norm", ct.c_double), ("elapsed_ms", ct.c_double), ("profiled_ms", ct.c_double),
                ("profiled_launches", ct.c_int64), ("profiled_points", ct.c_int64), ("profiled_precision", ct.c_int32),
                ("pad2", ct.c_int32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if not name.startswith("pad")}


class Till(ct.Structure):
    _fields_ = [("reference_effective_pressure", ct.c_double), ("overburden_fraction", ct.c_double),
                ("reference_void_ratio", ct.c_double), ("compressibility", ct.c_double),
                ("maximum_water_thickness", ct.c_double), ("drainage_rate", ct.c_double)]


def default_till(**kw):
    """TillEffectivePressure defaults (src/Thermodynamics/till_effective_pressure.jl:47-61), overridable by keyword."""
    t = Till(1000.0, 0.02, 0.69, 0.12, 2.0, 1e-3 / (365.25 * 86400.0))
    for k, v in kw.items():
        setattr(t, k, v)
    return t


class ServerConfig(ct.Structure):
    _fields_ = [("slots", ct.c_int32), ("pad", ct.c_int32), ("hbm_budget_bytes", ct.c_int64)]


class RequestInfo(ct.Structure):
    _fields_ = [("id", ct.c_int64), ("status", ct.c_int32), ("reused_handle", ct.c_int32),
                ("picard_iterations", ct.c_int32), ("total_linear_iterations", ct.c_int32),
                ("submit_ms", ct.c_double), ("admit_ms", ct.c_double), ("start_ms", ct.c_double),
                ("finish_ms", ct.c_double), ("device_bytes", ct.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def default_grid(Nx, Ny, dx, dy, topology=(BOUNDED, BOUNDED), halo=4, Nsigma=10, bed_refinement=2.0):
    return Grid(Nx, Ny, topology[0], topology[1], halo, Nsigma, dx, dy, bed_refinement)


def default_physics(**kw):
    """IceSheetConstants / SmoothIceMask / GlenFlowLaw / friction / DepthIntegrated defaults of the reference."""
    p = Physics(ice_density=918.0, ocean_density=1028.0, gravitational_acceleration=9.81,
                grounding=GROUNDING_NONE, masking=1, mask_minimum_thickness=1.0, mask_transition_thickness=1.0,
                glen_exponent=3.0, strain_rate_regularization=1e-13, rate_factor_depth=RATE_FACTOR_2D,
                friction=FRICTION_NONE, friction_exponent=1.0 / 3.0, coulomb_coefficient=0.0,
                velocity_regularization=1e-8, walls=FREE_SLIP, stress_balance=SSA)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def convergence(reltol=1.4901161193847656e-08, abstol=0.0, maxiter=1000):
    return Control(kind=CONVERGENCE, maxiter=maxiter, iterations=0, pad=0, reltol=reltol, abstol=abstol)


def fixed_iterations(n):
    return Control(kind=FIXED_ITERATIONS, maxiter=n, iterations=n, pad=0, reltol=0.0, abstol=0.0)


def default_options(**kw):
    """StressBalanceSolver and StressBalanceMultigridPreconditioner defaults of the reference."""
    o = SolverOptions(linear=convergence(1e-8, 0.0, 10_000), picard=convergence(1e-6, 0.0, 100),
                      preconditioner=PRECONDITIONER_DIAGONAL, smoother=POINT_JACOBI, smoothing_sweeps=2,
                      coarsest_sweeps=20, minimum_coarse_size=2, precision=PRECISION_FP64,
                      coarse_operator=COARSE_REDISCRETIZED, check_interval=8, profile_kernel=0, graphs=1, report_not_converged=0)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def field_shape(grid, name):
    """(nx, ny, nz) of a field: Face on a Bounded direction has N + 1 points."""
    lx, ly, sig = LOCATIONS[name]
    nx = grid.Nx + (1 if (lx == "f" and grid.topology_x == BOUNDED) else 0)
    ny = grid.Ny + (1 if (ly == "f" and grid.topology_y == BOUNDED) else 0)
    return nx, ny, sig


def packed_shape(grid, physics, name):
    """(nx, ny, nz) of a field as nk_field_shape reports it: σ fields have Nsigma layers, the rate factor (and its
    gradient) Nsigma layers when physics.rate_factor_depth is 3D."""
    nx, ny, sig = field_shape(grid, name)
    if sig is None:
        nz = grid.Nsigma if physics.rate_factor_depth == RATE_FACTOR_3D else 1
    else:
        nz = grid.Nsigma if sig else 1
    return nx, ny, nz


def node_coordinates(grid, name, x0=0.0, y0=0.0):
    """x, y node coordinates of a field's points (Oceananigans xnodes/ynodes), shaped (ny, nx)."""
    import numpy as np
    lx, ly, _ = LOCATIONS[name]
    nx, ny, _ = field_shape(grid, name)
    xi = np.arange(nx, dtype=np.float64)
    yj = np.arange(ny, dtype=np.float64)
    x = x0 + (xi if lx == "f" else xi + 0.5) * grid.dx
    y = y0 + (yj if ly == "f" else yj + 0.5) * grid.dy
    return np.meshgrid(x, y, indexing="xy")
