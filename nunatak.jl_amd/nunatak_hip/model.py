"""ctypes binding of libnunatak_hip.so (include/nunatak_hip.h).

``HIPModel`` owns one solver handle: the device-resident fields of one grid and the reference operations on them
(update_geometry!, compute_driving_stress!, coefficient update, ssa_operator!, precondition!, solve!,
picard_solve!, update_velocities!). It raises when the library is missing or a call fails; there is no CPU
fallback.
"""
import ctypes as ct
import os

import numpy as np

from . import abi

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libnunatak_hip.so")

# every symbol declared in include/nunatak_hip.h
EXPORTS = ["nk_abi_version", "nk_default_grid", "nk_default_physics", "nk_default_options", "nk_create", "nk_destroy",
           "nk_last_error", "nk_set_physics", "nk_set_options", "nk_field_shape", "nk_set_field", "nk_get_field",
           "nk_fill_field", "nk_get_classes", "nk_field_device_pointer", "nk_stream", "nk_synchronize",
           "nk_update_geometry", "nk_compute_driving_stress", "nk_update_coefficients", "nk_apply_operator",
           "nk_precondition", "nk_solve", "nk_picard_solve", "nk_update_velocities", "nk_compute_velocity_profile",
           "nk_thermal_diagnostics", "nk_compute_thickness_tendency", "nk_thickness_substep", "nk_cache_thickness",
           "nk_cell_advection_timescale", "nk_compute_strain_heating", "nk_energy_step", "nk_default_till",
           "nk_till_effective_pressure",
           "nk_rate_factor_from_temperature", "nk_multigrid_levels", "nk_adjoint", "nk_device_bytes",
           "nk_estimate_device_bytes", "nk_server_create", "nk_server_destroy", "nk_server_last_error",
           "nk_server_submit", "nk_server_poll", "nk_server_wait", "nk_apply_operator_noalloc", "nk_set_thickness_bounds"]

_lib = None


class NunatakError(RuntimeError):
    """A call into libnunatak_hip.so failed; ``status`` holds its nk_status (None when no call was made)."""

    def __init__(self, msg, status=None):
        super().__init__(msg)
        self.status = status


def load(path=LIB_PATH):
    """Load libnunatak_hip.so and declare its signatures. Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NunatakError(f"{path} is not built; run __graft_entry__.build() (make -C nunatak.jl_amd)")
    L = ct.CDLL(path)
    P, I32, I64, D = ct.c_void_p, ct.c_int32, ct.c_int64, ct.c_double
    dp = ct.POINTER(ct.c_double)
    L.nk_abi_version.restype = I32
    L.nk_default_grid.argtypes = [ct.POINTER(abi.Grid)]
    L.nk_default_physics.argtypes = [ct.POINTER(abi.Physics)]
    L.nk_default_options.argtypes = [ct.POINTER(abi.SolverOptions)]
    L.nk_create.argtypes = [ct.POINTER(abi.Grid), ct.POINTER(abi.Physics), ct.POINTER(abi.SolverOptions), ct.POINTER(P)]
    L.nk_destroy.argtypes = [P]
    L.nk_last_error.argtypes = [P]
    L.nk_last_error.restype = ct.c_char_p
    L.nk_set_physics.argtypes = [P, ct.POINTER(abi.Physics)]
    L.nk_set_options.argtypes = [P, ct.POINTER(abi.SolverOptions)]
    L.nk_field_shape.argtypes = [P, ct.c_int, ct.POINTER(ct.c_int32)]
    L.nk_set_field.argtypes = [P, ct.c_int, dp, I64, I64]
    L.nk_get_field.argtypes = [P, ct.c_int, dp, I64, I64]
    L.nk_fill_field.argtypes = [P, ct.c_int, D]
    L.nk_get_classes.argtypes = [P, P, I64, P, I64]
    L.nk_field_device_pointer.argtypes = [P, ct.c_int, ct.POINTER(dp), ct.POINTER(I64), ct.POINTER(I64)]
    L.nk_stream.argtypes = [P]
    L.nk_stream.restype = P
    for name in ("nk_synchronize", "nk_update_geometry", "nk_compute_driving_stress", "nk_update_coefficients",
                 "nk_apply_operator", "nk_precondition", "nk_rate_factor_from_temperature", "nk_compute_velocity_profile",
                 "nk_thermal_diagnostics", "nk_cache_thickness", "nk_compute_strain_heating"):
        getattr(L, name).argtypes = [P]
    L.nk_solve.argtypes = [P, ct.POINTER(abi.Info)]
    L.nk_picard_solve.argtypes = [P, ct.POINTER(abi.Info)]
    L.nk_update_velocities.argtypes = [P, I32, ct.POINTER(abi.Info)]
    L.nk_multigrid_levels.argtypes = [P, ct.POINTER(I32), ct.POINTER(I32), I32]
    L.nk_adjoint.argtypes = [P, ct.POINTER(abi.Control), ct.POINTER(abi.Info)]
    L.nk_compute_thickness_tendency.argtypes = [P, I32, I32]
    L.nk_thickness_substep.argtypes = [P, D, I32]
    L.nk_set_thickness_bounds.argtypes = [P, D, D]
    L.nk_cell_advection_timescale.argtypes = [P, ct.POINTER(D)]
    L.nk_energy_step.argtypes = [P, D, I32, ct.POINTER(abi.Till)]
    L.nk_till_effective_pressure.argtypes = [P, ct.POINTER(abi.Till)]
    L.nk_default_till.argtypes = [ct.POINTER(abi.Till)]
    L.nk_default_till.restype = None
    L.nk_device_bytes.argtypes = [P, ct.POINTER(I64)]
    L.nk_estimate_device_bytes.argtypes = [ct.POINTER(abi.Grid), ct.POINTER(abi.Physics), ct.POINTER(abi.SolverOptions),
                                           ct.POINTER(I64)]
    L.nk_server_create.argtypes = [ct.POINTER(abi.ServerConfig), ct.POINTER(P)]
    L.nk_server_destroy.argtypes = [P]
    L.nk_server_last_error.argtypes = [P]
    L.nk_server_last_error.restype = ct.c_char_p
    L.nk_server_submit.argtypes = [P, ct.POINTER(abi.Grid), ct.POINTER(abi.Physics), ct.POINTER(abi.SolverOptions), I32,
                                   ct.POINTER(I32), ct.POINTER(dp), ct.POINTER(I64)]
    L.nk_server_poll.argtypes = [P, I64, ct.POINTER(I32)]
    L.nk_server_wait.argtypes = [P, I64, dp, dp, ct.POINTER(abi.RequestInfo)]
    L.nk_apply_operator_noalloc.argtypes = [ct.POINTER(abi.Grid), I32] + [P] * 8 + [I64, P]
    for name in EXPORTS:
        if name not in ("nk_last_error", "nk_stream", "nk_abi_version", "nk_default_grid", "nk_default_physics",
                        "nk_default_options", "nk_server_last_error"):
            getattr(L, name).restype = ct.c_int
    _lib = L
    return L


def _dptr(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


class HIPModel:
    """One solver handle and its device fields (field-id protocol of include/nunatak_hip.h)."""

    def __init__(self, grid, physics, options):
        self.grid, self.physics, self.options = grid, physics, options
        self._L = load()
        h = ct.c_void_p()
        st = self._L.nk_create(ct.byref(grid), ct.byref(physics), ct.byref(options), ct.byref(h))
        self._h = h
        if st != abi.NK_OK:
            msg = self._L.nk_last_error(h).decode() if h.value else ""
            if h.value:
                self._L.nk_destroy(h)
                self._h = ct.c_void_p()
            raise NunatakError(f"nk_create: {abi.STATUS_NAMES.get(st, st)} {msg}", st)

    def _check(self, st, what):
        if st != abi.NK_OK:
            raise NunatakError(f"{what}: {abi.STATUS_NAMES.get(st, st)}: {self._L.nk_last_error(self._h).decode()}", st)

    def close(self):
        if self._h and self._h.value:
            self._L.nk_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_physics(self, physics):
        self._check(self._L.nk_set_physics(self._h, ct.byref(physics)), "nk_set_physics")
        self.physics = physics

    def set_options(self, options):
        self._check(self._L.nk_set_options(self._h, ct.byref(options)), "nk_set_options")
        self.options = options

    def shape(self, name):
        s = (ct.c_int32 * 3)()
        self._check(self._L.nk_field_shape(self._h, abi.FIELD_ID[name], s), "nk_field_shape")
        return tuple(s)

    def set(self, name, value):
        nx, ny, nz = self.shape(name)
        a = np.empty((nz, ny, nx), dtype=np.float64)
        a[...] = value
        a = np.ascontiguousarray(a)
        self._check(self._L.nk_set_field(self._h, abi.FIELD_ID[name], _dptr(a), nx, nx * ny), f"nk_set_field({name})")

    def get(self, name):
        nx, ny, nz = self.shape(name)
        a = np.empty((nz, ny, nx), dtype=np.float64)
        self._check(self._L.nk_get_field(self._h, abi.FIELD_ID[name], _dptr(a), nx, nx * ny), f"nk_get_field({name})")
        return a[0] if nz == 1 else a

    def fill(self, name, value):
        self._check(self._L.nk_fill_field(self._h, abi.FIELD_ID[name], float(value)), f"nk_fill_field({name})")

    def classes(self):
        nxu, nyu, _ = self.shape("U")
        nxv, nyv, _ = self.shape("V")
        cu = np.zeros((nyu, nxu), dtype=np.uint8)
        cv = np.zeros((nyv, nxv), dtype=np.uint8)
        self._check(self._L.nk_get_classes(self._h, cu.ctypes.data, nxu, cv.ctypes.data, nxv), "nk_get_classes")
        return cu, cv

    def device_pointer(self, name):
        p = ct.POINTER(ct.c_double)()
        pitch, layer = ct.c_int64(), ct.c_int64()
        self._check(self._L.nk_field_device_pointer(self._h, abi.FIELD_ID[name], ct.byref(p), ct.byref(pitch),
                                                    ct.byref(layer)), "nk_field_device_pointer")
        return ct.cast(p, ct.c_void_p).value, pitch.value, layer.value

    def synchronize(self):
        self._check(self._L.nk_synchronize(self._h), "nk_synchronize")

    def update_geometry(self):
        self._check(self._L.nk_update_geometry(self._h), "nk_update_geometry")

    def compute_driving_stress(self):
        self._check(self._L.nk_compute_driving_stress(self._h), "nk_compute_driving_stress")

    def update_coefficients(self):
        self._check(self._L.nk_update_coefficients(self._h), "nk_update_coefficients")

    def apply_operator(self):
        self._check(self._L.nk_apply_operator(self._h), "nk_apply_operator")

    def precondition(self):
        self._check(self._L.nk_precondition(self._h), "nk_precondition")

    def compute_thickness_tendency(self, advection=abi.ADVECTION_UPWIND1, grounded_melt_coupling=False):
        self._check(self._L.nk_compute_thickness_tendency(self._h, advection, 1 if grounded_melt_coupling else 0),
                    "nk_compute_thickness_tendency")

    def set_thickness_bounds(self, lower, upper):
        self._check(self._L.nk_set_thickness_bounds(self._h, lower, upper), "nk_set_thickness_bounds")

    def thickness_substep(self, dtau, final_stage):
        self._check(self._L.nk_thickness_substep(self._h, dtau, 1 if final_stage else 0), "nk_thickness_substep")

    def cache_thickness(self):
        self._check(self._L.nk_cache_thickness(self._h), "nk_cache_thickness")

    def cell_advection_timescale(self):
        t = ct.c_double()
        self._check(self._L.nk_cell_advection_timescale(self._h, ct.byref(t)), "nk_cell_advection_timescale")
        return t.value

    def compute_strain_heating(self):
        self._check(self._L.nk_compute_strain_heating(self._h), "nk_compute_strain_heating")

    def energy_step(self, dt, grounded_melt_coupling=False, till=None):
        self._check(self._L.nk_energy_step(self._h, dt, 1 if grounded_melt_coupling else 0,
                                           ct.byref(till) if till is not None else None), "nk_energy_step")

    def till_effective_pressure(self, till):
        self._check(self._L.nk_till_effective_pressure(self._h, ct.byref(till)), "nk_till_effective_pressure")

    def thermal_diagnostics(self):
        self._check(self._L.nk_thermal_diagnostics(self._h), "nk_thermal_diagnostics")

    def rate_factor_from_temperature(self):
        self._check(self._L.nk_rate_factor_from_temperature(self._h), "nk_rate_factor_from_temperature")

    def solve(self):
        info = abi.Info()
        self._check(self._L.nk_solve(self._h, ct.byref(info)), "nk_solve")
        return info.as_dict()

    def picard_solve(self):
        info = abi.Info()
        self._check(self._L.nk_picard_solve(self._h, ct.byref(info)), "nk_picard_solve")
        return info.as_dict()

    def compute_velocity_profile(self):
        self._check(self._L.nk_compute_velocity_profile(self._h), "nk_compute_velocity_profile")

    def update_velocities(self, geometry=True):
        info = abi.Info()
        self._check(self._L.nk_update_velocities(self._h, 1 if geometry else 0, ct.byref(info)), "nk_update_velocities")
        return info.as_dict()

    def adjoint(self, control=None):
        """nk_adjoint: λ, ∂J/∂C, ∂J/∂A from the cost gradient in UBAR, VBAR at the converged (U, V)."""
        info = abi.Info()
        c = ct.byref(control) if control is not None else None
        self._check(self._L.nk_adjoint(self._h, c, ct.byref(info)), "nk_adjoint")
        return info.as_dict()

    def device_bytes(self):
        b = ct.c_int64()
        self._check(self._L.nk_device_bytes(self._h, ct.byref(b)), "nk_device_bytes")
        return b.value

    def multigrid_levels(self):
        n = ct.c_int32()
        buf = (ct.c_int32 * 128)()
        self._check(self._L.nk_multigrid_levels(self._h, ct.byref(n), buf, 64), "nk_multigrid_levels")
        return [(buf[2 * k], buf[2 * k + 1]) for k in range(n.value)]


def estimate_device_bytes(grid, physics, options):
    """nk_estimate_device_bytes: device memory nk_create would allocate for these arguments."""
    L = load()
    b = ct.c_int64()
    st = L.nk_estimate_device_bytes(ct.byref(grid), ct.byref(physics), ct.byref(options), ct.byref(b))
    if st != abi.NK_OK:
        raise NunatakError(f"nk_estimate_device_bytes: {abi.STATUS_NAMES.get(st, st)}")
    return b.value
