"""Python binding of the CPU oracle (liboracle.so). TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker or the timed CPU
baseline. The product (nunatak_hip) never imports it.

``OracleModel`` exposes the same field-id protocol as ``nunatak_hip.model.HIPModel`` so a test can run the
same sequence of reference operations on both and compare field by field.
"""
import ctypes as ct
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "nunatak.jl_amd"))
from nunatak_hip import abi  # noqa: E402

LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ct.CDLL(LIB_PATH)
        P, I, D = ct.c_void_p, ct.c_int, ct.c_double
        dp = ct.POINTER(ct.c_double)
        L.nko_create.restype = P
        L.nko_create.argtypes = [ct.POINTER(abi.Grid), ct.POINTER(abi.Physics), ct.POINTER(abi.SolverOptions)]
        L.nko_destroy.argtypes = [P]
        L.nko_set_physics.argtypes = [P, ct.POINTER(abi.Physics)]
        L.nko_set_options.argtypes = [P, ct.POINTER(abi.SolverOptions)]
        L.nko_set_threads.argtypes = [I]
        L.nko_set_threads.restype = I
        L.nko_field_shape.argtypes = [P, I, ct.POINTER(ct.c_int32)]
        L.nko_set_field.argtypes = [P, I, dp, ct.c_int64, ct.c_int64]
        L.nko_get_field.argtypes = [P, I, dp, ct.c_int64, ct.c_int64]
        L.nko_get_classes.argtypes = [P, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_int64]
        for name in ["nko_update_geometry", "nko_compute_driving_stress", "nko_update_coefficients",
                     "nko_apply_operator", "nko_precondition", "nko_rate_factor_from_temperature",
                     "nko_thermal_diagnostics", "nko_cache_thickness", "nko_compute_strain_heating",
                     "nko_compute_velocity_profile"]:
            getattr(L, name).argtypes = [P]
        L.nko_solve.argtypes = [P, ct.POINTER(abi.Info)]
        L.nko_picard_solve.argtypes = [P, ct.POINTER(abi.Info)]
        L.nko_update_velocities.argtypes = [P, I, ct.POINTER(abi.Info)]
        L.nko_multigrid_levels.argtypes = [P, ct.POINTER(ct.c_int32), I]
        L.nko_multigrid_levels.restype = I
        L.nko_mg_transfer.argtypes = [P, I, dp, dp, dp, dp]
        L.nko_mg_transfer.restype = I
        L.nko_sigma_metrics.argtypes = [P, dp, dp, dp, dp]
        L.nko_enthalpy_conversions.argtypes = [D, D, D, dp, dp, dp]
        L.nko_thickness_tendency.argtypes = [P, I, I]
        L.nko_set_thickness_bounds.argtypes = [P, D, D]
        L.nko_thickness_substep.argtypes = [P, D, I]
        L.nko_cell_advection_timescale.argtypes = [P]
        L.nko_cell_advection_timescale.restype = D
        L.nko_energy_step.argtypes = [P, D, I, ct.c_void_p]
        L.nko_till_effective_pressure.argtypes = [P, ct.c_void_p]
        for name, args in [("nko_weno5z", [D] * 5), ("nko_weno3z", [D] * 3), ("nko_ice_mask_value", [D, D, D]), ("nko_segment_grounded_fraction", [D, D]),
                           ("nko_triangle_grounded_fraction", [D, D, D]), ("nko_arrhenius_rate_factor", [D, D]),
                           ("nko_basal_drag", [D, D, D, I, D, D])]:
            getattr(L, name).argtypes = args
            getattr(L, name).restype = D
        L.nko_grounding_line_stress.argtypes = [D] * 9 + [ct.POINTER(ct.c_int)]
        L.nko_grounding_line_stress.restype = D
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


class OracleModel:
    """CPU restatement of the reference's stress-balance state and operations (field-id protocol)."""

    def __init__(self, grid, physics, options):
        self.grid, self.physics, self.options = grid, physics, options
        self._h = lib().nko_create(ct.byref(grid), ct.byref(physics), ct.byref(options))

    def close(self):
        if self._h:
            lib().nko_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_physics(self, physics):
        self.physics = physics
        lib().nko_set_physics(self._h, ct.byref(physics))

    def set_options(self, options):
        self.options = options
        lib().nko_set_options(self._h, ct.byref(options))

    def shape(self, name):
        s = (ct.c_int32 * 3)()
        lib().nko_field_shape(self._h, abi.FIELD_ID[name], s)
        return tuple(s)

    def set(self, name, value):
        nx, ny, nz = self.shape(name)
        a = np.empty((nz, ny, nx), dtype=np.float64)
        a[...] = value
        a = np.ascontiguousarray(a)
        lib().nko_set_field(self._h, abi.FIELD_ID[name], _dptr(a), nx, nx * ny)

    def get(self, name):
        nx, ny, nz = self.shape(name)
        a = np.empty((nz, ny, nx), dtype=np.float64)
        lib().nko_get_field(self._h, abi.FIELD_ID[name], _dptr(a), nx, nx * ny)
        return a[0] if nz == 1 else a

    def classes(self):
        nxu, nyu, _ = self.shape("U")
        nxv, nyv, _ = self.shape("V")
        cu = np.zeros((nyu, nxu), dtype=np.uint8)
        cv = np.zeros((nyv, nxv), dtype=np.uint8)
        lib().nko_get_classes(self._h, cu.ctypes.data, nxu, cv.ctypes.data, nxv)
        return cu, cv

    def update_geometry(self):
        lib().nko_update_geometry(self._h)

    def compute_driving_stress(self):
        lib().nko_compute_driving_stress(self._h)

    def update_coefficients(self):
        lib().nko_update_coefficients(self._h)

    def apply_operator(self):
        lib().nko_apply_operator(self._h)

    def precondition(self):
        lib().nko_precondition(self._h)

    def rate_factor_from_temperature(self):
        lib().nko_rate_factor_from_temperature(self._h)

    def thermal_diagnostics(self):
        lib().nko_thermal_diagnostics(self._h)

    def compute_thickness_tendency(self, advection=0, grounded_melt_coupling=False):
        lib().nko_thickness_tendency(self._h, advection, 1 if grounded_melt_coupling else 0)

    def set_thickness_bounds(self, lower, upper):
        lib().nko_set_thickness_bounds(self._h, lower, upper)

    def thickness_substep(self, dtau, final_stage):
        lib().nko_thickness_substep(self._h, dtau, 1 if final_stage else 0)

    def cache_thickness(self):
        lib().nko_cache_thickness(self._h)

    def cell_advection_timescale(self):
        return lib().nko_cell_advection_timescale(self._h)

    def compute_strain_heating(self):
        lib().nko_compute_strain_heating(self._h)

    def energy_step(self, dt, grounded_melt_coupling=False, till=None):
        lib().nko_energy_step(self._h, dt, 1 if grounded_melt_coupling else 0, ct.byref(till) if till is not None else None)

    def till_effective_pressure(self, till):
        lib().nko_till_effective_pressure(self._h, ct.byref(till))

    def solve(self):
        info = abi.Info()
        lib().nko_solve(self._h, ct.byref(info))
        return info.as_dict()

    def picard_solve(self):
        info = abi.Info()
        lib().nko_picard_solve(self._h, ct.byref(info))
        return info.as_dict()

    def compute_velocity_profile(self):
        lib().nko_compute_velocity_profile(self._h)

    def update_velocities(self, geometry=True):
        info = abi.Info()
        lib().nko_update_velocities(self._h, 1 if geometry else 0, ct.byref(info))
        return info.as_dict()

    def multigrid_levels(self):
        buf = (ct.c_int32 * 128)()
        n = lib().nko_multigrid_levels(self._h, buf, 64)
        return [(buf[2 * k], buf[2 * k + 1]) for k in range(n)]

    def mg_transfer(self, direction, in_u, in_v, coarse_shape_u, coarse_shape_v):
        """direction 0: restrict fine (in) to the first coarse level; 1: prolongate coarse (in) to fine."""
        in_u = np.ascontiguousarray(in_u, dtype=np.float64)
        in_v = np.ascontiguousarray(in_v, dtype=np.float64)
        if direction == 0:
            out_u, out_v = np.zeros(coarse_shape_u), np.zeros(coarse_shape_v)
        else:
            out_u = np.zeros(self.shape("U")[1::-1])
            out_v = np.zeros(self.shape("V")[1::-1])
        rc = lib().nko_mg_transfer(self._h, direction, _dptr(in_u), _dptr(in_v), _dptr(out_u), _dptr(out_v))
        assert rc == 0
        return out_u, out_v

    def sigma_metrics(self):
        Ns = self.grid.Nsigma
        zf, zc, dz, w2 = (np.zeros(Ns + 1), np.zeros(Ns), np.zeros(Ns), np.zeros(Ns))
        lib().nko_sigma_metrics(self._h, _dptr(zf), _dptr(zc), _dptr(dz), _dptr(w2))
        return zf, zc, dz, w2


def set_threads(n):
    """Threads of the oracle's parallel row loops (0: leave as is); returns the count in effect."""
    return lib().nko_set_threads(int(n))


def weno5z(a, b, c, d, e):
    """Fifth-order WENO-Z value at the right face of cell c from cell averages (a, b, c, d, e)."""
    return lib().nko_weno5z(a, b, c, d, e)


def weno3z(a, b, c):
    return lib().nko_weno3z(a, b, c)


def ice_mask_value(H, hmin, htr):
    return lib().nko_ice_mask_value(H, hmin, htr)


def segment_grounded_fraction(f1, f2):
    return lib().nko_segment_grounded_fraction(f1, f2)


def triangle_grounded_fraction(f1, f2, f3):
    return lib().nko_triangle_grounded_fraction(f1, f2, f3)


def arrhenius_rate_factor(T, omega):
    return lib().nko_arrhenius_rate_factor(T, omega)


def enthalpy_conversions(T, omega, p):
    """(E = enthalpy(T, ω, p), temperature(E, p), water_fraction(E, p))"""
    E, Tb, wb = ct.c_double(), ct.c_double(), ct.c_double()
    lib().nko_enthalpy_conversions(T, omega, p, ct.byref(E), ct.byref(Tb), ct.byref(wb))
    return E.value, Tb.value, wb.value


def basal_drag(U, C, N, law, m, alpha2):
    return lib().nko_basal_drag(U, C, N, law, m, alpha2)


def grounding_line_stress(H1, H2, s1, s2, b1, b2, ri, rw, g):
    st = ct.c_int(0)
    t = lib().nko_grounding_line_stress(H1, H2, s1, s2, b1, b2, ri, rw, g, ct.byref(st))
    return bool(st.value), t
