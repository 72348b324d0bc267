"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's cpu_baseline leg, as the checker and the CPU baseline.  The product
never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# Field ids and element types mirror include/apic.h.
FIELDS = [
    "POS_X", "POS_Y", "VEL_X", "VEL_Y", "C_U_X", "C_U_Y", "C_V_X", "C_V_Y", "TIME_RESIDUAL",
    "SORT_KEY", "SORT_VAL", "MAC_U", "MAC_V", "VALID_U", "VALID_V", "DENSITY",
    "SMOOTHED_DENSITY", "PRESSURE", "FLUID", "OBSTACLE", "OPEN_U", "OPEN_V",
    "DIAG", "INV_DIAG", "PLUS_X", "PLUS_Y", "RESIDUAL", "ERROR", "SEARCH",
    "MATRIX_TIMES_SEARCH", "EXT_FORCE_U", "EXT_FORCE_V", "POST_FORCE_U", "POST_FORCE_V",
    "PRESSURE_F64", "RHS",
]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
U32_FIELDS = {"SORT_KEY", "SORT_VAL"}
U8_FIELDS = {"VALID_U", "VALID_V", "FLUID", "OBSTACLE", "OPEN_U", "OPEN_V"}
PARTICLE_FIELDS = FIELDS[:9]

STAGES = {
    "update_spatial_lookup": 0, "transfer_particles_to_grid": 1, "apply_external_forces": 2,
    "build_pressure_system": 3, "update_multilevel_grid": 4, "apply_preconditioner": 5,
    "solve_pcg": 6, "apply_pressure_gradient": 7, "apply_post_forces": 8,
    "extrapolate_velocity": 9, "transfer_grid_to_particles_and_advect": 10, "compute_keys": 11,
    "prolongate_base": 12, "restrict_base": 13, "smooth_base_forward": 14, "smooth_base_reverse": 15,
    "solve_coarse_exact": 16,
    "solve_coarse_relative": 17,
    "restrict_level": 18, "prolongate_level": 19,   # level index in `timestamp`
}


class WorldProperties(ctypes.Structure):
    _fields_ = [("gravity", ctypes.c_float), ("border_damping", ctypes.c_float),
                ("cfl", ctypes.c_float), ("solve_error", ctypes.c_float)]


class StepStats(ctypes.Structure):
    _fields_ = [("substeps", ctypes.c_uint32), ("pcg_iterations", ctypes.c_uint32),
                ("pcg_iterations_total", ctypes.c_uint32), ("pcg_exit_reason", ctypes.c_uint32),
                ("pcg_max_residual", ctypes.c_float), ("pcg_rel_residual", ctypes.c_double),
                ("fluid_cells", ctypes.c_uint32), ("timestamp", ctypes.c_uint32),
                ("prev_deltatime", ctypes.c_float), ("max_particle_velocity", ctypes.c_float)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                 ctypes.POINTER(WorldProperties), vp, ctypes.c_uint32,
                                 ctypes.POINTER(ctypes.c_int)]
        L.orc_destroy.argtypes = [vp]
        L.orc_update.argtypes = [vp, ctypes.c_int, ctypes.c_float]
        L.orc_stage.argtypes = [vp, ctypes.c_int, ctypes.c_float, ctypes.c_uint32]
        L.orc_field_length.restype = ctypes.c_longlong
        L.orc_field_length.argtypes = [vp, ctypes.c_int, ctypes.c_uint32]
        L.orc_read_field.argtypes = [vp, ctypes.c_int, ctypes.c_uint32, vp]
        L.orc_write_field.argtypes = [vp, ctypes.c_int, ctypes.c_uint32, vp]
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.orc_get_dims.argtypes = [vp, u32p, u32p, u32p, u32p, u32p, u32p]
        L.orc_get_level_counts.argtypes = [vp, u32p, u32p]
        L.orc_get_timestamp.argtypes = [vp, u32p, ctypes.POINTER(ctypes.c_double)]
        L.orc_set_timestamp.argtypes = [vp, ctypes.c_uint32, ctypes.c_double]
        L.orc_get_stats.argtypes = [vp, ctypes.POINTER(StepStats)]
        L.orc_set_p2g_shuffle.argtypes = [vp, ctypes.c_uint32]
        L.orc_set_threads.argtypes = [vp, ctypes.c_int]
        L.orc_set_solver.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_set_gravity_x.argtypes = [vp, ctypes.c_double]
        L.orc_max_particle_velocity.restype = ctypes.c_double
        L.orc_max_particle_velocity.argtypes = [vp]
        L.orc_add_radial_force.argtypes = [vp] + [ctypes.c_float] * 4
        L.orc_morton.restype = ctypes.c_uint32
        L.orc_morton.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.orc_random_chunk.argtypes = [ctypes.c_uint32, ctypes.c_uint32, vp]
        L.orc_temporal_w.restype = ctypes.c_float
        L.orc_temporal_w.argtypes = [ctypes.c_float]
        L.orc_theta.restype = ctypes.c_float
        L.orc_theta.argtypes = [ctypes.c_float] * 3
        L.orc_dot.restype = ctypes.c_float
        L.orc_dot.argtypes = [vp, vp, ctypes.c_size_t]
        _LIB = L
    return _LIB


def bench_props(solve_error=1e-5):
    """WorldProperties of benches/fluid_benchmark.rs:17-22."""
    return WorldProperties(-9.81, 0.1, 5.0, solve_error)


class Oracle:
    """One CPU simulation; precision 32 (reference mode) or 64 (FP64 oracle)."""

    def __init__(self, width, height, cell_size, props=None, xy=None, precision=32):
        L = lib()
        self.precision = precision
        props = props or bench_props()
        st = ctypes.c_int(0)
        if xy is not None:
            xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float32).reshape(-1))
            self._h = L.orc_create(precision, width, height, cell_size, ctypes.byref(props),
                                   xy.ctypes.data, xy.size // 2, ctypes.byref(st))
        else:
            self._h = L.orc_create(precision, width, height, cell_size, ctypes.byref(props),
                                   None, 0, ctypes.byref(st))
        self.status = st.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    def dims(self):
        v = [ctypes.c_uint32() for _ in range(6)]
        lib().orc_get_dims(self._h, *[ctypes.byref(x) for x in v])
        keys = ("width", "height", "num_particles", "num_chunks", "num_levels", "fluid_cells")
        out = dict(zip(keys, [x.value for x in v]))
        ref, rel = ctypes.c_uint32(), ctypes.c_uint32()
        lib().orc_get_level_counts(self._h, ctypes.byref(ref), ctypes.byref(rel))
        out["num_levels_relative"] = rel.value  # with the relative mode's deeper levels
        return out

    def update(self, frame_dt, disable_sort=False):
        lib().orc_update(self._h, int(disable_sort), frame_dt)

    def stage(self, name, dt=0.0, timestamp=0):
        rc = lib().orc_stage(self._h, STAGES[name], dt, timestamp)
        if rc:
            raise RuntimeError(f"oracle stage {name} failed ({rc})")

    def field_length(self, name, level=0):
        return lib().orc_field_length(self._h, FIELD_ID[name], level)

    def read(self, name, level=0):
        n = self.field_length(name, level)
        if n < 0:
            raise KeyError(name)
        dt = np.uint32 if name in U32_FIELDS else (np.uint8 if name in U8_FIELDS else np.float64)
        out = np.zeros(n, dtype=dt)
        lib().orc_read_field(self._h, FIELD_ID[name], level, out.ctypes.data)
        return out

    def write(self, name, values, level=0):
        n = self.field_length(name, level)
        dt = np.uint32 if name in U32_FIELDS else (np.uint8 if name in U8_FIELDS else np.float64)
        arr = np.ascontiguousarray(np.asarray(values).astype(dt).reshape(-1))
        if arr.size != n:
            raise ValueError(f"{name}: expected {n} elements, got {arr.size}")
        lib().orc_write_field(self._h, FIELD_ID[name], level, arr.ctypes.data)

    def timestamp(self):
        ts = ctypes.c_uint32()
        pd = ctypes.c_double()
        lib().orc_get_timestamp(self._h, ctypes.byref(ts), ctypes.byref(pd))
        return ts.value, pd.value

    def set_solver(self, mode=1, rel_tol=1e-6, max_iterations=500, fp64=True, recompute_every=25):
        """Solver rules: mode 0 the reference's, mode 1 relative (FP64 dots)."""
        lib().orc_set_solver(self._h, mode, rel_tol, max_iterations, int(fp64), recompute_every)

    def set_timestamp(self, ts, prev_dt):
        lib().orc_set_timestamp(self._h, ts, prev_dt)

    def set_gravity_x(self, gx):
        """The library's horizontal-gravity extension (0 = the reference)."""
        lib().orc_set_gravity_x(self._h, gx)

    def stats(self):
        s = StepStats()
        lib().orc_get_stats(self._h, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in StepStats._fields_}

    def set_threads(self, threads):
        """Threaded mode (the CPU baseline): the reference's parallel loops on
        `threads` OpenMP threads; 1 (the default) is the deterministic checker."""
        lib().orc_set_threads(self._h, int(threads))

    def set_p2g_shuffle(self, seed):
        lib().orc_set_p2g_shuffle(self._h, seed)

    def max_particle_velocity(self):
        return lib().orc_max_particle_velocity(self._h)

    def add_radial_force(self, x, y, radius, strength):
        lib().orc_add_radial_force(self._h, x, y, radius, strength)


def morton(x, y):
    return lib().orc_morton(x, y)


def random_chunk(chunk, seed):
    out = np.zeros(8, dtype=np.float32)
    lib().orc_random_chunk(chunk, seed, out.ctypes.data)
    return out


def temporal_w(tau):
    return lib().orc_temporal_w(tau)


def theta(rho, rho_center, target=9.0):
    return lib().orc_theta(rho, rho_center, target)


def dot(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return lib().orc_dot(a.ctypes.data, b.ctypes.data, a.size)
