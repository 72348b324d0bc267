"""Loader and ctypes signatures of libapic_hip.so (include/apic.h)."""
import ctypes
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libapic_hip.so")
_LIB = None

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
F64_FIELDS = {"PRESSURE_F64"}

STATUS = {0: "APIC_OK", 1: "APIC_ERR_INVALID_ARGUMENT", 2: "APIC_ERR_UNSUPPORTED_SIZE",
          3: "APIC_ERR_OUT_OF_MEMORY", 4: "APIC_ERR_HIP", 5: "APIC_ERR_NO_DEVICE"}

# every entry point of include/apic.h
SYMBOLS = [
    "apic_create", "apic_create_with_particles", "apic_destroy", "apic_last_error", "apic_update",
    "apic_add_radial_force", "apic_get_max_particle_velocity", "apic_is_fluid",
    "apic_set_solver_config", "apic_get_stats", "apic_get_dims", "apic_get_level_counts", "apic_get_timestamp",
    "apic_get_tuning", "apic_set_tuning", "apic_autotune",
    "apic_set_timestamp", "apic_field_length", "apic_read_field", "apic_write_field",
    "apic_read_particles_aosoa", "apic_write_particles_aosoa",
    "apic_update_spatial_lookup", "apic_transfer_particles_to_grid", "apic_apply_external_forces",
    "apic_build_pressure_system", "apic_update_multilevel_grid", "apic_apply_preconditioner",
    "apic_solve_pcg", "apic_apply_pressure_gradient", "apic_extrapolate_velocity",
    "apic_transfer_grid_to_particles_and_advect", "apic_set_profiling", "apic_get_timers",
    "apic_set_gravity_x", "apic_device_bytes", "apic_batch_create", "apic_batch_destroy",
    "apic_batch_create_sim", "apic_batch_update", "apic_device_memory", "apic_batch_update_noalloc",
]
TIMERS = ["sort", "p2g", "pressure_build", "mg_build", "pcg", "gradient", "extrapolate", "g2p",
          "kernel_p2g", "kernel_g2p", "kernel_spmv", "kernel_pcg_iteration"]


class WorldPropertiesC(ctypes.Structure):
    _fields_ = [("gravity", ctypes.c_float), ("border_damping", ctypes.c_float),
                ("cfl", ctypes.c_float), ("solve_error", ctypes.c_float)]


class DebugOptionsC(ctypes.Structure):
    _fields_ = [("disable_particle_sort", ctypes.c_uint8)]


class TuningC(ctypes.Structure):
    _fields_ = [("radix_bits", ctypes.c_uint32), ("few_tiles", ctypes.c_uint32),
                ("sort_ms", ctypes.c_float), ("vcycle_ms", ctypes.c_float)]


class SolverConfigC(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("rel_tol", ctypes.c_float), ("max_iterations", ctypes.c_int32),
                ("fp64_reductions", ctypes.c_int32), ("recompute_every", ctypes.c_int32)]


class StepStatsC(ctypes.Structure):
    _fields_ = [("substeps", ctypes.c_uint32), ("pcg_iterations", ctypes.c_uint32),
                ("pcg_iterations_total", ctypes.c_uint32), ("pcg_exit_reason", ctypes.c_uint32),
                ("pcg_max_residual", ctypes.c_float), ("pcg_rel_residual", ctypes.c_double),
                ("fluid_cells", ctypes.c_uint32), ("timestamp", ctypes.c_uint32),
                ("prev_deltatime", ctypes.c_float), ("max_particle_velocity", ctypes.c_float)]


class ApicError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


def lib():
    """Load libapic_hip.so; raise if it has not been built (there is no fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, f32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float, ctypes.c_int32
    u32p = ctypes.POINTER(u32)
    L.apic_create.argtypes = [u32, u32, f32, ctypes.POINTER(WorldPropertiesC), ctypes.POINTER(vp)]
    L.apic_create_with_particles.argtypes = [u32, u32, f32, ctypes.POINTER(WorldPropertiesC), vp, u32,
                                             ctypes.POINTER(vp)]
    L.apic_destroy.argtypes = [vp]
    L.apic_destroy.restype = None
    L.apic_last_error.restype = ctypes.c_char_p
    L.apic_last_error.argtypes = []
    L.apic_update.argtypes = [vp, ctypes.POINTER(DebugOptionsC), f32]
    L.apic_add_radial_force.argtypes = [vp, f32, f32, f32, f32]
    L.apic_get_max_particle_velocity.argtypes = [vp, ctypes.POINTER(f32)]
    L.apic_is_fluid.argtypes = [vp, i32, i32, ctypes.POINTER(ctypes.c_uint64)]
    L.apic_set_solver_config.argtypes = [vp, ctypes.POINTER(SolverConfigC)]
    L.apic_get_tuning.argtypes = [vp, ctypes.POINTER(TuningC)]
    L.apic_set_tuning.argtypes = [vp, ctypes.POINTER(TuningC)]
    L.apic_autotune.argtypes = [vp, u32, ctypes.POINTER(TuningC)]
    L.apic_get_stats.argtypes = [vp, ctypes.POINTER(StepStatsC)]
    L.apic_get_dims.argtypes = [vp, u32p, u32p, u32p, u32p, u32p]
    L.apic_get_level_counts.argtypes = [vp, u32p, u32p]
    L.apic_get_timestamp.argtypes = [vp, u32p, ctypes.POINTER(f32)]
    L.apic_set_timestamp.argtypes = [vp, u32, f32]
    L.apic_field_length.argtypes = [vp, ctypes.c_int, u32, ctypes.POINTER(ctypes.c_size_t)]
    L.apic_read_field.argtypes = [vp, ctypes.c_int, u32, vp, ctypes.c_size_t]
    L.apic_write_field.argtypes = [vp, ctypes.c_int, u32, vp, ctypes.c_size_t]
    L.apic_read_particles_aosoa.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t]
    L.apic_write_particles_aosoa.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t]
    for name in ("apic_update_spatial_lookup", "apic_build_pressure_system", "apic_update_multilevel_grid",
                 "apic_apply_preconditioner", "apic_solve_pcg", "apic_apply_pressure_gradient",
                 "apic_extrapolate_velocity"):
        getattr(L, name).argtypes = [vp]
    L.apic_transfer_particles_to_grid.argtypes = [vp, f32, u32]
    L.apic_apply_external_forces.argtypes = [vp, f32]
    L.apic_transfer_grid_to_particles_and_advect.argtypes = [vp, f32]
    L.apic_set_profiling.argtypes = [vp, ctypes.c_int]
    L.apic_get_timers.argtypes = [vp, ctypes.POINTER(ctypes.c_double), u32p, u32]
    L.apic_set_gravity_x.argtypes = [vp, f32]
    L.apic_device_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
    L.apic_batch_create.argtypes = [ctypes.POINTER(vp)]
    L.apic_device_memory.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    L.apic_batch_destroy.argtypes = [vp]
    L.apic_batch_create_sim.argtypes = [vp, u32, u32, f32, ctypes.POINTER(WorldPropertiesC), vp, u32,
                                        ctypes.POINTER(vp)]
    L.apic_batch_update.argtypes = [vp, ctypes.POINTER(vp), u32, ctypes.POINTER(DebugOptionsC),
                                    ctypes.POINTER(f32)]
    L.apic_batch_update_noalloc.argtypes = L.apic_batch_update.argtypes
    for name in SYMBOLS:
        fn = getattr(L, name)
        if name not in ("apic_destroy", "apic_last_error"):
            fn.restype = ctypes.c_int
    _LIB = L
    return L


def check(status):
    if status != 0:
        raise ApicError(status, lib().apic_last_error().decode(errors="replace"))
