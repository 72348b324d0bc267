"""`FLIP.apic` -- the reference's public simulation API (src/apic.rs) over libapic_hip.

Names, argument meaning and defaults follow the Rust module:
  Apic::new(width, height, cell_size, world_properties)      src/apic.rs:779
  Apic::update(&DebugOptions, frame_deltatime)                src/apic.rs:2429
  Apic::add_radial_force(x, y, radius, strength)              src/apic.rs:2352
  Apic::get_max_particle_velocity()                           src/apic.rs:2412
  Apic::is_fluid(i, j)                                        src/apic.rs:1015
  the pub stage methods                                       src/apic.rs:1043-2266
  WorldProperties {gravity, border_damping, cfl, solve_error} src/apic.rs:707-713
  DebugOptions {disable_particle_sort}                        src/apic.rs:397-400
The reference's pub fields are device-resident here; `read`/`write` copy a
field between HBM and a numpy array, and the particle views keep the
reference's flat chunk*8+lane order.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib


@dataclass
class WorldProperties:
    gravity: float = 0.0
    border_damping: float = 0.0
    cfl: float = 0.0
    solve_error: float = 0.0

    def _c(self):
        return _lib.WorldPropertiesC(self.gravity, self.border_damping, self.cfl, self.solve_error)


@dataclass
class DebugOptions:
    disable_particle_sort: bool = False


class Apic:
    """One simulation resident on the GPU."""

    def __init__(self, handle):
        self._h = handle

    # ---- construction ------------------------------------------------------
    @classmethod
    def new(cls, width, height, cell_size, world_properties):
        h = ctypes.c_void_p()
        check(lib().apic_create(width, height, cell_size, ctypes.byref(world_properties._c()), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def with_particles(cls, width, height, cell_size, world_properties, xy):
        """Apic::new with the particle set replaced by explicit positions (extension)."""
        xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float32).reshape(-1))
        h = ctypes.c_void_p()
        ptr = xy.ctypes.data if xy.size else None
        check(lib().apic_create_with_particles(width, height, cell_size, ctypes.byref(world_properties._c()),
                                               ptr, xy.size // 2, ctypes.byref(h)))
        return cls(h)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().apic_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- the step ----------------------------------------------------------
    def update(self, debug_options, frame_deltatime):
        opts = _lib.DebugOptionsC(1 if debug_options.disable_particle_sort else 0)
        check(lib().apic_update(self._h, ctypes.byref(opts), frame_deltatime))

    def add_radial_force(self, world_x, world_y, radius, strength):
        check(lib().apic_add_radial_force(self._h, world_x, world_y, radius, strength))

    def get_max_particle_velocity(self):
        v = ctypes.c_float()
        check(lib().apic_get_max_particle_velocity(self._h, ctypes.byref(v)))
        return v.value

    def is_fluid(self, i, j):
        v = ctypes.c_uint64()
        check(lib().apic_is_fluid(self._h, i, j, ctypes.byref(v)))
        return v.value

    def set_solver_config(self, mode=0, rel_tol=1e-6, max_iterations=500, fp64_reductions=True,
                          recompute_every=25):
        cfg = _lib.SolverConfigC(mode, rel_tol, max_iterations, int(fp64_reductions), recompute_every)
        check(lib().apic_set_solver_config(self._h, ctypes.byref(cfg)))

    def get_tuning(self):
        """The launch shapes (radix_bits, few_tiles) and the autotuner's timings."""
        t = _lib.TuningC()
        check(lib().apic_get_tuning(self._h, ctypes.byref(t)))
        return {"radix_bits": t.radix_bits, "few_tiles": t.few_tiles, "sort_ms": t.sort_ms, "vcycle_ms": t.vcycle_ms}

    def set_tuning(self, radix_bits=11, few_tiles=1024):
        """Sets the launch shapes by hand (they never change the result)."""
        t = _lib.TuningC(radix_bits, few_tiles, 0.0, 0.0)
        check(lib().apic_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, reps=5):
        """Times the candidate launch shapes on the current state, keeps the
        fastest and returns them (as get_tuning)."""
        t = _lib.TuningC()
        check(lib().apic_autotune(self._h, reps, ctypes.byref(t)))
        return {"radix_bits": t.radix_bits, "few_tiles": t.few_tiles, "sort_ms": t.sort_ms, "vcycle_ms": t.vcycle_ms}

    def set_gravity_x(self, gx):
        """Horizontal gravity added to u faces next to fluid (extension; 0 = the reference)."""
        check(lib().apic_set_gravity_x(self._h, gx))

    def device_bytes(self):
        """Device memory held by this simulation."""
        n = ctypes.c_size_t()
        check(lib().apic_device_bytes(self._h, ctypes.byref(n)))
        return n.value

    def stats(self):
        s = _lib.StepStatsC()
        check(lib().apic_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _lib.StepStatsC._fields_}

    def set_profiling(self, enable=True):
        """Record HIP events around every stage (True or 2), around the timed
        kernels only (1: P2G, G2P, SpMV), or not at all (False or 0).
        Extension; resets the timers."""
        level = 2 if enable is True else int(enable)
        check(lib().apic_set_profiling(self._h, level))

    def timers(self):
        """{timer: (total ms, count)} accumulated since profiling was enabled."""
        n = len(_lib.TIMERS)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_uint32 * n)()
        check(lib().apic_get_timers(self._h, ms, cnt, n))
        return {name: (ms[i], cnt[i]) for i, name in enumerate(_lib.TIMERS)}

    # ---- pub fields --------------------------------------------------------
    def dims(self):
        v = [ctypes.c_uint32() for _ in range(7)]
        check(lib().apic_get_dims(self._h, *[ctypes.byref(x) for x in v[:5]]))
        check(lib().apic_get_level_counts(self._h, None, ctypes.byref(v[6])))
        # num_levels: the reference's multigrid levels; num_levels_relative:
        # with the relative mode's deeper ones
        return dict(zip(("width", "height", "num_particles", "num_chunks", "num_levels", "num_levels_relative"),
                        [x.value for x in v[:5]] + [v[6].value]))

    @property
    def width(self):
        return self.dims()["width"]

    @property
    def height(self):
        return self.dims()["height"]

    @property
    def num_particles(self):
        return self.dims()["num_particles"]

    @property
    def num_chunks(self):
        return self.dims()["num_chunks"]

    @property
    def timestamp(self):
        ts = ctypes.c_uint32()
        pd = ctypes.c_float()
        check(lib().apic_get_timestamp(self._h, ctypes.byref(ts), ctypes.byref(pd)))
        return ts.value

    @property
    def prev_deltatime(self):
        ts = ctypes.c_uint32()
        pd = ctypes.c_float()
        check(lib().apic_get_timestamp(self._h, ctypes.byref(ts), ctypes.byref(pd)))
        return pd.value

    def set_timestamp(self, timestamp, prev_deltatime):
        check(lib().apic_set_timestamp(self._h, timestamp, prev_deltatime))

    def field_length(self, name, level=0):
        n = ctypes.c_size_t()
        check(lib().apic_field_length(self._h, _lib.FIELD_ID[name], level, ctypes.byref(n)))
        return n.value

    @staticmethod
    def _dtype(name):
        if name in _lib.F64_FIELDS:
            return np.float64
        return np.uint32 if name in _lib.U32_FIELDS else (np.uint8 if name in _lib.U8_FIELDS else np.float32)

    def read(self, name, level=0):
        n = self.field_length(name, level)
        out = np.zeros(n, dtype=self._dtype(name))
        check(lib().apic_read_field(self._h, _lib.FIELD_ID[name], level, out.ctypes.data if n else None, n))
        return out

    def write(self, name, values, level=0):
        n = self.field_length(name, level)
        arr = np.ascontiguousarray(np.asarray(values).astype(self._dtype(name)).reshape(-1))
        if arr.size != n:
            raise ValueError(f"{name}: expected {n} elements, got {arr.size}")
        check(lib().apic_write_field(self._h, _lib.FIELD_ID[name], level, arr.ctypes.data if n else None, n))

    # ---- the reference's pub particle vectors, in its chunk layout -------------
    AOSOA_GROUPS = {"part_positions": 0, "part_velocities": 1, "part_c_u": 2, "part_c_v": 3,
                    "part_time_residual": 4}

    def read_particles_aosoa(self, group):
        """Vec<Vec2x8> as (num_chunks, 2, 8) [x lanes then y lanes], or
        Vec<f32x8> as (num_chunks, 8) for part_time_residual (src/apic.rs:757-767)."""
        gid = self.AOSOA_GROUPS[group]
        nc = self.num_chunks
        shape = (nc, 8) if gid == 4 else (nc, 2, 8)
        out = np.zeros(shape, dtype=np.float32)
        check(lib().apic_read_particles_aosoa(self._h, gid, out.ctypes.data if out.size else None, out.size))
        return out

    def write_particles_aosoa(self, group, values):
        gid = self.AOSOA_GROUPS[group]
        nc = self.num_chunks
        shape = (nc, 8) if gid == 4 else (nc, 2, 8)
        arr = np.ascontiguousarray(np.asarray(values, dtype=np.float32).reshape(shape))
        check(lib().apic_write_particles_aosoa(self._h, gid, arr.ctypes.data if arr.size else None, arr.size))

    # ---- pub stage methods ---------------------------------------------------
    def update_spatial_lookup(self):
        check(lib().apic_update_spatial_lookup(self._h))

    def transfer_particles_to_grid(self, deltatime=0.0, timestamp=0):
        check(lib().apic_transfer_particles_to_grid(self._h, deltatime, timestamp))

    def apply_external_forces(self, deltatime):
        check(lib().apic_apply_external_forces(self._h, deltatime))

    def build_pressure_system(self):
        check(lib().apic_build_pressure_system(self._h))

    def update_multilevel_grid(self):
        check(lib().apic_update_multilevel_grid(self._h))

    def apply_preconditioner(self):
        check(lib().apic_apply_preconditioner(self._h))

    def solve_pcg(self):
        check(lib().apic_solve_pcg(self._h))

    def apply_pressure_gradient(self):
        check(lib().apic_apply_pressure_gradient(self._h))

    def extrapolate_velocity(self):
        check(lib().apic_extrapolate_velocity(self._h))

    def transfer_grid_to_particles_and_advect(self, deltatime):
        check(lib().apic_transfer_grid_to_particles_and_advect(self._h, deltatime))


class Batch:
    """Simulations on one stream, stepped together (extension; apic_batch_* of
    include/apic.h).  Each simulation of a batch is an ordinary `Apic`; a batch
    update of several of them computes, for each, exactly what `Apic.update`
    computes alone."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().apic_batch_create(ctypes.byref(h)))
        self._h = h

    def create_sim(self, width, height, cell_size, world_properties, xy=None):
        """Apic.new (xy None) or Apic.with_particles (xy: positions) inside the batch."""
        h = ctypes.c_void_p()
        if xy is None:
            check(lib().apic_batch_create_sim(self._h, width, height, cell_size,
                                              ctypes.byref(world_properties._c()), None, 0, ctypes.byref(h)))
        else:
            xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float32).reshape(-1))
            n = xy.size // 2
            if n == 0:
                xy = np.zeros(2, dtype=np.float32)  # a non-null pointer with n = 0: no particles
            ptr = ctypes.c_void_p(xy.ctypes.data)
            check(lib().apic_batch_create_sim(self._h, width, height, cell_size,
                                              ctypes.byref(world_properties._c()), ptr, n, ctypes.byref(h)))
        sim = Apic(h)
        sim._batch = self  # the batch outlives its simulations
        return sim

    def update(self, sims, debug_options, frame_deltatime, noalloc=False):
        """Apic.update(debug_options, dt) of every simulation in `sims`, together;
        frame_deltatime is one value or one per simulation.  noalloc: through
        apic_batch_update_noalloc (launch tables in the library's fixed
        device arena, no device allocation)."""
        n = len(sims)
        if n == 0:
            return
        dts = np.broadcast_to(np.asarray(frame_deltatime, dtype=np.float32), (n,)).copy()
        hs = (ctypes.c_void_p * n)(*[s._h.value for s in sims])
        opts = _lib.DebugOptionsC(1 if debug_options.disable_particle_sort else 0)
        fn = lib().apic_batch_update_noalloc if noalloc else lib().apic_batch_update
        check(fn(self._h, hs, n, ctypes.byref(opts), dts.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            check(lib().apic_batch_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
