"""Continuous-batching server over a resident pool of simulations (extension:
the reference steps one simulation per `Apic::update` call, src/apic.rs:2429).

A request asks for `steps` updates of one simulation of the pool.  The server
keeps every admitted simulation's state in HBM between requests and runs in
ticks: each tick admits waiting requests (first come, first served) while the
device-memory budget allows, then advances every simulation with an active
request by one update in a single batched update (`apic_batch_update`), so a
request that arrives while others run joins them at the next tick.  When a
simulation must be created and the budget is full, idle simulations (no
active or waiting request) are offloaded, least recently used first: their
particle state and step scalars go to host memory and come back, unchanged,
when they are admitted again.  A batched update computes for each simulation
exactly what `Apic.update` computes alone, so a served simulation's state
after n steps is that of n plain updates.
"""
import collections
import ctypes
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check, lib
from .apic import Apic, Batch, DebugOptions, WorldProperties

PARTICLE_FIELDS = ["POS_X", "POS_Y", "VEL_X", "VEL_Y", "C_U_X", "C_U_Y", "C_V_X", "C_V_Y", "TIME_RESIDUAL"]


def device_memory():
    """(free, total) bytes of the current HIP device."""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    check(lib().apic_device_memory(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


@dataclass
class SimSpec:
    """Construction of one simulation of the pool (Apic.with_particles plus the
    extensions it uses)."""
    width: int
    height: int
    dx: float
    xy: np.ndarray
    gravity: float = -9.81
    gravity_x: float = 0.0
    border_damping: float = 0.1
    cfl: float = 5.0
    solve_error: float = 1e-5
    solver: dict = field(default_factory=dict)  # Apic.set_solver_config arguments ({}: the reference's rules)

    def estimate_bytes(self):
        """Upper estimate of the device memory the simulation will hold."""
        n = int(self.xy.shape[0])
        nalloc = (n + 7) // 8 * 8
        side = 1 << max(int(self.width) - 1, int(self.height) - 1, 1).bit_length()
        return 72 * nalloc + 24 * n + 8 * side * side + 200 * self.width * self.height + (1 << 20)


@dataclass
class Request:
    rid: int
    sim: int
    steps: int
    t_arrival: float
    remaining: int = 0
    t_start: float = None
    t_done: float = None


class Server:
    """Resident pool + continuous batching (see the module docstring)."""

    def __init__(self, specs, hbm_budget=None, frame_dt=0.002, debug_options=None):
        self.specs = specs
        free, _ = device_memory()
        # leave room for the batch's launch tables and the runtime
        self.budget = int(min(hbm_budget, free) if hbm_budget else free * 0.85)
        self.frame_dt = frame_dt
        self.opts = debug_options or DebugOptions()
        self.batch = Batch()
        self.resident = collections.OrderedDict()  # sim id -> Apic, least recently stepped first
        self.sim_bytes = {}
        self.offloaded = {}                        # sim id -> host snapshot
        self.used = 0
        self.waiting = collections.deque()
        self.active = {}                           # sim id -> Request
        self.done = []
        self.steps_of = collections.Counter()      # updates applied per sim id
        self.particle_steps = 0
        self.offloads = 0
        self.ticks = 0     # batched updates issued
        self.stepped = 0   # simulation updates over all ticks

    # ---- pool ----------------------------------------------------------------
    def _create(self, sid):
        sp = self.specs[sid]
        wp = WorldProperties(sp.gravity, sp.border_damping, sp.cfl, sp.solve_error)
        snap = self.offloaded.pop(sid, None)
        xy = sp.xy if snap is None else np.stack([snap["POS_X"][:sp.xy.shape[0]],
                                                   snap["POS_Y"][:sp.xy.shape[0]]], axis=1)
        sim = self.batch.create_sim(sp.width, sp.height, sp.dx, wp, xy)
        if sp.gravity_x:
            sim.set_gravity_x(sp.gravity_x)
        if sp.solver:
            sim.set_solver_config(**sp.solver)
        if snap is not None:
            for name in PARTICLE_FIELDS:
                sim.write(name, snap[name])
            sim.set_timestamp(snap["timestamp"], snap["prev_deltatime"])
        self.resident[sid] = sim
        self.sim_bytes[sid] = sim.device_bytes()
        self.used += self.sim_bytes[sid]

    def _offload(self, sid):
        sim = self.resident.pop(sid)
        snap = {name: sim.read(name) for name in PARTICLE_FIELDS}
        snap["timestamp"], snap["prev_deltatime"] = sim.timestamp, sim.prev_deltatime
        self.offloaded[sid] = snap
        sim.close()
        self.used -= self.sim_bytes.pop(sid)
        self.offloads += 1

    def _make_room(self, need, keep, waiting_ids):
        """Offload idle simulations, least recently used first, until `need`
        more bytes fit: first those without a waiting request, then, if that
        is not enough, those with one (they are restored when admitted).
        False when the active simulations alone leave no room."""
        for spare_waiting in (True, False):
            for sid in list(self.resident):
                if self.used + need <= self.budget:
                    return True
                if sid in self.active or sid == keep or (spare_waiting and sid in waiting_ids):
                    continue
                self._offload(sid)
        return self.used + need <= self.budget

    def _admit(self):
        still = collections.deque()
        blocked = False
        while self.waiting:
            r = self.waiting.popleft()
            if blocked or r.sim in self.active:
                still.append(r)
                continue
            if r.sim not in self.resident:
                waiting_ids = {q.sim for q in still} | {q.sim for q in self.waiting}
                if not self._make_room(self.specs[r.sim].estimate_bytes(), r.sim, waiting_ids):
                    blocked = True  # first come, first served: later requests wait too
                    still.append(r)
                    continue
                self._create(r.sim)
            r.remaining = r.steps
            self.active[r.sim] = r
        self.waiting = still

    def preload(self, ids):
        """Creates simulations ahead of their first request while the budget allows."""
        for sid in ids:
            if sid in self.resident:
                continue
            if self.used + self.specs[sid].estimate_bytes() > self.budget:
                break
            self._create(sid)

    def state(self, sid):
        """Particle fields and step scalars of a simulation, resident or offloaded."""
        if sid in self.offloaded:
            return self.offloaded[sid]
        sim = self.resident[sid]
        out = {name: sim.read(name) for name in PARTICLE_FIELDS}
        out["timestamp"], out["prev_deltatime"] = sim.timestamp, sim.prev_deltatime
        return out

    # ---- requests --------------------------------------------------------------
    def submit(self, req):
        self.waiting.append(req)

    def idle(self):
        return not self.active and not self.waiting

    def tick(self, now=None):
        """Admit, then one batched update of every simulation with an active
        request; returns the number of simulations stepped."""
        self._admit()
        if not self.active:
            return 0
        ids = list(self.active)
        t0 = time.perf_counter() if now is None else now
        for sid in ids:
            if self.active[sid].t_start is None:
                self.active[sid].t_start = t0
        sims = [self.resident[sid] for sid in ids]
        self.batch.update(sims, self.opts, self.frame_dt)
        self.ticks += 1
        self.stepped += len(ids)
        sims[0].get_max_particle_velocity()  # the tick's work has finished
        t1 = time.perf_counter()
        for sid in ids:
            r = self.active[sid]
            r.remaining -= 1
            self.steps_of[sid] += 1
            self.particle_steps += int(self.specs[sid].xy.shape[0])
            self.resident.move_to_end(sid)
            if r.remaining == 0:
                r.t_done = t1
                self.done.append(r)
                del self.active[sid]
        return len(ids)

    def close(self):
        for sim in self.resident.values():
            sim.close()
        self.resident.clear()
        self.batch.close()


def poisson_trace(n_requests, rate, n_sims, seed, max_steps=20):
    """Requests with exponential inter-arrival times (rate per second), a
    uniformly drawn simulation and 1..max_steps updates each."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))
    sims = rng.integers(0, n_sims, size=n_requests)
    steps = rng.integers(1, max_steps + 1, size=n_requests)
    return [Request(i, int(s), int(k), float(a)) for i, (a, s, k) in enumerate(zip(t, sims, steps))]


def run_served(server, trace, clock=time.perf_counter):
    """Replays the trace in real time through the server; returns the wall
    time from the first arrival to the last completion."""
    t0 = clock()
    pending = collections.deque(sorted(trace, key=lambda r: r.t_arrival))
    for r in pending:
        r.t_arrival += t0
    while pending or not server.idle():
        now = clock()
        while pending and pending[0].t_arrival <= now:
            server.submit(pending.popleft())
        if server.tick() == 0 and pending:
            time.sleep(max(0.0, min(pending[0].t_arrival - clock(), 0.001)))
    return max(r.t_done for r in trace) - min(r.t_arrival for r in trace)


def run_one_at_a_time(sims, specs, trace, frame_dt=0.002, debug_options=None, clock=time.perf_counter):
    """The same trace with plain Apic.update calls, one request and one
    simulation at a time in arrival order; returns (wall time, particle-steps)."""
    opts = debug_options or DebugOptions()
    t0 = clock()
    order = sorted(trace, key=lambda r: r.t_arrival)
    for r in order:
        r.t_arrival += t0
    ps = 0
    for r in order:
        while clock() < r.t_arrival:
            time.sleep(min(r.t_arrival - clock(), 0.001))
        r.t_start = clock()
        sim = sims[r.sim]
        for _ in range(r.steps):
            sim.update(opts, frame_dt)
        sim.get_max_particle_velocity()  # this request's updates have finished
        r.t_done = clock()
        ps += int(specs[r.sim].xy.shape[0]) * r.steps
    return max(r.t_done for r in order) - min(r.t_arrival for r in order), ps
