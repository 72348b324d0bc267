"""The continuous-batching server (FLIP.serve): every simulation's state after
a served trace -- admitted among others, stepped in batches of whatever was
active, offloaded to host memory and restored when the device budget was
full -- equals, bit for bit, that of the same number of plain Apic.update
calls on the same simulation alone."""
import numpy as np
import pytest

from helpers import PARTICLE_FIELDS

pytestmark = pytest.mark.gpu

REL = dict(mode=1, rel_tol=1e-6, max_iterations=200, fp64_reductions=True, recompute_every=25)


def _specs():
    from FLIP import scenes
    from FLIP.serve import SimSpec
    out = []
    for i, (side, fill, ppc, ang, rel) in enumerate([(64, 0.3, 4, 0.0, False), (80, 0.5, 6, 9.0, True),
                                                      (96, 0.4, 5, -12.0, True), (128, 0.25, 4, 3.0, False),
                                                      (160, 0.45, 8, 6.0, True), (64, 0.6, 7, -4.0, True)]):
        dx = 1.0 / side
        xy = scenes.dam_break_fill(side, side, dx, fill, ppc, 40 + i)
        out.append(SimSpec(side, side, dx, xy, gravity=float(-9.81 * np.cos(np.radians(ang))),
                           gravity_x=float(9.81 * np.sin(np.radians(ang))), solver=REL if rel else {}))
    return out


def _alone(spec, steps, dt):
    from FLIP.apic import Apic, DebugOptions, WorldProperties
    sim = Apic.with_particles(spec.width, spec.height, spec.dx,
                              WorldProperties(spec.gravity, spec.border_damping, spec.cfl, spec.solve_error), spec.xy)
    if spec.gravity_x:
        sim.set_gravity_x(spec.gravity_x)
    if spec.solver:
        sim.set_solver_config(**spec.solver)
    for _ in range(steps):
        sim.update(DebugOptions(), dt)
    out = {name: sim.read(name) for name in PARTICLE_FIELDS}
    out["timestamp"], out["prev_deltatime"] = sim.timestamp, sim.prev_deltatime
    sim.close()
    return out


@pytest.mark.parametrize("tight", [False, True])
def test_served_equals_alone(tight):
    from FLIP.serve import Request, Server
    specs = _specs()
    rng = np.random.default_rng(7)
    trace = [Request(i, int(rng.integers(0, len(specs))), int(rng.integers(1, 6)), 0.0) for i in range(30)]
    # tight: room for the largest simulation and little more, so that the pool
    # must offload idle simulations and restore them later
    budget = max(sp.estimate_bytes() for sp in specs) + (1 << 20) if tight else None
    server = Server(specs, hbm_budget=budget)
    print(f"serving {len(trace)} requests over {len(specs)} simulations, tight={tight}", flush=True)
    pending = list(trace)
    ticks = 0
    while pending or not server.idle():
        # a few arrivals per tick, so that requests join while others run
        for _ in range(min(3, len(pending))):
            server.submit(pending.pop(0))
        server.tick()
        ticks += 1
        assert ticks < 1000
    assert len(server.done) == len(trace)
    assert all(r.remaining == 0 for r in trace)
    if tight:
        assert server.offloads > 0
        assert server.used <= server.budget
    for sid, spec in enumerate(specs):
        n = server.steps_of[sid]
        assert n == sum(r.steps for r in trace if r.sim == sid)
        if n == 0:
            continue
        got, ref = server.state(sid), _alone(spec, n, server.frame_dt)
        assert got["timestamp"] == ref["timestamp"] == n
        assert got["prev_deltatime"] == ref["prev_deltatime"]
        for name in PARTICLE_FIELDS:
            assert np.array_equal(got[name].view(np.uint32), ref[name].view(np.uint32)), (sid, name)
    server.close()
