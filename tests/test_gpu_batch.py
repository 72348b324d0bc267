"""Batched stepping (apic_batch_*): each simulation of a batch update computes
exactly what it computes when stepped alone with apic_update.

The batch runs one launch per kernel kind over every simulation's workgroups
(the PCG iterations in lockstep, a finished solve skipping the rest), so the
arithmetic of each simulation is that of the single path; the comparison is
bit for bit on every particle and grid field and on the solver counters.
"""
import numpy as np
import pytest

from helpers import PARTICLE_FIELDS

pytestmark = pytest.mark.gpu

GRID_FIELDS = ["MAC_U", "MAC_V", "VALID_U", "VALID_V", "DENSITY", "SMOOTHED_DENSITY", "PRESSURE", "FLUID"]
STAT_KEYS = ["substeps", "pcg_iterations", "pcg_iterations_total", "pcg_exit_reason", "pcg_max_residual",
             "pcg_rel_residual", "fluid_cells", "timestamp", "prev_deltatime", "max_particle_velocity"]

# (width, height, fill, ppc, seed, gravity angle in degrees, solver mode, frame dt)
SPECS = [
    (64, 64, 0.30, 4, 11, 0.0, 0, 0.002),
    (96, 96, 0.50, 6, 12, 10.0, 1, 0.002),
    (128, 64, 0.24, 4, 13, -7.0, 1, 0.002),
    (160, 160, 0.40, 8, 14, 3.0, 0, 0.002),
    (64, 128, 0.60, 5, 15, 0.0, 1, 0.03),   # several substeps per update (CFL-limited dt)
    (32, 32, 0.40, 4, 16, 0.0, 1, 0.002),   # one sort pass (10-bit keys)
    (80, 72, 0.40, 6, 17, 5.0, 0, 0.002),   # coarsest level 10 x 9: workgroup-wide coarse CG
    (320, 320, 0.30, 4, 18, -4.0, 1, 0.002),  # four levels above the one-workgroup bottom
]


def _configure(sim, spec):
    from FLIP.apic import WorldProperties  # noqa: F401
    w, h, fill, ppc, seed, ang, mode, dt = spec
    sim.set_gravity_x(float(9.81 * np.sin(np.radians(ang))))
    if mode == 1:
        sim.set_solver_config(mode=1, rel_tol=1e-6, max_iterations=200, fp64_reductions=True, recompute_every=7)


def _world(spec):
    from FLIP.apic import WorldProperties
    ang = spec[5]
    return WorldProperties(float(-9.81 * np.cos(np.radians(ang))), 0.1, 5.0, 1e-5)


def _scene(spec):
    from FLIP import scenes
    w, h, fill, ppc, seed = spec[:5]
    return scenes.dam_break_fill(w, h, 1.0 / max(w, h), fill, ppc, seed)


def _alone(specs):
    from FLIP.apic import Apic
    sims = []
    for sp in specs:
        s = Apic.with_particles(sp[0], sp[1], 1.0 / max(sp[0], sp[1]), _world(sp), _scene(sp))
        _configure(s, sp)
        sims.append(s)
    return sims


def _batched(batch, specs):
    sims = []
    for sp in specs:
        s = batch.create_sim(sp[0], sp[1], 1.0 / max(sp[0], sp[1]), _world(sp), _scene(sp))
        _configure(s, sp)
        sims.append(s)
    return sims


def _assert_same(a, b, label):
    for name in PARTICLE_FIELDS + GRID_FIELDS:
        x, y = a.read(name), b.read(name)
        assert x.shape == y.shape, (label, name)
        if x.dtype == np.float32:
            same = np.array_equal(x.view(np.uint32), y.view(np.uint32))
        else:
            same = np.array_equal(x, y)
        assert same, f"{label}: {name} differs at {np.flatnonzero(x != y)[:5]}"
    sa, sb = a.stats(), b.stats()
    for k in STAT_KEYS:
        assert sa[k] == sb[k] or (sa[k] != sa[k] and sb[k] != sb[k]), (label, k, sa[k], sb[k])


def _step_both(alone, batched, batch, specs, steps, disable_sort=False):
    from FLIP.apic import DebugOptions
    opts = DebugOptions(disable_particle_sort=disable_sort)
    dts = [sp[7] for sp in specs]
    for _ in range(steps):
        for s, dt in zip(alone, dts):
            s.update(opts, dt)
        batch.update(batched, opts, dts)


def test_batched_equals_alone_bit_for_bit():
    from FLIP.apic import Batch
    print(f"batch vs alone: {len(SPECS)} simulations, 12 updates", flush=True)
    alone = _alone(SPECS)
    batch = Batch()
    batched = _batched(batch, SPECS)
    _step_both(alone, batched, batch, SPECS, 12)
    for i, (a, b) in enumerate(zip(alone, batched)):
        _assert_same(a, b, f"sim {i} {SPECS[i][:2]}")
        assert a.stats()["pcg_iterations"] >= 1
    assert batched[4].stats()["substeps"] > 1
    for s in alone + batched:
        s.close()
    batch.close()


def test_batched_equals_alone_without_sort():
    from FLIP.apic import Batch
    specs = SPECS[:4]
    print(f"batch vs alone, sort disabled: {len(specs)} simulations, 6 updates", flush=True)
    alone = _alone(specs)
    batch = Batch()
    batched = _batched(batch, specs)
    _step_both(alone, batched, batch, specs, 6, disable_sort=True)
    for i, (a, b) in enumerate(zip(alone, batched)):
        _assert_same(a, b, f"sim {i}")
    for s in alone + batched:
        s.close()
    batch.close()


def test_subset_update_leaves_the_others():
    from FLIP.apic import Batch, DebugOptions
    specs = SPECS[:4]
    print("subset update: 4 simulations, 2 stepped", flush=True)
    batch = Batch()
    batched = _batched(batch, specs)
    before = {i: {n: batched[i].read(n).copy() for n in PARTICLE_FIELDS} for i in (1, 3)}
    alone = _alone([specs[0], specs[2]])
    opts = DebugOptions()
    for _ in range(3):
        batch.update([batched[0], batched[2]], opts, [specs[0][7], specs[2][7]])
        for s, sp in zip(alone, (specs[0], specs[2])):
            s.update(opts, sp[7])
    for i in (1, 3):
        for n in PARTICLE_FIELDS:
            assert np.array_equal(batched[i].read(n), before[i][n]), (i, n)
        assert batched[i].stats()["timestamp"] == 0
    _assert_same(alone[0], batched[0], "sim 0")
    _assert_same(alone[1], batched[2], "sim 2")
    # the untouched ones step alone afterwards like any simulation
    batched[1].update(opts, specs[1][7])
    ref = _alone([specs[1]])[0]
    ref.update(opts, specs[1][7])
    _assert_same(ref, batched[1], "sim 1 after a single update")
    for s in alone + batched + [ref]:
        s.close()
    batch.close()


def test_batch_argument_errors():
    from FLIP._lib import ApicError
    from FLIP.apic import Apic, Batch, DebugOptions
    batch = Batch()
    other = Batch()
    sp = SPECS[0]
    a = batch.create_sim(sp[0], sp[1], 1.0 / sp[0], _world(sp), _scene(sp))
    b = other.create_sim(sp[0], sp[1], 1.0 / sp[0], _world(sp), _scene(sp))
    lone = Apic.with_particles(sp[0], sp[1], 1.0 / sp[0], _world(sp), _scene(sp))
    with pytest.raises(ApicError):
        batch.update([a, b], DebugOptions(), 0.002)   # b belongs to another batch
    with pytest.raises(ApicError):
        batch.update([a, lone], DebugOptions(), 0.002)  # not in any batch
    with pytest.raises(ApicError):
        batch.update([a, a], DebugOptions(), 0.002)   # listed twice
    with pytest.raises(ApicError):
        batch.close()                                 # still has a simulation
    assert a.stats()["timestamp"] == 0                # nothing ran
    for s in (a, b, lone):
        s.close()
    batch.close()
    other.close()
