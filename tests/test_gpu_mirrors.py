"""Host mirrors of the reference's pub particle vectors (src/apic.rs:757-767:
Vec<Vec2x8> with eight x then eight y per chunk, Vec<f32x8>) and the
add_radial_force upload path (:2352-2410) against its FP64 closed form."""
import numpy as np
import pytest

from helpers import props
from radial_closed_form import radial_force

pytestmark = pytest.mark.gpu

PAIRS = {"part_positions": ("POS_X", "POS_Y"), "part_velocities": ("VEL_X", "VEL_Y"),
         "part_c_u": ("C_U_X", "C_U_Y"), "part_c_v": ("C_V_X", "C_V_Y")}


def ragged():
    from FLIP.apic import Apic, DebugOptions
    rng = np.random.default_rng(5)
    xy = rng.uniform(4.0, 20.0, size=(1237, 2)).astype(np.float32)  # 155 chunks, last one ragged
    g = Apic.with_particles(32, 32, 1.0, props(), xy)
    for _ in range(3):
        g.update(DebugOptions(False), 0.016)
    return g


def test_chunk_layout_matches_the_soa_fields():
    g = ragged()
    nc = g.num_chunks
    for group, (fx, fy) in PAIRS.items():
        a = g.read_particles_aosoa(group)
        assert a.shape == (nc, 2, 8)
        assert np.array_equal(a[:, 0, :].reshape(-1), g.read(fx)), group
        assert np.array_equal(a[:, 1, :].reshape(-1), g.read(fy)), group
    t = g.read_particles_aosoa("part_time_residual")
    assert t.shape == (nc, 8) and np.array_equal(t.reshape(-1), g.read("TIME_RESIDUAL"))
    g.close()


def test_chunk_layout_round_trip():
    g = ragged()
    rng = np.random.default_rng(9)
    for group in list(PAIRS) + ["part_time_residual"]:
        shape = g.read_particles_aosoa(group).shape
        v = rng.standard_normal(shape).astype(np.float32)
        g.write_particles_aosoa(group, v)
        assert np.array_equal(g.read_particles_aosoa(group), v), group
        if group in PAIRS:
            fx, fy = PAIRS[group]
            assert np.array_equal(g.read(fx), v[:, 0, :].reshape(-1))
    # a velocity write through the mirror refreshes max |v| (get_max_particle_velocity, :2412-2427)
    vel = g.read_particles_aosoa("part_velocities")
    expect = np.float32(np.sqrt(vel[:, 0, :] ** 2 + vel[:, 1, :] ** 2).max())
    assert abs(g.get_max_particle_velocity() - expect) <= 1e-6 * expect
    with pytest.raises(Exception):
        g.write_particles_aosoa("part_positions", np.zeros((3, 2, 8), np.float32))
    g.close()


@pytest.mark.parametrize("args", [(0.2, 0.3, 0.1, 50.0), (0.3, 0.2, 0.08, -40.0), (0.01, 0.5, 0.2, 5.0)])
def test_radial_force_closed_form(args):
    from FLIP import scenes
    from FLIP.apic import Apic
    xy = scenes.config_scene("cfg1_dam_break_128")
    g = Apic.with_particles(128, 128, 1.0 / 128, props(), xy)
    g.add_radial_force(*args)
    fu, fv = radial_force(128, 128, 1.0 / 128, *args)
    repel = args[3] > 0
    gu = g.read("POST_FORCE_U" if repel else "EXT_FORCE_U").astype(np.float64)
    gv = g.read("POST_FORCE_V" if repel else "EXT_FORCE_V").astype(np.float64)
    tol = 1e-5 * abs(args[3])
    assert np.max(np.abs(gu - fu)) <= tol and np.max(np.abs(gv - fv)) <= tol
    assert np.count_nonzero(gu) == np.count_nonzero(fu) and np.count_nonzero(gv) == np.count_nonzero(fv)
    g.close()
