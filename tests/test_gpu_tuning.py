"""Launch shapes (apic_set_tuning, apic_autotune): the radix sort's digit
width and the multigrid levels' tile side and pairing change launches, never
a bit of the result.  Simulations stepped with every extreme of the shapes
equal the default's, field for field and counter for counter, in the
relative mode on a 512^2 dam break (levels of 256 .. 1 tiles) and on a
2 x odd grid (254^2: the multi-workgroup coarse CG), and under the
reference's rules; the autotuner leaves the state it times as it was.
"""
import pytest

from helpers import props

pytestmark = pytest.mark.gpu

RELATIVE = dict(mode=1, rel_tol=1e-6, max_iterations=500, fp64_reductions=True, recompute_every=25)
# (side, solver mode)
CASES = [(512, 1), (254, 1), (256, 0)]
SHAPES = [(8, 0), (8, 16384), (11, 0), (9, 256), (10, 4096)]


def _sim(side, mode):
    from FLIP import scenes
    from FLIP.apic import Apic
    xy = scenes.dam_break_fill(side, side, 1.0 / side, 0.4, 6, 9)
    g = Apic.with_particles(side, side, 1.0 / side, props(), xy)
    if mode == 1:
        g.set_solver_config(**RELATIVE)
    return g


def _step(g, n, dt=0.002):
    from FLIP.apic import DebugOptions
    for _ in range(n):
        g.update(DebugOptions(False), dt)


@pytest.mark.parametrize("case", CASES)
def test_launch_shapes_leave_every_bit(case):
    from test_gpu_batch import _assert_same
    side, mode = case
    ref = _sim(side, mode)
    assert ref.get_tuning()["radix_bits"] == 11 and ref.get_tuning()["few_tiles"] == 1024
    _step(ref, 3)
    for bits, few in SHAPES:
        g = _sim(side, mode)
        g.set_tuning(bits, few)
        _step(g, 3)
        print(f"{side}^2 mode {mode}: radix_bits {bits}, few_tiles {few}: "
              f"{g.stats()['pcg_iterations_total']} PCG iterations", flush=True)
        _assert_same(g, ref, f"{side} {bits} {few}")
        g.close()
    ref.close()


@pytest.mark.parametrize("case", CASES[:2])
def test_autotune_leaves_the_state(case):
    from test_gpu_batch import _assert_same
    side, mode = case
    a, b = _sim(side, mode), _sim(side, mode)
    _step(a, 2)
    _step(b, 2)
    t = b.autotune(reps=3)
    print(f"{side}^2: autotuned {t}", flush=True)
    assert 8 <= t["radix_bits"] <= 11 and t["few_tiles"] in (0, 64, 256, 1024, 4096, 16384)
    assert t["sort_ms"] > 0 and t["vcycle_ms"] > 0
    assert b.get_tuning() == t
    _step(a, 2)
    _step(b, 2)
    _assert_same(b, a, f"{side} autotuned")
    a.close()
    b.close()


def test_tuning_validation():
    from FLIP._lib import ApicError
    g = _sim(64, 0)
    for bits in (7, 12):
        with pytest.raises(ApicError):
            g.set_tuning(bits, 1024)
    with pytest.raises(ApicError):
        g.autotune(reps=0)
    g.close()
