"""apic_batch_update_noalloc (no counterpart in the reference; for callers
that cannot allocate device memory while they step) against
apic_batch_update, bit for bit, through the C ABI: two calls in a row, the
second over a smaller set of simulations, then a batch whose launch tables
exceed the fixed arena is refused before it steps anything."""
import numpy as np
import pytest

from helpers import PARTICLE_FIELDS

pytestmark = pytest.mark.gpu

# (width, height, fill, ppc, seed, solver mode, frame dt)
SPECS = [
    (64, 64, 0.30, 4, 31, 0, 0.002),
    (96, 96, 0.50, 6, 32, 1, 0.002),
    (128, 64, 0.24, 4, 33, 1, 0.002),
    (160, 160, 0.40, 8, 34, 0, 0.002),
    (64, 128, 0.60, 5, 35, 1, 0.03),   # several substeps per update
]
FIELDS = PARTICLE_FIELDS + ["MAC_U", "MAC_V", "VALID_U", "VALID_V", "DENSITY", "PRESSURE", "FLUID"]


def _make(batch, sp):
    from FLIP import scenes
    from FLIP.apic import WorldProperties
    w, h, fill, ppc, seed, mode, _ = sp
    s = batch.create_sim(w, h, 1.0 / max(w, h), WorldProperties(-9.81, 0.1, 5.0, 1e-5),
                         scenes.dam_break_fill(w, h, 1.0 / max(w, h), fill, ppc, seed))
    if mode == 1:
        s.set_solver_config(mode=1, rel_tol=1e-6, max_iterations=200, fp64_reductions=True, recompute_every=7)
    return s


def _same(a, b, label):
    for name in FIELDS:
        x, y = a.read(name), b.read(name)
        same = np.array_equal(x.view(np.uint32), y.view(np.uint32)) if x.dtype == np.float32 else np.array_equal(x, y)
        assert same, f"{label}: {name} differs at {np.flatnonzero(x != y)[:5]}"
    sa, sb = a.stats(), b.stats()
    for k in ("substeps", "pcg_iterations", "pcg_iterations_total", "pcg_exit_reason", "timestamp"):
        assert sa[k] == sb[k], (label, k, sa[k], sb[k])


def test_noalloc_batch_update_equals_batch_update():
    from FLIP.apic import Batch, DebugOptions
    print(f"no-allocation batch update: {len(SPECS)} simulations, then a subset of 2", flush=True)
    ref_b, fix_b = Batch(), Batch()
    ref = [_make(ref_b, sp) for sp in SPECS]
    fix = [_make(fix_b, sp) for sp in SPECS]
    dts = [sp[6] for sp in SPECS]
    opts = DebugOptions(False)
    for call, idx in enumerate(([0, 1, 2, 3, 4], [1, 3])):   # second call: a smaller input
        for _ in range(3):
            ref_b.update([ref[i] for i in idx], opts, [dts[i] for i in idx])
            fix_b.update([fix[i] for i in idx], opts, [dts[i] for i in idx], noalloc=True)
        for i in idx:
            _same(ref[i], fix[i], f"call {call}, sim {i}")
    assert fix[4].stats()["substeps"] > 1  # the 0.03 frame took several substeps
    for s in ref + fix:
        s.close()
    ref_b.close()
    fix_b.close()


def test_noalloc_two_batches_in_turn():
    """Two batches take turns on the one arena: each call's stream waits for
    the previous call's kernels, so neither sees the other's tables."""
    from FLIP.apic import Batch, DebugOptions
    ref_b, a_b, c_b = Batch(), Batch(), Batch()
    ref = [_make(ref_b, sp) for sp in SPECS[:4]]
    a = [_make(a_b, sp) for sp in SPECS[:2]]
    c = [_make(c_b, sp) for sp in SPECS[2:4]]
    opts = DebugOptions(False)
    for _ in range(3):
        ref_b.update(ref, opts, 0.002)
        a_b.update(a, opts, 0.002, noalloc=True)
        c_b.update(c, opts, 0.002, noalloc=True)
    for i, s in enumerate(a + c):
        _same(ref[i], s, f"sim {i}")
    for s in ref + a + c:
        s.close()
    for b in (ref_b, a_b, c_b):
        b.close()
