"""The mixed-precision large case: a 4096 x 2048 grid (dx = 1/4096) with the
bottom half filled and a 512 x 512 block near the top, 8 particles per fluid
cell (35,602,448 particles), FP32 transfers and SpMV, FP64 PCG reductions,
residual recompute every 25 iterations, relative 1e-6 test in FP64.

  - At 1/8 scale (512 x 256, the same construction with a 64 x 64 block,
    550,928 particles) the GPU is compared with the FP32 oracle over three
    updates: every particle and grid field within 1e-5 relative L2, or within
    twice the oracle's own spread over two other valid P2G summation orders
    where that is larger; the FP64 oracle's distance is printed beside it.
  - At full size, properties that do not need the oracle: each solve exits by
    the tolerance with ||r|| / ||b|| <= 1e-6, and the true residual
    ||b - A p|| / ||b|| of the FP64 pressure, evaluated on the host in FP64,
    is within 1e-6 (1 + 1e-6) and within 1e-12 of the reported residual; the Morton sort (23-bit keys, three radix passes) equals
    a numpy stable argsort of keys computed independently from the positions
    (src/apic.rs:1049-1066), bit for bit; the particle-to-grid density sums to
    the sum of the particles' temporal weights (the cell stencil's weights sum
    to one, src/apic.rs:1260-1284).
"""
import numpy as np
import pytest

from helpers import PARTICLE_FIELDS, props, rel_l2
from test_oracle_invariants import morton_np

pytestmark = pytest.mark.gpu

RELATIVE = dict(mode=1, rel_tol=1e-6, max_iterations=500, fp64_reductions=True, recompute_every=25)


def _host_true_residual(g):
    W = g.width
    b = g.read("RHS").astype(np.float64)
    diag, px, py = (g.read(f).astype(np.float64) for f in ("DIAG", "PLUS_X", "PLUS_Y"))
    idx = np.nonzero(g.read("FLUID"))[0] + W
    pp = np.zeros(b.size)
    pp[W:W + g.width * g.height] = g.read("PRESSURE_F64")
    ap = diag[idx] * pp[idx] + px[idx] * pp[idx + 1] + px[idx - 1] * pp[idx - 1] + \
        py[idx] * pp[idx + W] + py[idx - W] * pp[idx - W]
    return float(np.linalg.norm(b[idx] - ap) / np.linalg.norm(b[idx]))


def test_falling_block_eighth_scale_against_oracle():
    from FLIP import scenes
    from FLIP.apic import Apic, DebugOptions
    from oracle import pyoracle as po
    c = scenes.CONFIGS["falling_block_512x256"]
    W, H, dx, dt = c["width"], c["height"], c["dx"], c["dt"]
    xy = scenes.config_scene("falling_block_512x256")
    print(f"falling block 512 x 256: {xy.shape[0]} particles, 3 updates", flush=True)

    def oracle(precision=32, shuffle=0):
        o = po.Oracle(W, H, dx, po.bench_props(), xy=xy, precision=precision)
        o.set_solver(1, 1e-6, 500, True, 25)
        if shuffle:
            o.set_p2g_shuffle(shuffle)
        return o

    o32, o64 = oracle(), oracle(64)
    others = [oracle(shuffle=s) for s in (12345, 777)]
    g = Apic.with_particles(W, H, dx, props(), xy)
    g.set_solver_config(**RELATIVE)
    for step in range(3):
        g.update(DebugOptions(False), dt)
        for o in [o32, o64] + others:
            o.update(dt)
        sg, so = g.stats(), o32.stats()
        print(f"  update {step}: PCG iterations GPU {sg['pcg_iterations']} oracle {so['pcg_iterations']}, "
              f"||r||/||b|| {sg['pcg_rel_residual']:.2e}", flush=True)
        assert sg["pcg_exit_reason"] == 3 and sg["pcg_rel_residual"] <= 1e-6
        assert abs(int(sg["pcg_iterations"]) - int(so["pcg_iterations"])) <= 1
    fields = PARTICLE_FIELDS + ["MAC_U", "MAC_V", "DENSITY", "PRESSURE"]
    for f in fields:
        err = rel_l2(g.read(f), o32.read(f))
        spread = max(rel_l2(x.read(f), o32.read(f)) for x in others)
        print(f"  {f}: GPU vs FP32 oracle {err:.2e} (oracle spread {spread:.2e}), "
              f"vs FP64 oracle {rel_l2(g.read(f), o64.read(f)):.2e}")
        assert err <= max(1e-5, 2.0 * spread), (f, err, spread)
    tr = _host_true_residual(g)
    assert tr <= 1e-6 * (1 + 1e-6) and abs(tr - g.stats()["pcg_rel_residual"]) <= 1e-12, tr
    g.close()


def test_falling_block_full_size_properties():
    from FLIP import scenes
    from FLIP.apic import Apic, DebugOptions
    name = "cfg3_falling_block_4096x2048"
    c = scenes.CONFIGS[name]
    W, H, dx, dt = c["width"], c["height"], c["dx"], c["dt"]
    xy = scenes.config_scene(name)
    n = xy.shape[0]
    assert n == 35602448
    print(f"falling block 4096 x 2048: {n} particles", flush=True)
    g = Apic.with_particles(W, H, dx, props(), xy)
    assert g.dims()["num_levels"] == 9
    g.set_solver_config(**RELATIVE)
    for step in range(2):
        g.update(DebugOptions(False), dt)
        st = g.stats()
        tr = _host_true_residual(g)
        print(f"  update {step}: {st['pcg_iterations']} PCG iterations, ||r||/||b|| {st['pcg_rel_residual']:.3e}, "
              f"true residual of the FP64 pressure {tr:.3e}", flush=True)
        assert st["pcg_exit_reason"] == 3 and 0.0 < st["pcg_rel_residual"] <= 1e-6
        assert tr <= 1e-6 * (1 + 1e-6) and abs(tr - st["pcg_rel_residual"]) <= 1e-12, (tr, st)
    # the sort of the current positions against numpy
    x, y = g.read("POS_X")[:n], g.read("POS_Y")[:n]
    fdx = np.float32(dx)
    ix = np.clip(x / fdx, 0, W - 1).astype(np.int64)
    iy = np.clip(y / fdx, 0, H - 1).astype(np.int64)
    keys = morton_np(ix, iy)
    order = np.argsort(keys, kind="stable")
    g.update_spatial_lookup()
    assert np.array_equal(g.read("SORT_KEY"), keys[order])
    assert np.array_equal(g.read("SORT_VAL"), order.astype(np.uint32))
    assert np.array_equal(g.read("POS_X")[:n], x[order])
    print("  sort: keys and permutation equal numpy's stable argsort", flush=True)
    # total density = total temporal weight
    g.transfer_particles_to_grid()
    tres = g.read("TIME_RESIDUAL")[:n].astype(np.float64)
    tau = -tres / g.prev_deltatime
    wt = np.maximum(2.1875 * np.maximum(1.0 - (tau - 0.5) ** 2, 0.0) ** 3, 1e-5)
    dens = g.read("DENSITY").sum(dtype=np.float64)
    print(f"  density sum {dens:.6e}, temporal weights {wt.sum():.6e}", flush=True)
    assert abs(dens - wt.sum()) <= 1e-5 * wt.sum()
    g.close()
