"""Oracle parity at the benchmarked size: the 1024 x 1024 dam break (dx =
1/1024, 8 particles per fluid cell, 2,005,736 particles, dt = 2.5e-4) under
the solver rules the benchmark times (relative 1e-6, FP64 reductions, cap
500).  These are the kernels a 1024^2 grid selects and smaller grids do not:
the 512-thread base tiles (VDown/VUp with 32 x 32 tiles and 512 threads, the
up-sweep with its FP64 gamma and flexible-beta partials), the level pairs
(VDown2/VUp2) above an eight-level hierarchy, the one-workgroup bottom, the
1024-thread SpmvDot, the update fused into the base down-sweep (VDownAx) and
the separate tile update (AxpyTile) with the true-residual kernel.

Bars:
  - apply_preconditioner (src/apic.rs:1788-1821): every level's error bit for
    bit (no reduction outside the coarse CG, which sums serially in the
    reference's order);
  - solve_pcg in the relative mode (the reference's loop, src/apic.rs:1865-1956,
    with a relative test): iteration count within one of the oracle's, exit by
    the tolerance, ||r||/||b|| <= 1e-6, pressure within 1e-5 relative L2 of the
    oracle's, and ||b - A p|| / ||b|| evaluated in FP64 from the GPU's own
    FP32 pressure within 1e-6 plus the FP32 rounding bound, and from its
    FP64-accumulated pressure within 1e-6 (1 + 1e-6) and within 1e-12 of the
    residual the solver reported;
  - one Apic::update (src/apic.rs:2429-2474): particles, MAC u and v, density
    and pressure within 1e-5 relative L2 of the FP32 oracle, or within twice
    the oracle's own spread over valid P2G summation orders where that is
    larger (the time residual; the FP64 oracle's distance is printed beside
    it);
  - a simulation stepped alone (fused update, level pairs) and in a batch
    (separate AxpyTile update, one launch per level) agree bit for bit.
"""
import numpy as np
import pytest

from helpers import PARTICLE_FIELDS, bit_equal, first_mismatch, props, rel_l2, to_gpu
from test_oracle_relative import assembled, residuals

pytestmark = pytest.mark.gpu

W = H = 1024
DX = 1.0 / 1024
DT = 2.5e-4
WARM = 12          # oracle updates before the state the tests start from
RELATIVE = dict(mode=1, rel_tol=1e-6, max_iterations=500, fp64_reductions=True, recompute_every=25)
PRE_SOLVE = ["update_spatial_lookup", "transfer_particles_to_grid", "apply_external_forces",
             "build_pressure_system", "update_multilevel_grid"]
LEVEL_FIELDS = ["FLUID", "DIAG", "INV_DIAG", "PLUS_X", "PLUS_Y"]


def _scene():
    from FLIP import scenes
    return scenes.config_scene("cfg2_dam_break_1024")


def _relative(o):
    o.set_solver(1, RELATIVE["rel_tol"], RELATIVE["max_iterations"], True, RELATIVE["recompute_every"])


@pytest.fixture(scope="module")
def warm():
    """The oracle after WARM relative-mode updates: its particle state and
    pressure, and the assembled system of the next substep (every level)."""
    from oracle import pyoracle as po
    xy = _scene()
    print(f"headline parity: oracle {WARM} updates of {xy.shape[0]} particles", flush=True)
    o = po.Oracle(W, H, DX, po.bench_props(), xy=xy)
    _relative(o)
    for _ in range(WARM):
        o.update(DT)
    state = {f: o.read(f) for f in PARTICLE_FIELDS + ["PRESSURE"]}
    state["timestamp"] = o.timestamp()
    for st in PRE_SOLVE:
        o.stage(st, DT)
    nl = o.dims()["num_levels_relative"]  # the relative mode's levels
    levels = [{f: o.read(f, l) for f in LEVEL_FIELDS} for l in range(nl + 1)]
    b = o.read("RESIDUAL")
    p0 = o.read("PRESSURE")
    sys_ = assembled(o)
    return dict(xy=xy, o=o, state=state, levels=levels, b=b, p0=p0, sys=sys_, nl=nl)


def _gpu_with_system(warm):
    from FLIP.apic import Apic
    g = Apic.with_particles(W, H, DX, props(), warm["xy"])
    assert g.dims()["num_levels_relative"] == warm["nl"]
    for lvl, fields in enumerate(warm["levels"]):
        for f in LEVEL_FIELDS:
            g.write(f, fields[f], lvl)
    g.write("RESIDUAL", warm["b"])
    g.write("PRESSURE", warm["p0"])
    return g


def _padded_mask(fluid, w, n):
    m = np.zeros(n, dtype=bool)
    m[w:w + fluid.size] = fluid.astype(bool)
    return m


def test_headline_preconditioner_bit_exact(warm):
    o = warm["o"]
    o.write("RESIDUAL", warm["b"])
    o.stage("apply_preconditioner")
    g = _gpu_with_system(warm)
    g.set_solver_config(**RELATIVE)  # (the coarse CG's relative exits)
    g.apply_preconditioner()
    for lvl in range(warm["nl"] + 1):
        e_o = o.read("ERROR", lvl)
        m = _padded_mask(warm["levels"][lvl]["FLUID"], W >> lvl, e_o.size)
        a, b = g.read("ERROR", lvl)[m], e_o[m]
        assert bit_equal(a, b), (lvl, first_mismatch(a, b))
    print(f"1024^2 V-cycle: {warm['nl'] + 1} levels bit-exact "
          f"({int(warm['levels'][0]['FLUID'].sum())} base fluid cells)")
    g.close()


@pytest.mark.parametrize("every", [25, 3])
def test_headline_relative_solve(warm, every):
    o = warm["o"]
    o.write("RESIDUAL", warm["b"])
    o.write("PRESSURE", warm["p0"])
    o.set_solver(1, 1e-6, 500, True, every)
    o.stage("solve_pcg")
    g = _gpu_with_system(warm)
    g.set_solver_config(**dict(RELATIVE, recompute_every=every))
    g.solve_pcg()
    so, sg = o.stats(), g.stats()
    res, bound = residuals(*warm["sys"], g.read("PRESSURE"))
    res64, _ = residuals(*warm["sys"], g.read("PRESSURE_F64"))
    perr = rel_l2(g.read("PRESSURE"), o.read("PRESSURE"))
    print(f"1024^2 relative solve, recompute every {every}: iterations GPU {sg['pcg_iterations']} "
          f"oracle {so['pcg_iterations']}; ||r||/||b|| GPU {sg['pcg_rel_residual']:.3e} "
          f"oracle {so['pcg_rel_residual']:.3e}; true residual of the GPU's FP32 pressure {res:.3e} "
          f"(FP32 bound {bound:.1e}), of its FP64 pressure {res64:.3e}; pressure vs oracle {perr:.2e}")
    assert sg["pcg_exit_reason"] == 3 and so["pcg_exit_reason"] == 3, (sg, so)
    assert abs(int(sg["pcg_iterations"]) - int(so["pcg_iterations"])) <= 1
    assert 0.0 <= sg["pcg_rel_residual"] <= 1e-6
    assert res <= 1e-6 + bound, (res, bound)
    # the FP64-accumulated pressure meets the tolerance itself: the residual
    # the solver tests, r + r_lo, is its true residual to FP64 rounding
    assert res64 <= 1e-6 * (1 + 1e-6), res64
    assert abs(res64 - sg["pcg_rel_residual"]) <= 1e-12, (res64, sg["pcg_rel_residual"])
    assert perr <= 1e-5, perr
    g.close()


def test_headline_one_update(warm):
    """One update against the FP32 oracle.  The reference's own P2G summation
    order is not fixed (parallel CAS adds), and the oracle run with two other
    valid orders measures how far that alone moves each field; a field passes
    within 1e-5 or within twice that spread, whichever is larger.  The time
    residual is the field where the spread exceeds 1e-5 at this size: it is
    dt + t_res - dt_act (src/apic.rs:2181-2186), a cancellation of numbers
    of size dt = 2.5e-4 into values of size 1e-7, so one ulp of dt (3e-11) in
    dt_act is a relative change of 1e-4 in that particle's value; its
    absolute difference is bounded by two ulps of 2 dt besides."""
    from FLIP.apic import Apic, DebugOptions
    from oracle import pyoracle as po
    st = warm["state"]

    def oracle(precision=32, shuffle=0):
        o = po.Oracle(W, H, DX, po.bench_props(), xy=warm["xy"], precision=precision)
        _relative(o)
        for f in PARTICLE_FIELDS + ["PRESSURE"]:
            o.write(f, st[f])
        o.set_timestamp(*st["timestamp"])
        if shuffle:
            o.set_p2g_shuffle(shuffle)
        o.update(DT)
        return o

    o32, o64 = oracle(), oracle(64)
    others = [oracle(shuffle=s) for s in (12345, 777)]
    g = Apic.with_particles(W, H, DX, props(), warm["xy"])
    g.set_solver_config(**RELATIVE)
    for f in PARTICLE_FIELDS + ["PRESSURE"]:
        g.write(f, st[f])
    g.set_timestamp(*st["timestamp"])
    g.update(DebugOptions(False), DT)
    so, sg = o32.stats(), g.stats()
    assert sg["substeps"] == so["substeps"] == 1
    assert abs(int(sg["pcg_iterations"]) - int(so["pcg_iterations"])) <= 1, (sg, so)
    assert sg["pcg_exit_reason"] == 3 and sg["pcg_rel_residual"] <= 1e-6
    fields = PARTICLE_FIELDS + ["MAC_U", "MAC_V", "DENSITY", "PRESSURE"]
    errs = {f: rel_l2(g.read(f), o32.read(f)) for f in fields}
    errs64 = {f: rel_l2(g.read(f), o64.read(f)) for f in fields}
    spread = {f: max(rel_l2(x.read(f), o32.read(f)) for x in others) for f in fields}
    for f in fields:
        print(f"1024^2 update, {f}: GPU vs FP32 oracle {errs[f]:.2e} (oracle spread over P2G orders "
              f"{spread[f]:.2e}); GPU vs FP64 oracle {errs64[f]:.2e}")
    print(f"1024^2 update: {sg['pcg_iterations']} PCG iterations (oracle {so['pcg_iterations']})")
    for f in fields:
        assert errs[f] <= max(1e-5, 2.0 * spread[f]), (f, errs[f], spread[f])
    n = warm["xy"].shape[0]
    dt_res = np.abs(g.read("TIME_RESIDUAL")[:n].astype(np.float64) - o32.read("TIME_RESIDUAL")[:n])
    assert dt_res.max() <= 2 * np.spacing(np.float32(2 * DT)), dt_res.max()
    g.close()


def test_headline_alone_equals_batched():
    """One 1024^2 simulation stepped alone (the update fused into the base
    down-sweep, level pairs) and the same simulation in a batch with a 256^2
    one (separate AxpyTile update, one launch per level): bit for bit after
    three updates, solver counters included."""
    from FLIP import scenes
    from FLIP.apic import Apic, Batch, DebugOptions
    from test_gpu_batch import GRID_FIELDS, STAT_KEYS
    xy = _scene()
    small = scenes.dam_break(256, 256, 1.0 / 256, 0.4, 0.6, 6, 21)
    alone = Apic.with_particles(W, H, DX, props(), xy)
    alone.set_solver_config(**RELATIVE)
    batch = Batch()
    big = batch.create_sim(W, H, DX, props(), xy)
    other = batch.create_sim(256, 256, 1.0 / 256, props(), small)
    for s in (big, other):
        s.set_solver_config(**RELATIVE)
    opts = DebugOptions(False)
    for _ in range(3):
        alone.update(opts, DT)
        batch.update([big, other], opts, [DT, 0.002])
    assert alone.stats()["pcg_iterations"] >= 5
    for name in PARTICLE_FIELDS + GRID_FIELDS:
        x, y = alone.read(name), big.read(name)
        same = np.array_equal(x.view(np.uint32), y.view(np.uint32)) if x.dtype == np.float32 else np.array_equal(x, y)
        assert same, f"{name} differs at {np.flatnonzero(x != y)[:5]}"
    sa, sb = alone.stats(), big.stats()
    for k in STAT_KEYS:
        assert sa[k] == sb[k], (k, sa[k], sb[k])
    for s in (alone, big, other):
        s.close()
    batch.close()
