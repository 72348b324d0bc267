"""The relative mode on grids whose side is 2 x odd: the reference's hierarchy
stops halving at an odd side (src/apic.rs:900-911), leaving a coarsest level
of up to (side / 2)^2 cells; the relative mode's V-cycle goes on below it,
each side halved and rounded up, until the coarsest has at most 64 cells or
a side of 2.  Sides 126 (63^2 -> 32^2 ... 8^2), 254 (127^2 -> 64^2 ... 8^2),
510 (255^2 -> 128^2 ... 8^2), 254 x 126 (127 x 63 -> 64 x 32 ... 8 x 4) and
2046 x 10 (1023 x 5 -> 512 x 3 -> 256 x 2) and 2052 x 8 (1026 x 4 -> 513 x 2:
a coarsest level of more than 1024 cells, solved by the one-workgroup CG in
global memory).

  - apply_preconditioner on the oracle's assembled system: e on every level
    bit-exact against the oracle (its coarse CG's FP64 sums run in ascending
    order, the GPU's in a tree; they round to the same FP32 step lengths
    unless an FP64 rounding tips one, about one chance in 10^5 per solve);
  - one relative solve (1e-6): exit by the tolerance within one iteration of
    the oracle, pressure within 1e-5 relative L2, the true residual of the
    FP64 pressure within the tolerance;
  - one update: particles, grid velocities and pressure within the mixed
    workload's bars (1e-5 positions and pressure);
  - a batch of such simulations beside others: every field and counter equal
    to stepping each alone, bit for bit.
"""
import numpy as np
import pytest

from helpers import PARTICLE_FIELDS, bit_equal, first_mismatch, props, rel_l2
from test_oracle_relative import assembled, residuals

pytestmark = pytest.mark.gpu

RELATIVE = dict(mode=1, rel_tol=1e-6, max_iterations=200, fp64_reductions=True, recompute_every=25)
PRE_SOLVE = ["update_spatial_lookup", "transfer_particles_to_grid", "apply_external_forces",
             "build_pressure_system", "update_multilevel_grid"]
LEVEL_FIELDS = ["FLUID", "DIAG", "INV_DIAG", "PLUS_X", "PLUS_Y"]
DT = 0.002
# (width, height, the relative mode's coarsest level)
GRIDS = [(126, 126, (8, 8)), (254, 254, (8, 8)), (510, 510, (8, 8)), (254, 126, (8, 4)), (2046, 10, (256, 2)),
         (2052, 8, (513, 2))]
TOL = {"POS_X": 1e-5, "POS_Y": 1e-5, "PRESSURE": 1e-5, "TIME_RESIDUAL": 2e-4, "VEL_X": 2e-5, "VEL_Y": 2e-5,
       "MAC_U": 2e-5, "MAC_V": 2e-5, "C_U_X": 2e-4, "C_U_Y": 2e-4, "C_V_X": 2e-4, "C_V_Y": 2e-4}


def _scene(w, h):
    from FLIP import scenes
    return scenes.dam_break_fill(w, h, 1.0 / max(w, h), 0.4, 4, 5)


def _oracle(w, h, xy):
    from oracle import pyoracle as po
    o = po.Oracle(w, h, 1.0 / max(w, h), po.bench_props(), xy=xy)
    o.set_solver(1, RELATIVE["rel_tol"], RELATIVE["max_iterations"], True, RELATIVE["recompute_every"])
    return o


def _gpu(w, h, xy):
    from FLIP.apic import Apic
    g = Apic.with_particles(w, h, 1.0 / max(w, h), props(), xy)
    g.set_solver_config(**RELATIVE)
    return g


def _system(w, h):
    """The oracle after one relative-mode update and the next substep's
    assembled system (every level), and a GPU simulation holding it."""
    xy = _scene(w, h)
    o = _oracle(w, h, xy)
    o.update(DT)
    for st in PRE_SOLVE:
        o.stage(st, DT)
    d = o.dims()
    nl = d["num_levels_relative"]
    g = _gpu(w, h, xy)
    assert g.dims()["num_levels"] == d["num_levels"] and g.dims()["num_levels_relative"] == nl
    for lvl in range(nl + 1):
        for f in LEVEL_FIELDS:
            g.write(f, o.read(f, lvl), lvl)
    b, p0 = o.read("RESIDUAL"), o.read("PRESSURE")
    g.write("RESIDUAL", b)
    g.write("PRESSURE", p0)
    return o, g, nl, b, p0


def _sides(w, h, nl):
    """Level sides: halved, rounded up (exact halves on the reference's levels)."""
    out = [(w, h)]
    for _ in range(nl):
        out.append(((out[-1][0] + 1) // 2, (out[-1][1] + 1) // 2))
    return out


@pytest.mark.parametrize("grid", GRIDS)
def test_odd_levels_preconditioner_bit_exact(grid):
    w, h, coarsest = grid
    o, g, nl, b, _ = _system(w, h)
    sides = _sides(w, h, nl)
    assert sides[-1] == coarsest, sides
    for lvl, (lw, lh) in enumerate(sides):
        assert g.field_length("FLUID", lvl) == lw * lh
    print(f"{w} x {h}: {o.dims()['num_levels'] + 1} levels of the reference, {nl + 1} in all, "
          f"{sides[o.dims()['num_levels']]} -> {coarsest}", flush=True)
    o.write("RESIDUAL", b)
    o.stage("apply_preconditioner")
    g.apply_preconditioner()
    for lvl in range(nl + 1):
        lw = sides[lvl][0]
        fl = o.read("FLUID", lvl).astype(bool)
        m = np.zeros(o.read("ERROR", lvl).size, dtype=bool)
        m[lw:lw + fl.size] = fl
        a, e = g.read("ERROR", lvl)[m], o.read("ERROR", lvl)[m]
        assert bit_equal(a, e), (lvl, first_mismatch(a, e))
    g.close()


@pytest.mark.parametrize("grid", GRIDS)
def test_odd_levels_relative_solve(grid):
    w, h, _ = grid
    o, g, _, b, p0 = _system(w, h)
    sys_ = assembled(o)
    o.write("RESIDUAL", b)
    o.write("PRESSURE", p0)
    o.stage("solve_pcg")
    g.solve_pcg()
    so, sg = o.stats(), g.stats()
    res64, _ = residuals(*sys_, g.read("PRESSURE_F64"))
    perr = rel_l2(g.read("PRESSURE"), o.read("PRESSURE"))
    print(f"{w} x {h} relative solve: iterations GPU {sg['pcg_iterations']} oracle {so['pcg_iterations']}, "
          f"||r||/||b|| {sg['pcg_rel_residual']:.3e}, true residual {res64:.3e}, pressure vs oracle {perr:.2e}")
    assert sg["pcg_exit_reason"] == 3 and so["pcg_exit_reason"] == 3, (sg, so)
    assert abs(int(sg["pcg_iterations"]) - int(so["pcg_iterations"])) <= 1
    assert res64 <= 1e-6 * (1 + 1e-6), res64
    assert perr <= 1e-5, perr
    g.close()


@pytest.mark.parametrize("grid", GRIDS)
def test_odd_levels_one_update(grid):
    from FLIP.apic import DebugOptions
    w, h, _ = grid
    xy = _scene(w, h)
    o, g = _oracle(w, h, xy), _gpu(w, h, xy)
    o.update(DT)
    g.update(DebugOptions(), DT)
    so, sg = o.stats(), g.stats()
    print(f"{w} x {h} update: iterations GPU {sg['pcg_iterations']} oracle {so['pcg_iterations']}", flush=True)
    assert sg["pcg_exit_reason"] == 3 and abs(int(sg["pcg_iterations"]) - int(so["pcg_iterations"])) <= 1
    for f in PARTICLE_FIELDS + ["PRESSURE", "MAC_U", "MAC_V"]:
        err = rel_l2(g.read(f), o.read(f))
        assert err <= TOL[f], (f, err)
    g.close()


def test_odd_levels_batched_equals_alone():
    from FLIP.apic import Apic, Batch, DebugOptions
    from test_gpu_batch import _assert_same
    # (width, height, solver mode): the relative mode's deeper levels on
    # three 2 x odd grids, beside a reference-rules one (the one-workgroup
    # serial CG on 63 x 63) and a grid that halves down to 3 x 3
    specs = [(254, 254, 1), (126, 126, 1), (190, 190, 1), (126, 126, 0), (96, 96, 1)]
    print(f"odd levels, batch vs alone: {len(specs)} simulations, 3 updates", flush=True)
    batch = Batch()
    alone, batched = [], []
    for w, h, mode in specs:
        xy = _scene(w, h)
        a = Apic.with_particles(w, h, 1.0 / w, props(), xy)
        s = batch.create_sim(w, h, 1.0 / w, props(), xy)
        if mode == 1:
            a.set_solver_config(**RELATIVE)
            s.set_solver_config(**RELATIVE)
        alone.append(a)
        batched.append(s)
    opts = DebugOptions()
    for _ in range(3):
        for a in alone:
            a.update(opts, DT)
        batch.update(batched, opts, [DT] * len(specs))
    for i, (a, b) in enumerate(zip(alone, batched)):
        _assert_same(a, b, f"sim {i} {specs[i]}")
        if specs[i][2] == 1:
            assert a.stats()["pcg_exit_reason"] == 3, (specs[i], a.stats())
    for s in alone + batched:
        s.close()
    batch.close()
