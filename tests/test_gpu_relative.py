"""The relative solver mode on the GPU (apic_solver_config.mode 1, an
extension; the reference's solve_pcg, src/apic.rs:1865-1956, has absolute
exits only) against the oracle's relative mode and the assembled system.

Bars: iteration counts within one of the FP32 oracle's; pressure within 1e-5
relative L2 of it (the FP64 sums run in another order); ||b - A p|| / ||b||
evaluated in FP64 from the GPU's own pressure within rel_tol plus the
rounding of an FP32 pressure (8 eps32 || |A| |p| || / ||b||), and from its
FP64-accumulated pressure within rel_tol (1 + 1e-6) and within 1e-6 rel_tol
of the residual the solver reported (its r + r_lo is the true residual to
FP64 rounding).
"""
import numpy as np
import pytest

from helpers import props, rel_l2, to_gpu
from test_gpu_stages import make_pair, oracle_until
from test_oracle_relative import assembled, residuals

pytestmark = pytest.mark.gpu


def solve_pair(name, tol, every, max_it=500):
    o, g, s = make_pair(name)
    oracle_until(o, "solve_pcg", s["dt"])
    nl = o.dims()["num_levels_relative"]  # the relative mode's levels
    for lvl in range(0, nl + 1):
        to_gpu(o, g, ["FLUID", "DIAG", "INV_DIAG", "PLUS_X", "PLUS_Y"], level=lvl)
    to_gpu(o, g, ["RESIDUAL", "PRESSURE"])
    sys_ = assembled(o)
    o.set_solver(1, tol, max_it, True, every)
    g.set_solver_config(1, tol, max_it, True, every)
    o.stage("solve_pcg")
    g.solve_pcg()
    return o, g, sys_


@pytest.mark.parametrize("name", ["ref64", "dam128"])
@pytest.mark.parametrize("every", [25, 3])
def test_relative_solve_matches_oracle(name, every):
    tol = 1e-6
    o, g, sys_ = solve_pair(name, tol, every)
    so, sg = o.stats(), g.stats()
    print(f"{name} every={every}: iterations GPU {sg['pcg_iterations']} oracle {so['pcg_iterations']}, "
          f"rel residual GPU {sg['pcg_rel_residual']:.3e} oracle {so['pcg_rel_residual']:.3e}")
    assert sg["pcg_exit_reason"] == 3 and so["pcg_exit_reason"] == 3, (sg, so)
    assert abs(int(sg["pcg_iterations"]) - int(so["pcg_iterations"])) <= 1
    assert 0.0 <= sg["pcg_rel_residual"] <= tol
    assert rel_l2(g.read("PRESSURE"), o.read("PRESSURE")) <= 1e-5
    res, bound = residuals(*sys_, g.read("PRESSURE"))
    assert res <= tol + bound, (res, bound)
    res64, _ = residuals(*sys_, g.read("PRESSURE_F64"))
    print(f"{name} every={every}: true residual of the FP64 pressure {res64:.3e}")
    assert res64 <= tol * (1 + 1e-6), res64
    assert abs(res64 - sg["pcg_rel_residual"]) <= 1e-6 * tol, (res64, sg["pcg_rel_residual"])
    g.close()


def test_relative_iteration_cap():
    o, g, _ = solve_pair("dam128", 1e-12, 25, max_it=3)
    assert g.stats()["pcg_iterations"] == 3 and g.stats()["pcg_exit_reason"] == 5
    assert o.stats()["pcg_iterations"] == 3
    g.close()


def test_relative_updates_track_the_oracle():
    """Ten updates with the sort off (no re-ordering chaos): the pressure
    solves agree in iteration count step by step."""
    from FLIP import scenes
    from FLIP.apic import Apic, DebugOptions
    from oracle import pyoracle as po
    xy = scenes.config_scene("cfg1_dam_break_128")
    g = Apic.with_particles(128, 128, 1.0 / 128, props(), xy)
    o = po.Oracle(128, 128, 1.0 / 128, po.bench_props(), xy=xy)
    g.set_solver_config(1, 1e-6, 500, True, 25)
    o.set_solver(1, 1e-6, 500, True, 25)
    its_g, its_o = [], []
    for _ in range(10):
        g.update(DebugOptions(True), 0.002)
        o.update(0.002, disable_sort=True)
        its_g.append(g.stats()["pcg_iterations"])
        its_o.append(o.stats()["pcg_iterations"])
    err = max(rel_l2(g.read(f), o.read(f)) for f in ("POS_X", "POS_Y", "VEL_X", "VEL_Y"))
    print(f"relative mode, 10 updates: iterations GPU {its_g} oracle {its_o}; particle error {err:.2e}")
    assert max(abs(a - b) for a, b in zip(its_g, its_o)) <= 1
    assert err <= 1e-4
    g.close()


def test_solver_config_validation():
    from FLIP._lib import ApicError
    from FLIP.apic import Apic
    g = Apic.new(64, 64, 1.0, props())
    for bad in (dict(mode=2), dict(mode=1, rel_tol=0.0), dict(mode=1, max_iterations=0),
                dict(mode=1, fp64_reductions=False), dict(mode=1, recompute_every=-1)):
        with pytest.raises(ApicError) as e:
            g.set_solver_config(**bad)
        assert e.value.status == 1, bad
    g.set_solver_config(0)
    g.close()
