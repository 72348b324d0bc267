"""The oracle's relative solver mode (an extension: the reference's
solve_pcg, src/apic.rs:1865-1956, has absolute exits only), pinned by the
mathematics rather than by the restatement itself:

  - the returned pressure satisfies ||b - A p||_2 <= rel_tol * ||b||_2 for the
    assembled system (A from DIAG/PLUS_X/PLUS_Y, computed here in FP64): the
    FP64-accumulated pressure exactly, the FP32 one up to its rounding (the
    exit certifies the true residual with a bound on the FP32 rounding of the
    recurrence residual);
  - FP32 vectors with FP64 dot products agree with the all-FP64 oracle;
  - a looser tolerance never needs more iterations than a tighter one;
  - recomputing the true residual (FP64, against the FP64-accumulated
    pressure) every few iterations changes the result only at rounding level
    and does not derail the iteration;
  - mode 0 (the reference's rules) is untouched by the extension.
"""
import numpy as np
import pytest

from helpers import rel_l2


def assembled(o):
    """b (RESIDUAL before the solve), the padded coefficients and the fluid mask."""
    d = o.dims()
    W, H = d["width"], d["height"]
    return (W, H, o.read("RESIDUAL").astype(np.float64), o.read("DIAG").astype(np.float64),
            o.read("PLUS_X").astype(np.float64), o.read("PLUS_Y").astype(np.float64),
            o.read("FLUID").astype(bool))


def residuals(W, H, b, diag, px, py, fluid, p):
    """||b - A p||_2 / ||b||_2 in FP64, and the FP32 rounding bound of a
    residual evaluated in single precision: 8 eps32 || |A| |p| ||_2 / ||b||_2."""
    n = W * H
    pp = np.zeros(n + 2 * W)
    pp[W:W + n] = p.astype(np.float64)[:n]
    idx = np.nonzero(fluid)[0] + W
    terms = [diag[idx] * pp[idx], px[idx] * pp[idx + 1], px[idx - 1] * pp[idx - 1],
             py[idx] * pp[idx + W], py[idx - W] * pp[idx - W]]
    r = b[idx] - sum(terms)
    nb = np.linalg.norm(b[idx])
    bound = 8 * 2.0 ** -24 * np.linalg.norm(sum(np.abs(t) for t in terms)) / nb
    return float(np.linalg.norm(r) / nb), float(bound)


def oracle_at_solve(precision=32):
    from FLIP import scenes
    from oracle import pyoracle as po
    xy = scenes.config_scene("cfg1_dam_break_128")
    o = po.Oracle(128, 128, 1.0 / 128, po.bench_props(), xy=xy, precision=precision)
    for _ in range(20):
        o.update(0.002)
    for st in ("update_spatial_lookup", "transfer_particles_to_grid", "apply_external_forces",
               "build_pressure_system", "update_multilevel_grid"):
        o.stage(st, 0.002)
    return o


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("tol", [1e-3, 1e-5, 1e-6, 1e-7, 1e-8])
def test_relative_solve_meets_the_tolerance(tol, precision):
    """FP64: the true residual meets the tolerance.  FP32 vectors: it meets the
    tolerance up to the rounding of a single-precision residual."""
    o = oracle_at_solve(precision)
    sys_ = assembled(o)
    o.set_solver(1, tol, 500, True, 25)
    o.stage("solve_pcg")
    s = o.stats()
    assert s["pcg_exit_reason"] == 3, s
    assert s["pcg_rel_residual"] <= tol
    res, bound = residuals(*sys_, o.read("PRESSURE"))
    slack = 1e-3 * tol if precision == 64 else bound
    assert res <= tol + slack, (tol, res, bound, s["pcg_rel_residual"])
    # the residual the test reads, r + r_lo, is the true residual of the
    # FP64-accumulated pressure up to FP64 rounding, whatever the precision
    # of the vectors: the true residual meets the tolerance and the reported
    # one agrees with it to a millionth of the tolerance
    res64, _ = residuals(*sys_, o.read("PRESSURE_F64"))
    assert res64 <= tol * (1 + 1e-6), (tol, res64)
    assert abs(res64 - s["pcg_rel_residual"]) <= 1e-6 * tol, (res64, s["pcg_rel_residual"])


def test_fp32_vectors_fp64_dots_match_fp64_oracle():
    o32, o64 = oracle_at_solve(32), oracle_at_solve(64)
    for o in (o32, o64):
        o.set_solver(1, 1e-6, 500, True, 25)
        o.stage("solve_pcg")
    i32, i64 = o32.stats()["pcg_iterations"], o64.stats()["pcg_iterations"]
    assert abs(i32 - i64) <= 1, (i32, i64)
    assert rel_l2(o32.read("PRESSURE"), o64.read("PRESSURE")) <= 1e-5


def test_iterations_monotone_in_tolerance():
    its = []
    for tol in (1e-2, 1e-4, 1e-6):
        o = oracle_at_solve()
        o.set_solver(1, tol, 500, True, 25)
        o.stage("solve_pcg")
        its.append(o.stats()["pcg_iterations"])
    assert its == sorted(its) and its[0] < its[-1], its


def test_residual_recomputation_is_rounding_level():
    ps, its = [], []
    for every in (0, 3):
        o = oracle_at_solve()
        o.set_solver(1, 1e-6, 500, True, every)
        o.stage("solve_pcg")
        ps.append(o.read("PRESSURE"))
        its.append(o.stats()["pcg_iterations"])
        assert o.stats()["pcg_exit_reason"] == 3
    assert rel_l2(ps[1], ps[0]) <= 1e-5
    assert abs(its[1] - its[0]) <= 1


def test_iteration_cap():
    o = oracle_at_solve()
    o.set_solver(1, 1e-12, 4, True, 25)
    o.stage("solve_pcg")
    s = o.stats()
    assert s["pcg_iterations"] == 4 and s["pcg_exit_reason"] == 5


def test_reference_mode_unchanged():
    a, b = oracle_at_solve(), oracle_at_solve()
    b.set_solver(1, 1e-6, 500, True, 25)
    b.set_solver(0)
    a.stage("solve_pcg")
    b.stage("solve_pcg")
    assert np.array_equal(a.read("PRESSURE"), b.read("PRESSURE"))
    assert a.stats()["pcg_iterations"] == b.stats()["pcg_iterations"]
