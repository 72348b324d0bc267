"""Pins of the oracle by invariants, closed forms and independent library routines.

What each test pins (reference lines):
  sort        std::stable_sort order == numpy's stable argsort of the Morton
              keys of trunc(clamp(x/dx)) (src/apic.rs:1043-1111)
  P2G         total density == sum of temporal weights (bilinear weights sum
              to one; clamped corners keep their weight) (:1237-1426)
  P2G->G2P    one interior particle round-trips u_p and C (APIC with linear
              weights reproduces affine fields) (:1237-1426, :2057-2264)
  G2P         an affine grid field is interpolated exactly and C equals its
              gradient (:2139-2247)
  MG build    coarse operator == 1/2 P^T A P for piecewise-constant P (:1499-1602)
  restrict    coarse r == P^T (b - A e) on coarse fluid cells (:537-594)
  prolongate  e_f += e_c(parent), the base level's last fluid cell 1 + pad
              times (the padded-list quirk, :1487-1494, :596-616)
  V-cycle     symmetric operator when the padded-list quirk is inactive
              (:1788-1821, red-black forward / black-red reverse sweeps)
  PCG         converges to the dense solution of the assembled system (:1865-1956)
  extrapolate one valid face spreads to its Manhattan diamond of radius 4 (:2266-2350)
  gradient    face updates by hand on a small grid (:1958-2055)
  forces      gravity on v faces next to fluid only (:1604-1626)

Parity unpinned (no closed form, invariant or runnable reference pins the bit
pattern): the summation orders of the reference's parallel CAS scatters and
rayon reductions, wide's reduce_add lane order, and rdst's tie order.  These
are checked for consistency only (FP32 vs FP64, shuffled scatter order).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import pyoracle as po


def make(w, h, cell=1.0, xy=None, precision=32, solve_error=1e-5):
    o = po.Oracle(w, h, cell, po.bench_props(solve_error), xy=xy, precision=precision)
    assert o.status == 0
    return o


def morton_np(x, y):
    k = np.zeros_like(x, dtype=np.uint64)
    for b in range(16):
        k |= ((x >> b) & 1).astype(np.uint64) << np.uint64(2 * b)
        k |= ((y >> b) & 1).astype(np.uint64) << np.uint64(2 * b + 1)
    return k.astype(np.uint32)


def test_sort_is_stable_morton_order():
    rng = np.random.default_rng(1)
    xy = rng.uniform(1.0, 31.0, size=(3001, 2)).astype(np.float32)
    xy[100:200] = xy[100]  # many equal keys
    o = make(32, 32, 1.0, xy=xy)
    o.stage("update_spatial_lookup")
    ix = np.clip(xy[:, 0], 0, 31).astype(np.int64)
    iy = np.clip(xy[:, 1], 0, 31).astype(np.int64)
    keys = morton_np(ix, iy)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(o.read("SORT_KEY"), keys[order])
    assert np.array_equal(o.read("SORT_VAL"), order.astype(np.uint32))
    n = xy.shape[0]
    assert np.array_equal(o.read("POS_X")[:n], xy[order, 0].astype(np.float64))
    # tail lanes of the last chunk are zeroed by the gather (src/apic.rs:1077-1103)
    assert np.all(o.read("POS_X")[n:] == 0.0)


def test_p2g_density_conserves_temporal_weight():
    rng = np.random.default_rng(2)
    n = 2000
    xy = rng.uniform(0.5, 31.5, size=(n, 2)).astype(np.float32)
    o = make(32, 32, 1.0, xy=xy)
    tres = rng.uniform(-0.02, 0.02, size=o.field_length("TIME_RESIDUAL")).astype(np.float32)
    o.write("TIME_RESIDUAL", tres)
    o.stage("transfer_particles_to_grid")
    prev = np.float32(1.0 / 60.0)
    wt = np.array([po.temporal_w(np.float32(-t) / prev) for t in tres[:n]], dtype=np.float64)
    assert abs(o.read("DENSITY").sum() - wt.sum()) <= 1e-5 * wt.sum()
    # every face with weight is valid; valid faces cover all stencils
    assert o.read("VALID_U").sum() > 0 and o.read("VALID_V").sum() > 0


@pytest.mark.parametrize("pos", [(7.3, 8.6), (10.55, 4.2), (5.9, 11.45)])
def test_single_particle_apic_round_trip(pos):
    o = make(16, 16, 1.0, xy=np.array([pos], dtype=np.float32))
    n8 = o.field_length("VEL_X")
    vals = {"VEL_X": 0.7, "VEL_Y": -0.4, "C_U_X": 0.3, "C_U_Y": -0.2, "C_V_X": 0.1, "C_V_Y": 0.25}
    for k, v in vals.items():
        o.write(k, np.full(n8, v))
    o.stage("transfer_particles_to_grid")
    o.stage("transfer_grid_to_particles_and_advect", 1e-4)
    for k, v in vals.items():
        got = o.read(k)[0]
        assert abs(got - v) <= 2e-5, (k, got, v)


def test_g2p_reproduces_affine_field():
    rng = np.random.default_rng(4)
    xy = rng.uniform(3.0, 13.0, size=(64, 2)).astype(np.float32)
    o = make(16, 16, 1.0, xy=xy)
    W = H = 16
    a, b, c = 0.5, 0.25, -0.125
    d, e, f = -0.3, 0.0625, 0.375
    iu, ju = np.meshgrid(np.arange(W + 1), np.arange(H), indexing="xy")
    o.write("MAC_U", (a + b * iu + c * (ju + 0.5)).reshape(-1))
    iv, jv = np.meshgrid(np.arange(W), np.arange(H + 1), indexing="xy")
    o.write("MAC_V", (d + e * (iv + 0.5) + f * jv).reshape(-1))
    o.stage("transfer_grid_to_particles_and_advect", 1e-4)
    n = xy.shape[0]
    vx, vy = o.read("VEL_X")[:n], o.read("VEL_Y")[:n]
    assert np.allclose(vx, a + b * xy[:, 0] + c * xy[:, 1], atol=1e-5)
    assert np.allclose(vy, d + e * xy[:, 0] + f * xy[:, 1], atol=1e-5)
    assert np.allclose(o.read("C_U_X")[:n], b, atol=2e-4) and np.allclose(o.read("C_U_Y")[:n], c, atol=2e-4)
    assert np.allclose(o.read("C_V_X")[:n], e, atol=2e-4) and np.allclose(o.read("C_V_Y")[:n], f, atol=2e-4)


def fluid_state(steps=60):
    """The reference constructor's 64^2 scene after impact, stepped through pressure assembly."""
    o = make(64, 64, 1.0)
    for _ in range(steps):
        o.update(0.016)
    for st in ("update_spatial_lookup", "transfer_particles_to_grid"):
        o.stage(st)
    o.stage("apply_external_forces", 0.016)
    o.stage("build_pressure_system")
    o.stage("update_multilevel_grid")
    return o


def level_matrix(o, level):
    """Sparse operator of one level from its padded coefficient arrays, over raw cells."""
    w = 64 >> level
    fl = o.read("FLUID", level).astype(bool)
    n = fl.size
    diag, px, py = o.read("DIAG", level), o.read("PLUS_X", level), o.read("PLUS_Y", level)
    rows, cols, vals = [], [], []
    for c in np.nonzero(fl)[0]:
        p = c + w
        rows.append(c); cols.append(c); vals.append(diag[p])
        for nb, coef in ((c + 1, px[p]), (c - 1, px[p - 1]), (c + w, py[p]), (c - w, py[p - w])):
            if coef != 0.0:
                rows.append(c); cols.append(nb); vals.append(coef)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n)), fl


def test_coarse_operator_is_half_galerkin():
    o = fluid_state()
    for level in range(0, o.dims()["num_levels"]):
        A, fl = level_matrix(o, level)
        w = 64 >> level
        cw = w // 2
        ci = (np.arange(fl.size) // w // 2) * cw + (np.arange(fl.size) % w) // 2
        fine = np.nonzero(fl)[0]
        P = sp.csr_matrix((np.ones(fine.size), (fine, ci[fine])), shape=(fl.size, fl.size // 4))
        G = (0.5 * (P.T @ A @ P)).toarray()
        Ac, flc = level_matrix(o, level + 1)
        assert np.array_equal(flc, np.bincount(ci[fine], minlength=fl.size // 4) > 0)
        Acd = Ac.toarray()
        idx = np.nonzero(flc)[0]
        assert np.allclose(Acd[np.ix_(idx, idx)], G[np.ix_(idx, idx)], rtol=1e-5, atol=1e-5), level


def test_restrict_is_transpose_of_injection():
    o = fluid_state()
    A, fl = level_matrix(o, 0)
    rng = np.random.default_rng(5)
    n = fl.size
    e = np.where(fl, rng.standard_normal(n), 0.0).astype(np.float32).astype(np.float64)
    b = np.where(fl, rng.standard_normal(n), 0.0).astype(np.float32).astype(np.float64)
    pad = np.zeros(o.field_length("ERROR")); pad[64:64 + n] = e
    o.write("ERROR", pad)
    pad = np.zeros(o.field_length("RESIDUAL")); pad[64:64 + n] = b
    o.write("RESIDUAL", pad)
    o.stage("restrict_base")
    ci = (np.arange(n) // 64 // 2) * 32 + (np.arange(n) % 64) // 2
    fine = np.nonzero(fl)[0]
    P = sp.csr_matrix((np.ones(fine.size), (fine, ci[fine])), shape=(n, n // 4))
    expect = P.T @ (b - A @ e)
    flc = o.read("FLUID", 1).astype(bool)
    got = o.read("RESIDUAL", 1)[32:32 + n // 4]
    assert np.allclose(got[flc], expect[flc], rtol=1e-4, atol=1e-4)
    assert np.all(o.read("ERROR", 1)[32:32 + n // 4] == 0.0)


def test_prolongate_padded_list_quirk():
    o = fluid_state()
    fl = o.read("FLUID", 0).astype(bool)
    nf = int(fl.sum())
    n = fl.size
    rng = np.random.default_rng(6)
    ec = rng.standard_normal(n // 4).astype(np.float32).astype(np.float64)
    pad = np.zeros(o.field_length("ERROR", 1)); pad[32:32 + n // 4] = ec
    o.write("ERROR", pad, 1)
    o.write("ERROR", np.zeros(o.field_length("ERROR", 0)), 0)
    o.stage("prolongate_base")
    got = o.read("ERROR", 0)[64:64 + n]
    ci = (np.arange(n) // 64 // 2) * 32 + (np.arange(n) % 64) // 2
    expect = np.where(fl, ec[ci], 0.0)
    last = int(np.nonzero(fl)[0][-1])
    reps = 1 + (8 - nf % 8) % 8
    acc = np.float32(0.0)
    for _ in range(reps):
        acc = np.float32(acc + np.float32(ec[ci[last]]))
    expect[last] = acc
    assert np.array_equal(got.astype(np.float32), expect.astype(np.float32))


def rect_system(cells_x, cells_y, x0=6, y0=6, W=32, H=32):
    """A rectangle of fluid cells written through the FLUID field, with zero velocity
    and density 9, assembled and coarsened by the oracle."""
    o = make(W, H, 1.0, xy=np.array([[x0 + 0.5, y0 + 0.5]], dtype=np.float32))
    fl = np.zeros(W * H, dtype=np.uint8)
    for y in range(y0, y0 + cells_y):
        fl[y * W + x0:y * W + x0 + cells_x] = 1
    o.write("FLUID", fl)
    o.write("DENSITY", np.where(fl > 0, 9.0, 0.0))
    o.stage("build_pressure_system")
    o.stage("update_multilevel_grid")
    return o, fl.astype(bool)


def precondition(o, r, W=32):
    pad = np.zeros(o.field_length("RESIDUAL")); pad[W:W + r.size] = r
    o.write("RESIDUAL", pad)
    o.stage("apply_preconditioner")
    return o.read("ERROR")[W:W + r.size]


def test_vcycle_symmetric_without_padding_quirk():
    o, fl = rect_system(8, 10)  # 80 fluid cells: the padded list has no duplicates
    rng = np.random.default_rng(8)
    r1 = np.where(fl, rng.standard_normal(fl.size), 0.0).astype(np.float32).astype(np.float64)
    r2 = np.where(fl, rng.standard_normal(fl.size), 0.0).astype(np.float32).astype(np.float64)
    z1, z2 = precondition(o, r1), precondition(o, r2)
    a, b = float(r2 @ z1), float(r1 @ z2)
    assert abs(a - b) <= 1e-5 * max(abs(a), abs(b)), (a, b)
    assert float(r1 @ z1) > 0 and float(r2 @ z2) > 0  # positive definite on these samples


def test_vcycle_asymmetric_with_padding_quirk():
    o, fl = rect_system(9, 9)  # 81 fluid cells: the last cell is prolongated 8 times
    rng = np.random.default_rng(9)
    r1 = np.where(fl, rng.standard_normal(fl.size), 0.0).astype(np.float32).astype(np.float64)
    r2 = np.where(fl, rng.standard_normal(fl.size), 0.0).astype(np.float32).astype(np.float64)
    z1, z2 = precondition(o, r1), precondition(o, r2)
    a, b = float(r2 @ z1), float(r1 @ z2)
    assert abs(a - b) > 1e-4 * max(abs(a), abs(b)), (a, b)


def test_pcg_matches_dense_solve():
    o, fl = rect_system(10, 8)
    A, _ = level_matrix_w(o, 32)
    rng = np.random.default_rng(10)
    b = np.where(fl, rng.standard_normal(fl.size), 0.0).astype(np.float32).astype(np.float64)
    pad = np.zeros(o.field_length("RESIDUAL")); pad[32:32 + fl.size] = b
    o.write("RESIDUAL", pad)
    o.write("PRESSURE", np.zeros(fl.size))
    o.stage("solve_pcg")
    st = o.stats()
    p = o.read("PRESSURE")
    idx = np.nonzero(fl)[0]
    Ad = A.toarray()[np.ix_(idx, idx)]
    x = np.linalg.solve(Ad, b[idx])
    res = b[idx] - Ad @ p[idx]
    assert st["pcg_exit_reason"] in (3, 4), st
    assert np.abs(res).max() < 1e-3, np.abs(res).max()
    assert np.linalg.norm(p[idx] - x) <= 1e-3 * np.linalg.norm(x)
    # the recursively updated residual tracks the true one
    r = o.read("RESIDUAL")[32:32 + fl.size][idx]
    assert np.abs(r - res).max() < 1e-4


def level_matrix_w(o, w):
    fl = o.read("FLUID", 0).astype(bool)
    n = fl.size
    diag, px, py = o.read("DIAG"), o.read("PLUS_X"), o.read("PLUS_Y")
    rows, cols, vals = [], [], []
    for c in np.nonzero(fl)[0]:
        p = c + w
        rows.append(c); cols.append(c); vals.append(diag[p])
        for nb, coef in ((c + 1, px[p]), (c - 1, px[p - 1]), (c + w, py[p]), (c - w, py[p - w])):
            if coef != 0.0:
                rows.append(c); cols.append(nb); vals.append(coef)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n)), fl


def test_assembled_operator_symmetric_and_dominant():
    o = fluid_state()
    A, fl = level_matrix(o, 0)
    assert abs(A - A.T).max() == 0.0
    d = A.diagonal()
    off = np.asarray(abs(A).sum(axis=1)).ravel() - np.abs(d)
    idx = np.nonzero(fl)[0]
    assert np.all(d[idx] >= off[idx] - 1e-6)


def test_extrapolation_diamond():
    o = make(20, 20, 1.0)
    W = H = 20
    vu = np.zeros((H, W + 1), dtype=np.uint8)
    mu = np.zeros((H, W + 1))
    vu[10, 10] = 1
    mu[10, 10] = 3.5
    o.write("VALID_U", vu.reshape(-1)); o.write("MAC_U", mu.reshape(-1))
    o.write("VALID_V", np.zeros(o.field_length("VALID_V"), dtype=np.uint8))
    o.write("MAC_V", np.zeros(o.field_length("MAC_V")))
    o.stage("extrapolate_velocity")
    jj, ii = np.meshgrid(np.arange(H), np.arange(W + 1), indexing="ij")
    diamond = (np.abs(jj - 10) + np.abs(ii - 10)) <= 4
    got_v = o.read("VALID_U").reshape(H, W + 1).astype(bool)
    got_u = o.read("MAC_U").reshape(H, W + 1)
    assert np.array_equal(got_v, diamond)
    assert np.all(got_u[diamond] == 3.5) and np.all(got_u[~diamond] == 0.0)
    assert not o.read("VALID_V").any()


def test_pressure_gradient_by_hand():
    o = make(8, 8, 0.5)
    W, H, dx = 8, 8, 0.5
    fl = np.zeros(W * H, dtype=np.uint8)
    fl[3 * W + 3] = fl[3 * W + 4] = 1  # two fluid cells side by side
    o.write("FLUID", fl)
    p = np.zeros(W * H); p[3 * W + 3] = 2.0; p[3 * W + 4] = 5.0
    o.write("PRESSURE", p)
    rho = np.full(W * H, 9.0); rho[3 * W + 5] = 3.0  # air cell right of (4,3)
    o.write("SMOOTHED_DENSITY", rho)
    o.write("MAC_U", np.ones(o.field_length("MAC_U")))
    o.write("MAC_V", np.ones(o.field_length("MAC_V")))
    o.stage("apply_pressure_gradient")
    u = o.read("MAC_U").reshape(H, W + 1)
    v = o.read("MAC_V").reshape(H + 1, W)
    assert u[3, 4] == np.float32(1.0 - (5.0 - 2.0) * (1.0 / dx))          # fluid | fluid
    th = np.float32(min(max((9.0 - 4.5) / (9.0 - 3.0), 0.1), 1.0))
    assert u[3, 5] == np.float32(1.0 - np.float32(-5.0 / th) * (1.0 / dx))  # fluid | air
    assert u[3, 3] == np.float32(1.0 - np.float32(2.0 / 1.0) * (1.0 / dx))  # air | fluid, equal densities
    assert u[5, 5] == 1.0                                                # air | air
    assert np.all(u[1:H - 1, 1] == 0.0) and np.all(u[1:H - 1, W - 1] == 0.0)  # wall | interior
    assert np.all(u[0, :] == 1.0)                                        # wall | wall: unchanged
    assert np.all(v[1, 1:W - 1] == 0.0) and np.all(v[H - 1, 1:W - 1] == 0.0)
    assert v[3, 3] == np.float32(1.0 - 2.0 * 2.0)                        # air below fluid (3,3)
    assert np.all(o.read("VALID_U").reshape(H, W + 1)[1:H - 1, 1])


def test_external_forces_gravity_next_to_fluid():
    o = make(8, 8, 1.0)
    W, H = 8, 8
    fl = np.zeros(W * H, dtype=np.uint8); fl[4 * W + 3] = 1
    o.write("FLUID", fl)
    o.write("MAC_U", np.zeros(o.field_length("MAC_U")))
    o.write("MAC_V", np.zeros(o.field_length("MAC_V")))
    o.stage("apply_external_forces", 0.01)
    v = o.read("MAC_V").reshape(H + 1, W)
    g = np.float32(np.float32(-9.81) * np.float32(0.01))
    expect = np.zeros((H + 1, W)); expect[4, 3] = g; expect[5, 3] = g
    assert np.array_equal(v.astype(np.float32), expect.astype(np.float32))
    assert not o.read("MAC_U").any()


def test_fp32_against_fp64_oracle():
    from FLIP import scenes
    xy = scenes.config_scene("cfg1_dam_break_128")
    a = make(128, 128, 1.0 / 128, xy=xy, precision=32)
    b = make(128, 128, 1.0 / 128, xy=xy, precision=64)
    for _ in range(10):
        a.update(0.002)
        b.update(0.002)
    for f in ("POS_X", "POS_Y"):
        x, y = a.read(f), b.read(f)
        assert np.linalg.norm(x - y) <= 1e-5 * np.linalg.norm(y), f


def test_particles_stay_inside_walls():
    o = make(64, 64, 1.0)
    for _ in range(80):
        o.update(0.016)
    x, y = o.read("POS_X"), o.read("POS_Y")
    lo, hi = np.float32(1.001), np.float32((64 - 1.001))
    assert x.min() >= lo and x.max() <= hi and y.min() >= lo and y.max() <= hi
    ts, pd = o.timestamp()
    assert ts == 80 and abs(pd - 0.016) < 1e-9
