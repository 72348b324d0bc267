"""Known answers for the oracle functions that the invariant tests leave open:
build_pressure_system (src/apic.rs:1640-1785) -- its three density-smoothing
passes, the divergence and its obstacle masking, the density correction and
the ghost-fluid 1/theta terms -- the G2P advection step (:2057-2264) and the
coarsest level's plain CG with its exits (:618-704).

Every expected value below is worked out by hand from the cited lines for a
small configuration (numbers in the test, not the oracle's formula typed
again), or is a closed form of the operation (eigenvalues of the smoothing
average, a dense solve of the coarse system), chosen so that a dropped term,
a flipped sign or a shifted index changes the answer.  Both the FP64 oracle
and the FP32 oracle are checked, to 1e-6 relative (the FP64 oracle keeps the
reference's f32 constants, 1.15f, 0.1f, 1.001f).
"""
import numpy as np
import pytest

from oracle import pyoracle as po

TOL = {64: 1e-6, 32: 2e-6}  # the FP64 oracle keeps the reference's f32 literals (1.15f, 0.1f, 1.001f)


def oracle(w, h, cell, precision, xy=None):
    if xy is None:
        xy = np.zeros((0, 2), np.float32)
    o = po.Oracle(w, h, cell, po.bench_props(), xy=xy, precision=precision)
    assert o.status == 0
    return o


def fluid_flags(w, h, cells):
    f = np.zeros(w * h, np.uint8)
    for x, y in cells:
        f[y * w + x] = 1
    return f


def near(a, b, precision):
    return abs(float(a) - float(b)) <= TOL[precision] * max(1.0, abs(float(b)))


def padded(o, name, x, y, w):
    return o.read(name)[w + y * w + x]


# ---- build_pressure_system ---------------------------------------------------

@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("rho,b_expected", [(12.35, 0.05), (30.0, 0.225), (10.0, 0.0)])
def test_density_correction_known_answers(precision, rho, b_expected):
    """Every interior cell of an 8 x 8 grid (h = 0.5) fluid, at rest, with a
    uniform density rho: the smoothing keeps it, and b = clamp(rho - 9*1.15, 0,
    9) * 0.1 * h^2: 2.0 * 0.1 * 0.25 = 0.05; the cap 9 * 0.1 * 0.25 = 0.225; 0
    below the threshold.  The operator is the Neumann Laplacian: diagonal =
    number of non-wall neighbours, -1 towards a fluid neighbour right/top, and
    A 1 = 0 on every row."""
    W = H = 8
    o = oracle(W, H, 0.5, precision)
    interior = [(x, y) for y in range(1, 7) for x in range(1, 7)]
    o.write("FLUID", fluid_flags(W, H, interior))
    o.write("DENSITY", np.full(W * H, rho))
    o.stage("build_pressure_system")
    b = o.read("RESIDUAL")
    for x, y in interior:
        assert near(b[W + y * W + x], b_expected, precision), (x, y, b[W + y * W + x])
    # hand diagonals: a corner, an edge and an inner cell
    assert padded(o, "DIAG", 1, 1, W) == 2.0 and padded(o, "DIAG", 1, 3, W) == 3.0
    assert padded(o, "DIAG", 3, 3, W) == 4.0
    assert padded(o, "PLUS_X", 3, 3, W) == -1.0 and padded(o, "PLUS_X", 6, 3, W) == 0.0
    assert padded(o, "PLUS_Y", 3, 6, W) == 0.0 and padded(o, "PLUS_Y", 3, 5, W) == -1.0
    diag, px, py = (o.read(n) for n in ("DIAG", "PLUS_X", "PLUS_Y"))
    for x, y in interior:
        i = W + y * W + x
        assert diag[i] + px[i] + px[i - 1] + py[i] + py[i - W] == 0.0, (x, y)


@pytest.mark.parametrize("precision", [64, 32])
def test_divergence_sign_and_obstacle_masking(precision):
    """u(i, j) = 0.1 i on u faces, v(i, j) = 0.2 j on v faces, 8 x 8 grid,
    h = 0.5, density 9 (no correction), an obstacle at (4, 4).  b = -div * h
    with the faces next to an obstacle taken as zero:
      (3,3): div = 0.1 + 0.2 = 0.3              -> b = -0.15
      (1,3): u(1) masked: 0.2 + 0.2 = 0.4       -> b = -0.2
      (6,3): u(7) masked: -0.6 + 0.2 = -0.4     -> b = +0.2
      (3,1): v(1) masked: 0.1 + 0.4 = 0.5       -> b = -0.25
      (3,6): v(7) masked: 0.1 - 1.2 = -1.1      -> b = +0.55
      (1,1): 0.2 + 0.4 = 0.6                    -> b = -0.3
      (3,4): u(4) masked by the obstacle: -0.3 + 0.2 = -0.1 -> b = +0.05,
             diagonal 3 (no term towards the obstacle), plus_x 0."""
    W = H = 8
    o = oracle(W, H, 0.5, precision)
    interior = [(x, y) for y in range(1, 7) for x in range(1, 7) if (x, y) != (4, 4)]
    obst = o.read("OBSTACLE").copy()
    obst[4 * W + 4] = 1
    o.write("OBSTACLE", obst)
    o.write("FLUID", fluid_flags(W, H, interior))
    o.write("DENSITY", np.full(W * H, 9.0))
    u = np.array([0.1 * i for j in range(H) for i in range(W + 1)])
    v = np.array([0.2 * j for j in range(H + 1) for i in range(W)])
    o.write("MAC_U", u)
    o.write("MAC_V", v)
    o.stage("build_pressure_system")
    expect = {(3, 3): -0.15, (1, 3): -0.2, (6, 3): 0.2, (3, 1): -0.25, (3, 6): 0.55, (1, 1): -0.3, (3, 4): 0.05}
    for (x, y), bv in expect.items():
        assert near(padded(o, "RESIDUAL", x, y, W), bv, precision), ((x, y), padded(o, "RESIDUAL", x, y, W))
    assert padded(o, "DIAG", 3, 4, W) == 3.0 and padded(o, "PLUS_X", 3, 4, W) == 0.0


@pytest.mark.parametrize("precision", [64, 32])
def test_ghost_fluid_theta_terms(precision):
    """One fluid cell (2, 2) in a 6 x 6 grid (h = 1), density 9, its four air
    neighbours at densities 3, 6, 9 and -100 (non-fluid cells keep their raw
    density through the smoothing).  theta = clamp((9 - 4.5) / (9 - rho), 0.1,
    1) or 1 when the densities agree within 1e-5:
      left 3    -> 0.75        -> 1/theta = 4/3   (the arguments swapped: 4)
      right 6   -> 1.5 -> 1    -> 1
      bottom 9  -> equal       -> 1
      top -100  -> 0.0413 -> 0.1 -> 10
    diagonal = 13 1/3, inverse 3/40, no off-diagonal term, b = 0 at rest."""
    W = H = 6
    o = oracle(W, H, 1.0, precision)
    o.write("FLUID", fluid_flags(W, H, [(2, 2)]))
    rho = np.full(W * H, 50.0)
    rho[2 * W + 2], rho[2 * W + 1], rho[2 * W + 3] = 9.0, 3.0, 6.0
    rho[1 * W + 2], rho[3 * W + 2] = 9.0, -100.0
    o.write("DENSITY", rho)
    o.stage("build_pressure_system")
    assert near(padded(o, "DIAG", 2, 2, W), 40.0 / 3.0, precision)
    assert near(padded(o, "INV_DIAG", 2, 2, W), 3.0 / 40.0, precision)
    assert padded(o, "PLUS_X", 2, 2, W) == 0.0 and padded(o, "PLUS_Y", 2, 2, W) == 0.0
    assert padded(o, "RESIDUAL", 2, 2, W) == 0.0
    sm = o.read("SMOOTHED_DENSITY")
    assert sm[2 * W + 2] == 9.0 and sm[3 * W + 2] == -100.0 and sm[0] == 50.0


@pytest.mark.parametrize("precision", [64, 32])
def test_smoothing_closed_forms(precision):
    """The three smoothing passes rho' = (4 rho + sum of fluid neighbours) /
    (4 + their count) are a linear map; on small isolated fluid patterns its
    eigenvectors give the result in closed form.
      - two adjacent cells (a, b) = (20, 2): the mean 11 is kept and the
        difference shrinks by (4 - 1)/5 = 3/5 a pass: 11 +- 9 (3/5)^3 =
        12.944 / 9.056;
      - a 2 x 2 block in a checkerboard (10 +- 9): the alternating mode of the
        4-cycle has eigenvalue (4 - 2)/6 = 1/3: 10 +- 9/27;
      - non-fluid cells keep their raw density (here 100), and do not enter
        their fluid neighbours' averages."""
    W = H = 10
    o = oracle(W, H, 1.0, precision)
    pair = [(2, 2), (3, 2)]
    block = [(6, 6), (7, 6), (6, 7), (7, 7)]
    o.write("FLUID", fluid_flags(W, H, pair + block))
    rho = np.full(W * H, 100.0)
    rho[2 * W + 2], rho[2 * W + 3] = 20.0, 2.0
    for (x, y), s in zip(block, (1, -1, -1, 1)):
        rho[y * W + x] = 10.0 + 9.0 * s
    o.write("DENSITY", rho)
    o.stage("build_pressure_system")
    sm = o.read("SMOOTHED_DENSITY")
    d = 9.0 * (3.0 / 5.0) ** 3
    assert near(sm[2 * W + 2], 11.0 + d, precision) and near(sm[2 * W + 3], 11.0 - d, precision)
    for (x, y), s in zip(block, (1, -1, -1, 1)):
        assert near(sm[y * W + x], 10.0 + s * 9.0 / 27.0, precision), (x, y, sm[y * W + x])
    assert sm[5 * W + 5] == 100.0 and sm[2 * W + 4] == 100.0


# ---- G2P advection -----------------------------------------------------------

# jitter of particle index i at timestamp 0 (lowbias32 of i, pinned by the
# known answers of random_simd in test_oracle_kat.py): lanes 0, 1, 3, 4
RAND0 = {0: -0.5, 1: -0.4252701, 3: 0.3899494, 4: 0.4872686}


@pytest.mark.parametrize("precision", [64, 32])
def test_g2p_advection_known_answers(precision):
    """A 16 x 16 grid (h = 1), u(i, j) = -3 + 0.25 (j + 1/2) on u faces and
    v(i, j) = 0.5 (i + 1/2) on v faces (affine: APIC gives C_u = (0, 0.25) and
    C_v = (0.5, 0) with the unmasked weights), dt = 0.5, timestamp 0.
      P0 (8.3, 8.4): v = (-0.9, 4.15), |v| dt / h > 1 so gamma = 1 and
          dt_act = dt (1 + rand) = 0.25: x = 8.075, y = 9.4375, residual 0.25.
      P1 (1.05, 8.4): the faces on the wall side are closed, the interpolation
          keeps the open ones: v = (-0.9, 0.75); it crosses x = 1.001 h, so x
          is clamped there, u reflected and damped: -0.9 * -0.1 = 0.09, and
          C_u.y and C_v.x change sign; its time is conserved:
          (y' - y) / 0.75 + residual' = dt.
      P3 (5.3, 5.6), residual -2 dt: dt + res + jitter < 0, so dt_act = 0:
          it does not move and its residual becomes dt + res = -0.5.
      P4 (10.2, 6.7), residual 3 dt: dt_act = 2 dt = 1: x = 8.875,
          y = 11.8, residual 0.5 + 1.5 - 1 = 1.
      P5 (8.2, 14.9): the top faces are closed, v = (0.625, 4.1); it
          crosses y = 14.999, so y is clamped, v reflected to -0.41, and the
          same two C components change sign."""
    W = H = 16
    dt = 0.5
    xy = np.array([[8.3, 8.4], [1.05, 8.4], [6.0, 6.0], [5.3, 5.6], [10.2, 6.7], [8.2, 14.9]], np.float32)
    o = oracle(W, H, 1.0, precision, xy=xy)
    u = np.array([-3.0 + 0.25 * (j + 0.5) for j in range(H) for i in range(W + 1)])
    v = np.array([0.5 * (i + 0.5) for j in range(H + 1) for i in range(W)])
    o.write("MAC_U", u)
    o.write("MAC_V", v)
    res = o.read("TIME_RESIDUAL")
    res[3], res[4] = -2 * dt, 3 * dt
    o.write("TIME_RESIDUAL", res)
    o.set_timestamp(0, 1.0 / 60.0)
    o.stage("transfer_grid_to_particles_and_advect", dt)
    X, Y, VX, VY = (o.read(f) for f in ("POS_X", "POS_Y", "VEL_X", "VEL_Y"))
    CUX, CUY, CVX, CVY, R = (o.read(f) for f in ("C_U_X", "C_U_Y", "C_V_X", "C_V_Y", "TIME_RESIDUAL"))
    p = precision
    # P0
    assert near(X[0], 8.075, p) and near(Y[0], 9.4375, p) and near(R[0], 0.25, p)
    assert near(VX[0], -0.9, p) and near(VY[0], 4.15, p)
    assert near(CUX[0], 0.0, p) and near(CUY[0], 0.25, p) and near(CVX[0], 0.5, p) and near(CVY[0], 0.0, p)
    # P1: clamp, reflection, C sign flips, time conservation
    assert near(X[1], 1.001, p) and near(VX[1], 0.09, p) and near(VY[1], 0.75, p)
    assert near(CUY[1], -0.25, p) and near(CVX[1], -0.5, p) and near(CUX[1], 0.0, p) and near(CVY[1], 0.0, p)
    dt_act1 = (Y[1] - 8.4) / 0.75
    assert 0.0 <= dt_act1 <= 2 * dt and near(dt_act1 + R[1], dt, 32)
    assert abs(dt_act1 - dt) <= 0.5 * dt  # |jitter| <= gamma * |rand| * dt <= dt / 2
    # P3: lower clamp of dt_act
    assert X[3] == np.float32(5.3) and Y[3] == np.float32(5.6) and near(R[3], -0.5, p)
    # P4: upper clamp of dt_act
    assert near(X[4], 10.2 - 1.325, p) and near(Y[4], 6.7 + 5.1, p) and near(R[4], 1.0, p)
    # P5: top wall
    assert near(Y[5], 14.999, p) and near(VY[5], -0.41, p) and near(VX[5], 0.625, p)
    assert near(CUY[5], -0.25, p) and near(CVX[5], -0.5, p)


# ---- coarsest-level CG (solve_coarse_exact) --------------------------------

def coarse_oracle(precision, b, fluid5):
    """A 10 x 10 grid has one multigrid level of 5 x 5 cells, the coarsest;
    its operator here is the 5-point Laplacian with a Dirichlet boundary
    (diagonal 4, -1 towards fluid neighbours), b given per cell."""
    o = oracle(10, 10, 1.0, precision, xy=np.array([[5.0, 5.0]], np.float32))
    assert o.dims()["num_levels"] == 1
    w = 5
    n = o.field_length("DIAG", 1)
    diag, px, py, r = (np.zeros(n) for _ in range(4))
    fl = np.zeros(25, np.uint8)
    for (x, y) in fluid5:
        fl[y * w + x] = 1
    for (x, y) in fluid5:
        i = w + y * w + x
        diag[i] = 4.0
        if (x + 1, y) in fluid5:
            px[i] = -1.0
        if (x, y + 1) in fluid5:
            py[i] = -1.0
        r[i] = b[(x, y)]
    o.write("FLUID", fl, 1)
    for name, a in (("DIAG", diag), ("INV_DIAG", np.where(diag > 0, 1.0 / np.maximum(diag, 1), 0.0)),
                    ("PLUS_X", px), ("PLUS_Y", py), ("RESIDUAL", r)):
        o.write(name, a, 1)
    return o, diag, px, py


@pytest.mark.parametrize("precision", [64, 32])
def test_coarse_cg_matches_dense_solve(precision):
    """Plain CG on n <= 100 unknowns, at most n iterations, stopped when max|r|
    < 1e-6 or |s.As| < 1e-5 (src/apic.rs:636-681): with b of size 1e6 the
    absolute exits come late, and the error it returns is the solution of the
    system (numpy.linalg.solve on the dense matrix) to 1e-8 relative (FP64)
    or float32 rounding (FP32)."""
    fluid5 = [(x, y) for y in range(5) for x in range(5) if (x, y) not in ((0, 0), (4, 2), (2, 4))]
    rng = np.random.default_rng(3)
    b = {c: float(v) for c, v in zip(fluid5, rng.uniform(-1e6, 1e6, len(fluid5)))}
    o, diag, px, py = coarse_oracle(precision, b, fluid5)
    o.stage("solve_coarse_exact")
    e = o.read("ERROR", 1)
    idx = {c: k for k, c in enumerate(fluid5)}
    A = np.zeros((len(fluid5), len(fluid5)))
    for (x, y), k in idx.items():
        A[k, k] = 4.0
        for nb in ((x + 1, y), (x - 1, y), (x, y + 1), (x, y - 1)):
            if nb in idx:
                A[k, idx[nb]] = -1.0
    rhs = np.array([b[c] for c in fluid5])
    sol = np.linalg.solve(A, rhs)
    got = np.array([e[5 + y * 5 + x] for (x, y) in fluid5])
    err = np.abs(got - sol).max() / np.abs(sol).max()
    assert err <= (1e-8 if precision == 64 else 1e-5), err


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("bval,moves", [(1e-7, False), (1e-3, False), (1e-2, True)])
def test_coarse_cg_exits(precision, bval, moves):
    """b nonzero on one cell only.  gamma = b^2 < 1e-12 (b = 1e-7): exit
    before the loop, e = 0.  b = 1e-3: gamma = 1e-6 passes, but delta = s.As
    = 4 b^2 = 4e-6 < 1e-5 ends the loop before the first update, e = 0.
    b = 1e-2: delta = 4e-4, the solve moves."""
    fluid5 = [(x, y) for y in range(5) for x in range(5)]
    b = {c: 0.0 for c in fluid5}
    b[(2, 2)] = bval
    o, *_ = coarse_oracle(precision, b, fluid5)
    o.stage("solve_coarse_exact")
    e = o.read("ERROR", 1)
    assert bool(np.any(e != 0.0)) == moves, e[5:30]


# ---- the relative mode's coarse CG ------------------------------------------
# solve_coarse_exact's iteration with FP64 dots, a search vector that is zero
# off the fluid cells, and exits relative to the level's right-hand side
# (r.r = 0 at the start, s.As <= 0, ||r|| <= 1e-2 ||r0||), which the relative
# mode runs on its coarsest level whatever its size.

def _halvings_up(m):
    """Sides of the levels below an m x m one: halved, rounded up, while the
    level has more than 64 cells and its side exceeds 2."""
    out = []
    while m * m > 64 and m > 2:
        m = (m + 1) // 2
        out.append(m)
    return out


def square_coarse_oracle(precision, m, diag_value, b):
    """A (2m) x (2m) grid with m odd: one multigrid level of m x m cells, the
    reference's coarsest, every cell fluid; operator diag_value on the
    diagonal, -1 towards neighbours (Dirichlet outside), b per cell.  (The
    relative mode's deeper levels, below it when m^2 > 64, are not touched by
    the coarse-solve stages.)"""
    o = oracle(2 * m, 2 * m, 1.0, precision, xy=np.array([[float(m), float(m)]], np.float32))
    assert o.dims()["num_levels"] == 1
    assert o.dims()["num_levels_relative"] == 1 + len(_halvings_up(m))
    n = o.field_length("DIAG", 1)
    diag, px, py, r = (np.zeros(n) for _ in range(4))
    idx = m + np.arange(m * m)
    x = np.arange(m * m) % m
    y = np.arange(m * m) // m
    diag[idx] = diag_value
    px[idx] = np.where(x + 1 < m, -1.0, 0.0)
    py[idx] = np.where(y + 1 < m, -1.0, 0.0)
    r[idx] = b
    o.write("FLUID", np.ones(m * m, np.uint8), 1)
    for name, a in (("DIAG", diag), ("INV_DIAG", np.where(diag > 0, 1.0 / np.maximum(diag, 1), 0.0)),
                    ("PLUS_X", px), ("PLUS_Y", py), ("RESIDUAL", r)):
        o.write(name, a, 1)
    A = np.zeros((m * m, m * m))
    for k in range(m * m):
        A[k, k] = diag_value
        for j, ok in ((k + 1, x[k] + 1 < m), (k - 1, x[k] > 0), (k + m, y[k] + 1 < m), (k - m, y[k] > 0)):
            if ok:
                A[k, j] = -1.0
    return o, A


def _rel_residual(A, b, e):
    return np.linalg.norm(b - A @ e) / np.linalg.norm(b)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("bval", [1e-7, 1e-3])
def test_relative_coarse_cg_moves_where_absolute_exits_stop(precision, bval):
    """The single-cell right-hand sides on which the reference's exits return
    e = 0 (test_coarse_cg_exits): the relative exits run the solve until the
    residual ||b - A e|| (evaluated here in FP64 from the dense matrix) is a
    hundredth of ||b||, for b = 1e-7 as for 1e-3."""
    m = 5
    b = np.zeros(m * m)
    b[12] = bval
    o, A = square_coarse_oracle(precision, m, 4.0, b)
    o.stage("solve_coarse_relative")
    e = o.read("ERROR", 1)[m:m + m * m].astype(np.float64)
    assert np.any(e != 0.0)
    assert _rel_residual(A, b, e) <= 1e-2 * (1 + 1e-3)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("m", [5, 35])
def test_relative_coarse_cg_is_scale_invariant(precision, m):
    """With exits relative to the right-hand side, scaling b by a power of two
    scales every operation of the solve exactly: e(2^-20 b) = 2^-20 e(b) bit
    for bit (the absolute exits would stop at once on the small one:
    |s.As| < 1e-5)."""
    rng = np.random.default_rng(7)
    b = rng.uniform(-1.0, 1.0, m * m)
    outs = []
    for scale in (1.0, 2.0 ** -20):
        o, A = square_coarse_oracle(precision, m, 4.0, b * scale)
        o.stage("solve_coarse_relative")
        outs.append(o.read("ERROR", 1)[m:m + m * m].astype(np.float64))
    assert np.array_equal(outs[1], outs[0] * 2.0 ** -20)
    assert _rel_residual(A, b, outs[0]) <= 1e-2 * (1 + 1e-3)


@pytest.mark.parametrize("precision", [64, 32])
def test_relative_coarse_cg_on_a_well_conditioned_system(precision):
    """On a system with condition number below 3 (diagonal 8) the hundredfold
    residual reduction is an error below 3e-2 of the dense solution, and CG
    reaches it in a few iterations; on 35 x 35 Dirichlet Poisson (condition
    number about 500) it is reached within the 100-iteration cap."""
    rng = np.random.default_rng(11)
    m = 35
    b = rng.uniform(-1.0, 1.0, m * m)
    o, A = square_coarse_oracle(precision, m, 8.0, b)
    o.stage("solve_coarse_relative")
    e = o.read("ERROR", 1)[m:m + m * m].astype(np.float64)
    sol = np.linalg.solve(A, b)
    assert np.abs(e - sol).max() <= 3e-2 * np.abs(sol).max()
    o, A = square_coarse_oracle(precision, m, 4.0, b)
    o.stage("solve_coarse_relative")
    e = o.read("ERROR", 1)[m:m + m * m].astype(np.float64)
    assert _rel_residual(A, b, e) <= 1e-2 * (1 + 1e-3)
