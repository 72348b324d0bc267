"""Shared helpers: copy state between the oracle and the GPU library, compare fields."""
import numpy as np

PARTICLE_FIELDS = ["POS_X", "POS_Y", "VEL_X", "VEL_Y", "C_U_X", "C_U_Y", "C_V_X", "C_V_Y", "TIME_RESIDUAL"]
SOLVER_FIELDS = ["DIAG", "INV_DIAG", "PLUS_X", "PLUS_Y", "RESIDUAL", "ERROR"]


def props(solve_error=1e-5):
    from FLIP.apic import WorldProperties
    return WorldProperties(-9.81, 0.1, 5.0, solve_error)


def to_gpu(o, g, names, level=0):
    """Copy oracle fields into the GPU simulation (float64 -> float32 is exact for f32 oracles)."""
    for n in names:
        g.write(n, o.read(n, level), level)


def copy_state(o, g):
    """Copy the particle state, timestamp and prev_deltatime."""
    to_gpu(o, g, PARTICLE_FIELDS)
    ts, pd = o.timestamp()
    g.set_timestamp(ts, pd)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    if den == 0.0:
        return float(np.linalg.norm(a - b))
    return float(np.linalg.norm(a - b) / den)


def bit_equal(a, b):
    """Equality of float32 values (treating +0 and -0 as equal)."""
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def first_mismatch(a, b):
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    bad = np.nonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))[0]
    if bad.size == 0:
        return None
    i = int(bad[0])
    return f"{bad.size} mismatches; first at {i}: gpu={a[i]!r} oracle={b[i]!r}"


def fluid_padded_mask(o, level=0):
    """Boolean mask over a padded solver array that selects the level's fluid cells."""
    fl = o.read("FLUID", level).astype(bool)
    w = o.read("FLUID", level).size
    dims_w = None
    # padded index = raw + width; width recovered from the padded length
    n = o.field_length("DIAG", level)
    return fl, n


def advance_oracle(o, steps, dt):
    for _ in range(steps):
        o.update(dt)
