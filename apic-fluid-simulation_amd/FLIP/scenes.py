"""Seeded scene construction for the dam-break workloads (host-side input data).

Particles of a fluid cell are placed at (cell + U[0,1)^2) * dx with U drawn
from SplitMix64 (Steele, Lea & Flood, 2014) as a 24-bit mantissa,
(r >> 40) * 2^-24; cells are enumerated y-outer, x-inner as the reference's
constructor does (src/apic.rs:832-845).  A fluid cell is an interior cell
(not a wall) whose centre lies in the block.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, offset=0):
    """The first n outputs of SplitMix64 seeded with `seed` (after `offset` draws)."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform24(seed, n):
    r = splitmix64(seed, n)
    return ((r >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))).astype(np.float32)


def block_cells(width, height, x0, x1, y0, y1):
    """Interior cells (x, y) whose centres (in cell units) lie in [x0,x1) x [y0,y1), y-outer order."""
    ys, xs = np.meshgrid(np.arange(1, height - 1), np.arange(1, width - 1), indexing="ij")
    cx, cy = xs + 0.5, ys + 0.5
    m = (cx >= x0) & (cx < x1) & (cy >= y0) & (cy < y1)
    return xs[m].astype(np.int64), ys[m].astype(np.int64)


def seed_cells(xs, ys, ppc, dx, seed):
    """ppc jittered particles per listed cell; returns float32 (N, 2) positions."""
    n = xs.size * ppc
    u = uniform24(seed, 2 * n).reshape(n, 2)
    cx = np.repeat(xs, ppc).astype(np.float32)
    cy = np.repeat(ys, ppc).astype(np.float32)
    dx = np.float32(dx)
    px = (cx + u[:, 0]) * dx
    py = (cy + u[:, 1]) * dx
    return np.stack([px, py], axis=1).astype(np.float32)


def dam_break(width, height, dx, fill_x=0.4, fill_y=0.6, ppc=4, seed=1):
    """Dam break: the block fills the left fill_x of the width and bottom fill_y of the height."""
    xs, ys = block_cells(width, height, 0.0, fill_x * width, 0.0, fill_y * height)
    return seed_cells(xs, ys, ppc, dx, seed)


def falling_block(width, height, dx, ppc=8, seed=3, pool=0.5, block=512):
    """Mixed-precision large case: fluid in the bottom `pool` of the height plus a
    block x block square seeded near the top centre."""
    xs0, ys0 = block_cells(width, height, 0.0, width, 0.0, pool * height)
    bx0 = (width - block) / 2.0
    by1 = height - 1 - 16
    xs1, ys1 = block_cells(width, height, bx0, bx0 + block, by1 - block, by1)
    xs = np.concatenate([xs0, xs1])
    ys = np.concatenate([ys0, ys1])
    order = np.lexsort((xs, ys))
    return seed_cells(xs[order], ys[order], ppc, dx, seed)


def dam_break_fill(width, height, dx, fill, ppc, seed):
    """Dam break whose block covers the fraction `fill` of the domain with the
    2:3 aspect of the 0.4 x 0.6 block (fill = 0.24 gives that block)."""
    fx = float(np.sqrt(fill * 2.0 / 3.0))
    fy = float(np.sqrt(fill * 3.0 / 2.0))
    return dam_break(width, height, dx, fx, fy, ppc, seed)


def instance_mix(count, side_min, side_max, seed, progress=None):
    """Independent dam breaks of mixed sizes: square sides log-uniform in
    [side_min, side_max] (both even), rounded to the nearest even side; fill
    fraction U[0.2, 0.6], 4-8 particles per cell, gravity angle U[-15, 15]
    degrees.  Half the even sides are 2 x odd, whose multigrid hierarchy
    stops at (side / 2)^2 cells.  Returns a list of dicts (width, height, dx,
    ppc, fill, angle_deg, gravity, gravity_x, seed, xy).  progress(i), when
    given, is called after every 256th instance."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        side = float(np.exp(rng.uniform(np.log(side_min), np.log(side_max))))
        side = int(min(max(2 * round(side / 2.0), side_min), side_max))
        fill = float(rng.uniform(0.2, 0.6))
        ppc = int(rng.integers(4, 9))
        angle = float(rng.uniform(-15.0, 15.0))
        g = 9.81
        inst_seed = int(seed) * 100003 + i
        dx = 1.0 / side
        out.append(dict(width=side, height=side, dx=dx, ppc=ppc, fill=fill, angle_deg=angle,
                        gravity=float(-g * np.cos(np.radians(angle))), gravity_x=float(g * np.sin(np.radians(angle))),
                        seed=inst_seed, xy=dam_break_fill(side, side, dx, fill, ppc, inst_seed)))
        if progress is not None and (i + 1) % 256 == 0:
            progress(i + 1)
    return out


CONFIGS = {
    # dam breaks: the block fills the left 40% x bottom 60%
    "cfg1_dam_break_128": dict(width=128, height=128, dx=1.0 / 128, dt=0.002, steps=100, ppc=4, seed=1),
    "cfg2_dam_break_1024": dict(width=1024, height=1024, dx=1.0 / 1024, dt=2.5e-4, steps=200, ppc=8, seed=2),
    # the mixed-precision large case: the bottom half filled plus a 512 x 512
    # block near the top (35,602,448 particles); the workload names no dt, so
    # it is cfg2's scaled with the cell size (the same CFL number)
    "cfg3_falling_block_4096x2048": dict(width=4096, height=2048, dx=1.0 / 4096, dt=6.25e-5, steps=50, ppc=8,
                                         seed=3, scene="falling_block", block=512),
    # the same construction at 1/8 scale, small enough for the oracle in tests
    "falling_block_512x256": dict(width=512, height=256, dx=1.0 / 512, dt=5e-4, steps=50, ppc=8, seed=3,
                                  scene="falling_block", block=64),
}


def config_scene(name):
    c = CONFIGS[name]
    if c.get("scene") == "falling_block":
        return falling_block(c["width"], c["height"], c["dx"], ppc=c["ppc"], seed=c["seed"], block=c["block"])
    return dam_break(c["width"], c["height"], c["dx"], 0.4, 0.6, c["ppc"], c["seed"])
