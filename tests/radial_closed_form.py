"""add_radial_force (src/apic.rs:2352-2410) as a closed form in FP64: every
u face (x, y) at (x dx, (y + 1/2) dx) and v face at ((x + 1/2) dx, y dx)
within the radius R of the centre c (and farther than sqrt(1e-5)) receives
the radial field strength (1 - d/R) (p - c)/d, its x component on u faces and
its y component on v faces; repelling forces (strength > 0) go to the
post-projection arrays, attracting ones to the external-force arrays."""
import numpy as np


def radial_force(W, H, dx, cx, cy, radius, strength):
    """(u field, v field) in FP64, on the (W+1) x H u faces and W x (H+1) v faces."""
    yu, xu = np.mgrid[0:H, 0:W + 1]
    yv, xv = np.mgrid[0:H + 1, 0:W]
    out = []
    for (gx, gy, comp) in ((xu * dx - cx, (yu + 0.5) * dx - cy, 0), ((xv + 0.5) * dx - cx, yv * dx - cy, 1)):
        d2 = gx * gx + gy * gy
        inside = (d2 < radius * radius) & (d2 > 1e-5)
        d = np.sqrt(np.where(inside, d2, 1.0))
        comp_v = gx if comp == 0 else gy
        out.append(np.where(inside, strength * (1.0 - d / radius) * comp_v / d, 0.0).reshape(-1))
    return out
