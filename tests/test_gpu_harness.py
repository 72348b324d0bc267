"""The reference's own benchmark workload (benches/fluid_benchmark.rs:10-39):
Apic::new(size, size, 1.0, props) -- the constructor's own particle seeding,
cell size 1.0 -- under the reference's solver rules, 15 warm-up updates of
dt = 0.016, then the timed updates.  At 512^2 (1,136,277 particles) and
1024^2 (4,631,445 particles) the GPU runs the warm-up and one more update
beside the oracle, and the particle positions and velocities, the MAC
velocities, the density and the pressure must be within 1e-5 relative L2 of
it, the time residuals within 2e-4 (dt's rounding and the jitter, as in the
mixed workload's bars); the affine matrices C within 1e-5 of the larger of
their own norm and the velocities' norm over one cell (the scene falls
uniformly, so C_V_Y is rounding noise: the FP32 oracle's differs from the
FP64 oracle's by its whole norm); the particle count, the substeps and the
PCG iteration counts must be equal.
"""
import numpy as np
import pytest

from helpers import PARTICLE_FIELDS, props, rel_l2

pytestmark = pytest.mark.gpu

WARMUP = 15
DT = 0.016


@pytest.mark.parametrize("size,particles", [(512, 1136277), (1024, 4631445)])
def test_harness_workload_matches_oracle(size, particles):
    from FLIP.apic import Apic, DebugOptions
    from oracle import pyoracle as po
    g = Apic.new(size, size, 1.0, props())
    o = po.Oracle(size, size, 1.0, po.bench_props())
    assert g.dims()["num_particles"] == o.dims()["num_particles"] == particles
    print(f"harness {size}^2: {particles} particles, {WARMUP} warm-up updates + 1", flush=True)
    for k in range(WARMUP + 1):
        g.update(DebugOptions(False), DT)
        o.update(DT)
        sg, so = g.stats(), o.stats()
        assert sg["substeps"] == so["substeps"], (k, sg, so)
        assert sg["pcg_iterations_total"] == so["pcg_iterations_total"], (k, sg, so)
    print(f"  last update: {sg['substeps']} substeps, {sg['pcg_iterations_total']} PCG iterations, "
          f"exit {sg['pcg_exit_reason']}", flush=True)
    vnorm = float(np.hypot(np.linalg.norm(o.read("VEL_X")), np.linalg.norm(o.read("VEL_Y"))))  # / cell 1.0
    for f in PARTICLE_FIELDS + ["MAC_U", "MAC_V", "DENSITY", "PRESSURE"]:
        a, b = g.read(f).astype(np.float64), o.read(f).astype(np.float64)
        if f.startswith("C_"):
            err = np.linalg.norm(a - b) / max(np.linalg.norm(b), vnorm)
        else:
            err = rel_l2(a, b)
        print(f"  {f}: {err:.2e}")
        assert err <= (2e-4 if f == "TIME_RESIDUAL" else 1e-5), (f, err)
    g.close()
