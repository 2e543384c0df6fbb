"""GPU parity of the fluid sources (fluidsource.cpp:75-92 via
FluidSimulator.cpp:725-726) against the oracle.

The advection stage ends with the sources' update: new particles at the end of
the particle list, in source, cell and particle order, jittered from the same
31-bit SplitMix64 stream (the GPU jumps to each particle's draws, the oracle
walks the stream).  Positions, velocities, ids and material are bit-exact;
P2G over the grown particle set is within 1e-12 (sum order), and full
substeps stay within the trajectory tolerances of the dam break.
"""
import sys

import numpy as np
import pytest

from oracle_lib import FIELD, STAGE
from scenes import copy_state, gpu_like, oracle_dam_break, small_params

pytestmark = pytest.mark.gpu

DX = 0.125
SOURCES = [
    # a jet into the dam's open side, and a source overlapping the top wall
    ((18 * DX, 9 * DX, 6 * DX), (1.5 * DX, 2 * DX, 3 * DX), (-6.0, 0.0, 0.5)),
    ((10 * DX, 15.2 * DX, 4 * DX), (2 * DX, 1 * DX, 1 * DX), (0.0, -3.0, 0.0)),
]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def xf():
    import xfluid_amd
    if xfluid_amd.device_count() < 1:
        pytest.fail("no HIP device visible to libxfluid_hip.so")
    return xfluid_amd


def _exact(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, f"{what}: shapes {a.shape} vs {b.shape}"
    if not np.array_equal(a, b):
        d = np.abs(a.astype(np.float64) - b.astype(np.float64))
        raise AssertionError(f"{what}: {np.count_nonzero(d)} entries differ, max {d.max():.3e}")


def _close(a, b, rtol, what):
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max() / scale
    assert err <= rtol, f"{what}: relative error {err:.3e} > {rtol:.1e}"


def _scene(seed=5):
    o = oracle_dam_break(24, 16, 12, DX, (0.35, 0.6, 1.0), 0, small_params())
    g = gpu_like(o)
    for c, w, v in SOURCES:
        o.add_fluid_source(c, *w, v)
        g.addFluidSource(c, *w, v)
    o.set_source_seed(seed)
    g.set_source_seed(seed)
    return o, g


def test_source_stage_bitexact(xf):
    say("[source] stage parity 24x16x12, 2 sources, 3 substeps")
    o, g = _scene()
    n0 = len(o.particles()[0])
    for step, dt in enumerate((1.0 / 60, 1.0 / 90, 1.0 / 45)):
        copy_state(o, g)
        for s in ("MATERIAL", "SAVE"):
            o.stage(STAGE[s]); g.stage(STAGE[s])
        o.stage(STAGE["FORCES"], dt); g.stage(STAGE["FORCES"], dt)
        assert o.stage(STAGE["SOLVE"]) == 0 and g.stage(STAGE["SOLVE"]) == 0
        g.set_field(FIELD["PRESSURE"], o.get(FIELD["PRESSURE"]))
        for s in ("APPLY", "EXTRAPOLATE"):
            o.stage(STAGE[s]); g.stage(STAGE[s])
        assert o.stage(STAGE["ADVECT"], dt) == 0
        assert g.stage(STAGE["ADVECT"], dt) == 0
        Po, Vo = o.particles()
        Pg, Vg = g.particles()
        assert len(Po) > n0, "the sources appended nothing"
        _exact(Pg, Po, f"step{step} positions after advection + sources")
        _exact(Vg, Vo, f"step{step} velocities after advection + sources")
        _exact(g.get_field(FIELD["MATERIAL"]), o.get(FIELD["MATERIAL"]), f"step{step} material")
        n0 = len(Po)
        o.stage(STAGE["P2G"]); g.stage(STAGE["P2G"])
        for f in ("U", "V", "W"):
            _close(g.get_field(FIELD[f]), o.get(FIELD[f]), 1e-12, f"step{step} p2g {f}")


def test_source_trajectory(xf):
    say("[source] 6 full substeps")
    o, g = _scene(seed=11)
    dt = 1.0 / 60
    for _ in range(6):
        assert o.time_step(dt) == 0
        assert g.timeStep(dt) == 0
    Po, Vo = o.particles()
    Pg, Vg = g.particles()
    assert Pg.shape == Po.shape
    _close(Pg, Po, 1e-9, "positions after 6 substeps")
    _close(Vg, Vo, 1e-6, "velocities after 6 substeps")
    _exact(g.get_field(FIELD["MATERIAL"]), o.get(FIELD["MATERIAL"]), "material after 6 substeps")


def test_dense_cells_p2g(xf):
    """Cells holding hundreds of particles (a source's first frame, a compressed
    pool): the per-cell ordering and the P2G slabs too full for LDS (read from
    global memory instead) give the oracle's result (P2G to 1e-12, positions
    exact)."""
    say("[dense] 24x16x12, 60000 particles in 6^3 cells")
    o = oracle_dam_break(24, 16, 12, DX, (0.35, 0.6, 1.0), 0, small_params())
    rng = np.random.default_rng(3)
    n = 60000
    P = (rng.random((n, 3)) * 6.0 + 4.0) * DX
    V = rng.standard_normal((n, 3))
    o.set_particles(P, V)
    g = gpu_like(o)
    o.stage(STAGE["MATERIAL"]); g.stage(STAGE["MATERIAL"])
    _exact(g.get_field(FIELD["MATERIAL"]), o.get(FIELD["MATERIAL"]), "dense material")
    o.stage(STAGE["P2G"]); g.stage(STAGE["P2G"])
    for f in ("U", "V", "W"):
        _close(g.get_field(FIELD[f]), o.get(FIELD[f]), 1e-12, f"dense p2g {f}")
    Pg, Vg = g.particles()
    _exact(Pg, P, "dense positions (sorted and restored)")
    _exact(Vg, V, "dense velocities")


def test_source_demo_frames(xf):
    """The reference's demo_source scene at half resolution (32^3, dx 0.25, no
    initial fluid): three frameStep(1/30) calls, CFL-chosen substeps, particle
    count exact and the particles within the trajectory tolerances."""
    say("[source] demo scene 32^3, 3 frames")
    from oracle_lib import OracleSim, default_params
    dx = 0.25
    c, w, v = (2 * dx, 15 * dx, 15 * dx), (1 * dx, 2.5 * dx, 5 * dx), (10.0, 0.0, 0.0)
    o = OracleSim(32, 32, 32, dx, default_params())
    o.lib.xo_add_constant_force(o.h, 0.0, -20.0, 0.0)
    o.add_fluid_source(c, *w, v)
    o.lib.xo_set_solid_boundary(o.h)
    o.set_source_seed(3)
    g = xf.FluidSimulator(32, 32, 32, xf.default_params())
    g.setCellWidth(dx)
    g.addConstantForce(0.0, -20.0, 0.0)
    g.addFluidSource(c, *w, v)
    g.set_source_seed(3)
    g.init()
    for f in range(3):
        assert o.frame_step(1.0 / 30) >= 0
        assert g.frameStep(1.0 / 30) >= 0
        Po, Vo = o.particles()
        Pg, Vg = g.particles()
        assert Pg.shape == Po.shape and len(Po) > 0, (f, Pg.shape, Po.shape)
        _close(Pg, Po, 1e-9, f"frame {f} positions")
        _close(Vg, Vo, 1e-6, f"frame {f} velocities")
