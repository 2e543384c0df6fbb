"""GPU parity at the benchmark workloads' own sizes.

* The library's init path (setSolidBoundary, addCubicFluid, addSpereFluid,
  initParticles; FluidSimulator.cpp:64-140, 168-222, 347-370) against the
  oracle's: material and every particle position bit for bit, at 64x32x32
  (90,720 particles) and 128^3 (3,175,200 particles).
"""
import sys

import numpy as np
import pytest

from oracle_lib import FIELD, OracleSim, default_params

pytestmark = pytest.mark.gpu

DAMS = {
    # name: (ni, nj, nk, dx, fluid fractions, seed, fluid cells)
    "dam64x32x32": (64, 32, 32, 0.125, (0.35, 0.6, 1.0), 0, 11340),
    "dam128": (128, 128, 128, 1.0 / 128, (0.4, 0.5, 1.0), 1, 396900),
}


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
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if not np.array_equal(a, b):
        d = np.abs(a.astype(np.float64) - b.astype(np.float64))
        raise AssertionError(f"{what}: {np.count_nonzero(d)} entries differ, max {d.max():.3e}")


@pytest.mark.parametrize("name", list(DAMS))
@pytest.mark.parametrize("mode", [0, 1])
def test_library_init_matches_oracle(xf, name, mode):
    ni, nj, nk, dx, frac, seed, cells = DAMS[name]
    say(f"[init] {name} seed mode {mode}")
    rd = np.array([frac[0] * ni * dx, frac[1] * nj * dx, frac[2] * nk * dx])
    g = xf.FluidSimulator(ni, nj, nk, xf.default_params())
    g.setCellWidth(dx)
    g.setSolidBoundary()
    g.addCubicFluid((0.0, 0.0, 0.0), rd)
    g.initParticles(mode, seed)
    o = OracleSim(ni, nj, nk, dx, default_params())
    o.lib.xo_set_solid_boundary(o.h)
    o.lib.xo_add_cubic_fluid(o.h, np.zeros(3), rd)
    o.lib.xo_seed_particles(o.h, mode, seed)
    m = g.get_field(FIELD["MATERIAL"])
    _exact(m, o.get(FIELD["MATERIAL"]), f"{name} material")
    assert int((m == 1).sum()) == cells
    assert g.num_particles() == 8 * cells
    Pg, Vg = g.particles()
    Po, Vo = o.particles()
    _exact(Pg, Po, f"{name} particle positions")
    assert not Vg.any()


def test_library_init_sphere_and_frame_defaults(xf):
    """addSpereFluid (float r^2, FluidSimulator.cpp:103-121) + init(), which sets the
    walls and seeds in the reference mode (seed 0) when no particles were given."""
    ni, nj, nk, dx = 40, 36, 28, 0.05
    say("[init] sphere + init()")
    g = xf.FluidSimulator(ni, nj, nk, xf.default_params())
    g.setCellWidth(dx)
    g.addSpereFluid(1.0, 0.9, 0.7, 0.45)
    g.addCubicFluid((0.0, 0.0, 0.0), (2.0, 0.3, 1.4))
    g.init()
    o = OracleSim(ni, nj, nk, dx, default_params())
    o.lib.xo_add_sphere_fluid(o.h, 1.0, 0.9, 0.7, 0.45)
    o.lib.xo_add_cubic_fluid(o.h, np.zeros(3), np.array([2.0, 0.3, 1.4]))
    o.lib.xo_set_solid_boundary(o.h)
    o.lib.xo_seed_particles(o.h, 0, 0)
    _exact(g.get_field(FIELD["MATERIAL"]), o.get(FIELD["MATERIAL"]), "sphere material")
    _exact(g.particles()[0], o.particles()[0], "sphere particles")
