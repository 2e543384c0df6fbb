"""The init path (FluidSimulator.cpp:64-140, 168-222, 347-385): walls, cubic
and spherical fluid regions, and the 8-particles-per-cell seeding.

Against the reference's own build (oracle/_ref): the material after
setSolidBoundary / addCubicFluid / addSpereFluid is bit-identical, and the
particles of initParticles after srand(seed) equal the oracle's seeding fed
with the same rand() values (cell order i slowest, the 8 points c -+ dx/4 in
_addParticle's order, randomDouble(-0.01, 0.01) jitter).  The reference draws a
particle's three jitter values as the arguments of one Vec3d constructor,
whose evaluation order C++ leaves open; the test finds the order this build
used.  Without the reference build, the fluid-cell and particle counts of the
two dam-break workloads are pinned by the cell-centre rule (11,340 cells /
90,720 particles at 64x32x32; 396,900 / 3,175,200 at 128^3).
"""
import numpy as np
import pytest

from oracle_lib import FIELD, OracleSim, RefSim, default_params, ref_available, ref_rand

SCENES = {
    # name: (ni, nj, nk, dx, cubic (lu, rd) or None, sphere (x, y, z, r) or None)
    "dam64x32x32": (64, 32, 32, 0.125, ((0.0, 0.0, 0.0), (0.35 * 8.0, 0.6 * 4.0, 4.0)), None),
    "sphere_and_block": (24, 20, 16, 0.1, ((0.0, 0.0, 0.0), (2.4, 0.45, 1.6)), (1.2, 1.1, 0.8, 0.55)),
    "odd_sphere": (17, 13, 11, 0.07, None, (0.6, 0.45, 0.38, 0.3)),
}


def _build(cls, name, walls_first=True):
    ni, nj, nk, dx, cube, sph = SCENES[name]
    s = cls(ni, nj, nk, dx, default_params())
    L = s.lib
    wall = L.xr_set_solid_boundary if cls is RefSim else L.xo_set_solid_boundary
    if walls_first:
        wall(s.h)
    if cube:
        (L.xr_add_cubic_fluid if cls is RefSim else L.xo_add_cubic_fluid)(s.h, np.array(cube[0]), np.array(cube[1]))
    if sph:
        (L.xr_add_sphere_fluid if cls is RefSim else L.xo_add_sphere_fluid)(s.h, *sph)
    if not walls_first:
        wall(s.h)
    return s


needs_ref = pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")


@needs_ref
@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("walls_first", [True, False])
def test_material_matches_reference(name, walls_first):
    o, r = _build(OracleSim, name, walls_first), _build(RefSim, name, walls_first)
    mo, mr = o.get(FIELD["MATERIAL"]), r.get(FIELD["MATERIAL"])
    assert (mo == 1).sum() > 100
    np.testing.assert_array_equal(mo, mr)


@needs_ref
@pytest.mark.parametrize("name", list(SCENES))
def test_particles_match_reference(name):
    seed = 77
    r = _build(RefSim, name)
    r.lib.xr_init_particles(r.h, seed)
    Pr, Vr = r.particles()
    n = len(Pr)
    assert n == 8 * int((r.get(FIELD["MATERIAL"]) == 1).sum())
    raw, rmax = ref_rand(seed, 3 * n)
    assert rmax == 2147483647
    raw = raw.reshape(n, 3)
    for order in ((0, 1, 2), (2, 1, 0)):
        o = _build(OracleSim, name)
        d = np.ascontiguousarray(raw[:, order]).ravel()
        o.lib.xo_seed_particles_draws(o.h, d.ctypes.data)
        Po, Vo = o.particles()
        if np.array_equal(Po, Pr):
            break
    else:
        raise AssertionError(f"{name}: seeded particles differ from the reference's "
                             f"(max {np.abs(Po - Pr).max() if Po.shape == Pr.shape else 'count'})")
    assert np.array_equal(Vo, Vr) and not Vo.any()


def _cube_count(ni, nj, nk, dx, lu, rd):
    c = [(np.arange(1, n - 1) + 0.5) * dx for n in (ni, nj, nk)]
    return int(np.prod([((x >= a) & (x <= b)).sum() for x, a, b in zip(c, lu, rd)]))


@pytest.mark.parametrize("dims,frac,cells", [((64, 32, 32, 0.125), (0.35, 0.6, 1.0), 11340),
                                             ((128, 128, 128, 1.0 / 128), (0.4, 0.5, 1.0), 396900)])
def test_dam_break_counts(dims, frac, cells):
    ni, nj, nk, dx = dims
    rd = (frac[0] * ni * dx, frac[1] * nj * dx, frac[2] * nk * dx)
    assert _cube_count(ni, nj, nk, dx, (0, 0, 0), rd) == cells
    o = OracleSim(ni, nj, nk, dx, default_params())
    o.lib.xo_set_solid_boundary(o.h)
    o.lib.xo_add_cubic_fluid(o.h, np.zeros(3), np.array(rd))
    assert int((o.get(FIELD["MATERIAL"]) == 1).sum()) == cells
    for mode in (0, 1):
        o.lib.xo_seed_particles(o.h, mode, 1)
        P, V = o.particles()
        assert len(P) == 8 * cells
        m = o.get(FIELD["MATERIAL"])
        if mode == 0:
            # cells i slowest, then the 8 points c -+ dx/4 in _addParticle's order,
            # each within the +-0.01 absolute jitter (more than dx/4 at 128^3)
            fl = np.argwhere(m.transpose(2, 1, 0) == 1)
            sg = np.array([(-1, -1, -1), (1, -1, -1), (1, -1, 1), (-1, -1, 1),
                           (-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1)])
            pts = ((fl[:, None, :] + 0.5) * dx + sg[None] * 0.25 * dx).reshape(-1, 3)
            assert np.abs(P - pts).max() <= 0.01 and np.abs(P - pts).max() > 0.009
            continue
        # mode 1: every particle lies in its own fluid cell, one per 2x2x2 sub-cell
        cell = np.floor(P / dx).astype(int)
        assert np.all(m[cell[:, 2], cell[:, 1], cell[:, 0]] == 1)
        sub = np.floor(P / (0.5 * dx)).astype(np.int64)
        key = sub[:, 0] + 2 * ni * (sub[:, 1] + 2 * nj * sub[:, 2])
        assert len(np.unique(key)) == len(P)
