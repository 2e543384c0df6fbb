"""Fluid sources: the oracle's CubicFluidSource against the reference's own.

The reference's CubicFluidSource (fluidsource.cpp:8-33, 75-92) draws its
jitter from rand().  The reference build in oracle/_ref runs one
_applyFluidSource after srand(seed); the oracle runs the same update with the
raw rand() values of that seed fed in as its draws.  Cell list and order,
particle count per cell ((int)(8 |v| dt / dx)), the randomDouble arithmetic
(xfluidutil.cpp:67-70), velocities and material must then agree bit for bit.
The reference draws a particle's three values as the arguments of one Vec3d
constructor, whose evaluation order C++ leaves open; the test finds the
order this build used and hands the oracle the draws in x, y, z order.

Without the reference build, the stream tests still pin the oracle's
SplitMix64 stream (the one the GPU reproduces by jumping ahead) and the
counts against hand-computed values.
"""
import numpy as np
import pytest

from oracle_lib import FIELD, OracleSim, RefSim, default_params, ref_available, ref_rand

GOLDEN = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1


def splitmix_draw(seed, d):
    """Draw d (0-based) of the 31-bit SplitMix64 stream: mix(seed + (d + 1) gamma) >> 33."""
    z = (seed + (d + 1) * GOLDEN) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return (z ^ (z >> 31)) >> 33


SOURCES = [
    # centre, half-widths, velocity: one inside the walls, one crossing the
    # domain's low corner (pos2Index truncates -0.5 -> 0 and -1.5 -> -1), one
    # overlapping the solid wall at high i
    ((4 * 0.125, 8 * 0.125, 8 * 0.125), (2 * 0.125, 3 * 0.125, 4 * 0.125), (10.0, 0.0, 0.0)),
    ((0.5 * 0.125, 1.0 * 0.125, 0.0), (1.0 * 0.125, 2.5 * 0.125, 0.5 * 0.125), (3.0, 4.0, 0.0)),
    ((15.2 * 0.125, 6.0 * 0.125, 5.0 * 0.125), (1.0 * 0.125, 1.0 * 0.125, 1.0 * 0.125), (0.0, -2.0, 1.0)),
]
DT = 0.0125


def _scene(cls, walls_first):
    s = cls(16, 14, 12, 0.125, default_params())
    if walls_first:
        s.lib.xr_set_solid_boundary(s.h) if cls is RefSim else s.lib.xo_set_solid_boundary(s.h)
    for c, w, v in SOURCES:
        s.add_fluid_source(c, *w, v)
    return s


def _counts(walls_first=True):
    """Particles each source appends in one update, from the stated rules."""
    o = _scene(OracleSim, walls_first)
    mat = o.get(FIELD["MATERIAL"])
    out = []
    for c, w, v in SOURCES:
        per = int(8.0 * np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * DT / 0.125)
        lo = [int((c[a] - w[a]) / 0.125) for a in range(3)]
        hi = [int((c[a] + w[a]) / 0.125) for a in range(3)]
        cells = 0
        for i in range(lo[0], hi[0] + 1):
            for j in range(lo[1], hi[1] + 1):
                for k in range(lo[2], hi[2] + 1):
                    if 0 <= i < 16 and 0 <= j < 14 and 0 <= k < 12 and mat[k, j, i] != 2:
                        cells += 1
        out.append(per * cells)
    return out


def test_source_stream_matches_jump_ahead():
    """The oracle's sequential stream equals draw d = mix(seed + (d+1) gamma) >> 33 at every d."""
    o = _scene(OracleSim, True)
    seed = 12345
    o.set_source_seed(seed)
    assert o.apply_sources(DT) == 0
    P, V = o.particles()
    n = len(P)
    assert n == sum(_counts()) and n > 0
    # rebuild the positions from the explicit draws: same values
    o2 = _scene(OracleSim, True)
    draws = np.array([splitmix_draw(seed, d) for d in range(3 * n)], np.uint32)
    assert o2.apply_sources(DT, draws) == 0
    P2, V2 = o2.particles()
    assert np.array_equal(P, P2) and np.array_equal(V, V2)


def test_source_particles_inside_cells():
    """Every new particle lies within 0.4999 dx of its cell centre per axis, cells in i, j, k order."""
    o = _scene(OracleSim, True)
    o.set_source_seed(7)
    o.apply_sources(DT)
    P, V = o.particles()
    cell = np.floor(P / 0.125).astype(int)
    off = P - (cell + 0.5) * 0.125
    assert np.all(np.abs(off) <= 0.4999 * 0.125 + 1e-15)
    # the first source's cells come first, i slowest, then j, then k
    first = cell[: _counts()[0]]
    key = first[:, 0] * 10000 + first[:, 1] * 100 + first[:, 2]
    assert np.all(np.diff(key) >= 0)
    mat = o.get(FIELD["MATERIAL"])
    assert np.all(mat[cell[:, 2], cell[:, 1], cell[:, 0]] == 1)


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("walls_first", [True, False])
def test_source_matches_reference(walls_first):
    seed = 2024
    r = _scene(RefSim, walls_first)
    r.apply_sources(DT, seed)
    Pr, Vr = r.particles()
    n = len(Pr)
    assert n > 0
    raw, rmax = ref_rand(seed, 3 * n)
    assert rmax == 2147483647, "the stream restatement assumes glibc's RAND_MAX"
    raw = raw.reshape(n, 3)
    # which order did this build evaluate the Vec3d(randomDouble x, y, z) arguments in?
    for order in ((0, 1, 2), (2, 1, 0)):
        d = np.ascontiguousarray(raw[:, order]).ravel()
        o_try = _scene(OracleSim, walls_first)
        assert o_try.apply_sources(DT, d) == 0
        Po, Vo = o_try.particles()
        if np.array_equal(Po, Pr):
            break
    else:
        diff = np.abs(Po - Pr).max() if Po.shape == Pr.shape else None
        raise AssertionError(f"oracle source particles differ from the reference's (max {diff})")
    assert np.array_equal(Vo, Vr)
    np.testing.assert_array_equal(o_try.get(FIELD["MATERIAL"]), r.get(FIELD["MATERIAL"]))


def test_source_golden():
    """The committed fixture of the reference's update (tests/golden/make_golden.py
    `sources`): the oracle fed the same raw draws reproduces it bit for bit."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sources16x14x12.npz")
    g = np.load(path, allow_pickle=False)
    dims = tuple(int(x) for x in g["dims"])
    o = OracleSim(*dims, float(g["dx"]), default_params())
    o.lib.xo_set_solid_boundary(o.h)
    for c, w, v in zip(g["centers"], g["halves"], g["vels"]):
        o.add_fluid_source(c, *w, v)
    draws = np.ascontiguousarray(g["raw"].reshape(-1, 3)[:, tuple(int(x) for x in g["order"])]).ravel()
    assert o.apply_sources(float(g["dt"]), draws) == 0
    P, V = o.particles()
    assert np.array_equal(P, g["P"]) and np.array_equal(V, g["V"])
    assert np.array_equal(o.get(FIELD["MATERIAL"]), g["material"])
