"""GPU parity at the benchmark workloads' own sizes.

* The library's init path (setSolidBoundary, addCubicFluid, addSpereFluid,
  initParticles; FluidSimulator.cpp:64-140, 168-222, 347-370) against the
  oracle's: material and every particle position bit for bit, at 64x32x32
  (90,720 particles) and 128^3 (3,175,200 particles).
* The 64x32x32 dam break (dx 0.125, CFL 5, FLIP 0.95, FP64 PCG to 1e-6, blob
  radius 1.0 dx, iso 0.5): 30 x frameStep(1/30) (FluidSimulator.cpp:224-262)
  on the GPU and the oracle side by side: positions within 1e-10 relative
  after every frame (observed: 1.6e-15 after frame 30), particle count and
  final material exact; the surface cells, cube cases and the mesh of the
  GPU's final density field equal the oracle's on that same field.
* The 128^3 dam break (3,175,200 particles, dt 1/240): one substep stage by
  stage on identical inputs (material, forces, RHS, MIC(0) pivots, apply,
  extrapolation, advection bit-exact, one 8x8-tile preconditioner application
  bit-exact against the serial sweeps, P2G and density 1e-12, pressure 1e-5),
  then three whole substeps in FP64 and mixed precision with the error of
  pressure, u, v, w and positions against the FP64 oracle.
* The 256^3 sphere-scene projection (VariationalPressureSolver::solve +
  applyPressure, variational_pressuresolver.cpp:50-220), FP64 and mixed, to
  1e-6 and 1e-9: the true residual |A p - b| (the oracle's assembly and
  SpMV) within the tolerance, and the divergence of the projected velocity on
  liquid cells with no solid neighbour, which equals minus the residual row
  (div u' = div u + (A p)_c = (A p - b)_c), equally small.
"""
import sys

import numpy as np
import pytest

from oracle_lib import FIELD, STAGE, OracleSim, default_params, oracle
from scenes import copy_state, gpu_like, oracle_dam_break

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


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def _c1_params(lib):
    return lib.default_params(cfl=5.0, pcg_tol=1e-6, pcg_max_iter=200, blob_radius=0.125, iso=0.5, flip_ratio=0.95)


def test_c1_dam64x32x32_frames(xf, tmp_path):
    ni, nj, nk, dx, frac, seed, cells = DAMS["dam64x32x32"]
    import oracle_lib
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, _c1_params(oracle_lib))
    rd = np.array([frac[0] * ni * dx, frac[1] * nj * dx, frac[2] * nk * dx])
    g = xf.FluidSimulator(ni, nj, nk, _c1_params(xf))
    g.setCellWidth(dx)
    g.setSolidBoundary()
    g.addCubicFluid((0.0, 0.0, 0.0), rd)
    g.initParticles(xf.SEED_SUBCELL, seed)
    g.addConstantForce(0.0, -9.81, 0.0)
    g.set_mesh_output(str(tmp_path) + "/frame_")
    _exact(g.particles()[0], o.particles()[0], "C1 initial particles")
    n0 = g.num_particles()
    assert n0 == 8 * cells
    for f in range(30):
        assert o.frame_step(1.0 / 30) == 0
        assert g.frameStep(1.0 / 30) == 0
        Po, Pg = o.particles()[0], g.particles()[0]
        err = _rel(Pg, Po)
        if f in (0, 2, 9, 19, 29):
            say(f"[C1] frame {f + 1}: positions rel err {err:.3e}, PCG iterations {g.last_solve().iterations}")
        assert err <= 1e-10, f"positions after {f + 1} frames: {err:.3e}"
    assert g.num_particles() == n0 == len(Po)
    _exact(g.get_field(FIELD["MATERIAL"]), o.get(FIELD["MATERIAL"]), "C1 material after 30 frames")
    assert (tmp_path / "frame_29.ply").exists()
    # classification and mesh on the same (the GPU's) density field
    phi = g.get_field(FIELD["PHI"])
    o.set(FIELD["PHI"], phi)
    _exact(g.surface_cells()[0], o.surface_cells()[0], "C1 surface cells")
    _exact(g.surface_cells()[1], o.surface_cells()[1], "C1 cube cases")
    vg, tg, ng = g.generate_mesh(True)
    vo, to, no = o.mesh(True)
    assert len(tg) > 1000
    _exact(tg, to, "C1 triangles")
    _exact(vg, vo, "C1 vertices")


def _dam128(xf, precision, prm_mod):
    ni, nj, nk, dx, frac, seed, cells = DAMS["dam128"]
    p = prm_mod.default_params(flip_ratio=0.95, pcg_tol=1e-6, pcg_max_iter=200, precision=precision) \
        if prm_mod is xf else prm_mod.default_params(flip_ratio=0.95, pcg_tol=1e-6, pcg_max_iter=200)
    return ni, nj, nk, dx, frac, seed, p


def test_c2_dam128_stages(xf):
    ni, nj, nk, dx, frac, seed, p = _dam128(xf, 0, __import__("oracle_lib"))
    say("[C2] 128^3 stage by stage")
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, p)
    g = gpu_like(o)
    dt = 1.0 / 240
    assert o.time_step(dt) == 0  # one substep first: non-zero velocities everywhere
    copy_state(o, g)
    for st in ("MATERIAL", "SAVE"):
        o.stage(STAGE[st]); g.stage(STAGE[st])
    _exact(g.get_field(FIELD["MATERIAL"]), o.get(FIELD["MATERIAL"]), "C2 material")
    o.stage(STAGE["FORCES"], dt); g.stage(STAGE["FORCES"], dt)
    for f in ("U", "V", "W"):
        _exact(g.get_field(FIELD[f]), o.get(FIELD[f]), f"C2 forces {f}")
    assert o.stage(STAGE["SOLVE"]) == 0 and g.stage(STAGE["SOLVE"]) == 0
    _exact(g.get_field(FIELD["RHS"]), o.get(FIELD["RHS"]), "C2 rhs")
    inv = o.get(FIELD["MIC_INV"])
    _exact(g.get_field(FIELD["MIC_INV"]), inv, "C2 MIC(0) pivots")
    st = g.last_solve()
    assert st.residual <= 1e-6 * st.rhs_norm and st.fallback == 0
    say(f"[C2] PCG {st.iterations} iterations (oracle {o.last_solve().iterations}), box {st.box_cells} cells")
    assert _rel(g.get_field(FIELD["PRESSURE"]), o.get(FIELD["PRESSURE"])) <= 1e-5
    # one preconditioner application (the 8x8-tile pipeline, fused forward + backward) vs the serial sweeps
    N = ni * nj * nk
    O = oracle()
    A7, b = np.empty(7 * N), np.empty(N)
    O.xo_assemble(ni, nj, nk, o.get(FIELD["MATERIAL"]).ravel(), o.get(FIELD["U"]).ravel(), o.get(FIELD["V"]).ravel(),
                  o.get(FIELD["W"]).ravel(), A7, b)
    x = np.random.default_rng(5).standard_normal(N)
    x[inv.ravel() == 0] = 0.0
    q, z = np.empty(N), np.empty(N)
    O.xo_sweep_fwd(ni, nj, nk, A7, inv.ravel(), x, q)
    O.xo_sweep_bwd(ni, nj, nk, A7, inv.ravel(), q, z)
    _exact(g.apply_precond(x), z, "C2 preconditioner application")
    g.set_field(FIELD["PRESSURE"], o.get(FIELD["PRESSURE"]))
    for st in ("APPLY", "EXTRAPOLATE"):
        o.stage(STAGE[st]); g.stage(STAGE[st])
        for f in ("U", "V", "W"):
            _exact(g.get_field(FIELD[f]), o.get(FIELD[f]), f"C2 {st.lower()} {f}")
    o.stage(STAGE["ADVECT"], dt); g.stage(STAGE["ADVECT"], dt)
    Po, Vo = o.particles()
    Pg, Vg = g.particles()
    _exact(Pg, Po, "C2 advected positions")
    _exact(Vg, Vo, "C2 G2P velocities")
    o.stage(STAGE["P2G"]); g.stage(STAGE["P2G"])
    for f in ("U", "V", "W"):
        assert _rel(g.get_field(FIELD[f]), o.get(FIELD[f])) <= 1e-12, f"C2 p2g {f}"
    o.stage(STAGE["DENSITY"]); g.stage(STAGE["DENSITY"])
    assert _rel(g.get_field(FIELD["PHI"]), o.get(FIELD["PHI"])) <= 1e-12, "C2 density"


def test_c2_dam128_substeps_fp64_and_mixed(xf):
    import oracle_lib
    ni, nj, nk, dx, frac, seed, po = _dam128(xf, 0, oracle_lib)
    say("[C2] 128^3: 3 substeps, oracle vs GPU FP64 and mixed")
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, po)
    dt = 1.0 / 240
    for _ in range(3):
        assert o.time_step(dt) == 0
    ref = {f: o.get(FIELD[f]) for f in ("PRESSURE", "U", "V", "W")}
    ref["positions"] = o.particles()[0]
    bounds = {0: dict(pressure=1e-10, u=1e-10, v=1e-10, w=1e-10, positions=1e-12),
              1: dict(pressure=1e-5, u=1e-5, v=1e-5, w=1e-5, positions=1e-8)}
    for prec in (0, 1):
        _, _, _, _, _, _, p = _dam128(xf, prec, xf)
        g = xf.FluidSimulator(ni, nj, nk, p)
        g.setCellWidth(dx)
        g.setSolidBoundary()
        g.addCubicFluid((0.0, 0.0, 0.0), (frac[0] * ni * dx, frac[1] * nj * dx, frac[2] * nk * dx))
        g.initParticles(xf.SEED_SUBCELL, seed)
        g.addConstantForce(0.0, -9.81, 0.0)
        for _ in range(3):
            assert g.timeStep(dt) == 0
            st = g.last_solve()
            assert st.residual <= 1e-6 * st.rhs_norm and st.fallback == 0
        got = {f: g.get_field(FIELD[f]) for f in ("PRESSURE", "U", "V", "W")}
        got["positions"] = g.particles()[0]
        errs = {k.lower(): _rel(got[k], ref[k]) for k in ref}
        say(f"[C2] {'mixed' if prec else 'fp64'} per-field error vs FP64 oracle: " +
            ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
        for k, v in errs.items():
            assert v <= bounds[prec][k], (prec, k, v)
        g.close()


@pytest.fixture(scope="module")
def sphere256():
    from xfluid_amd.scenes import sphere_scene
    return sphere_scene(256)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("tol", [1e-6, 1e-9])
def test_c3_sphere256_projection(xf, sphere256, prec, tol):
    mat, u, v, w = sphere256
    n = 256
    say(f"[C3] 256^3 sphere scene, {'mixed' if prec else 'fp64'} to {tol:g}")
    p = xf.default_params(pcg_tol=tol, pcg_max_iter=2000, precision=prec)
    u2, v2, w2, pres, st = xf.project(mat, u, v, w, p)
    assert st.status == 0 and st.fallback == 0
    O = oracle()
    N = n ** 3
    A7, b = np.empty(7 * N), np.empty(N)
    O.xo_assemble(n, n, n, mat.ravel(), u.ravel(), v.ravel(), w.ravel(), A7, b)
    Ap = np.empty(N)
    O.xo_spmv(n, n, n, A7, pres.ravel(), Ap)
    bn = np.abs(b).max()
    true = np.abs(Ap - b).max() / bn
    say(f"[C3] {st.iterations} iterations ({st.outer} outer), {st.ms:.1f} ms, reported {st.residual / bn:.2e}, "
        f"true residual {true:.2e}, box {st.box_cells} cells, fused sweeps {st.sweep_fused}")
    assert true <= (1.0 if prec else 3.0) * tol + 1e-15
    # divergence of the projected field on liquid cells away from solid = -(A p - b)_c
    div = (u2[:, :, 1:] - u2[:, :, :-1]) + (v2[:, 1:, :] - v2[:, :-1, :]) + (w2[1:, :, :] - w2[:-1, :, :])
    sol = mat == 2
    near = np.zeros_like(sol)
    near[1:-1, 1:-1, 1:-1] = (sol[2:, 1:-1, 1:-1] | sol[:-2, 1:-1, 1:-1] | sol[1:-1, 2:, 1:-1] |
                              sol[1:-1, :-2, 1:-1] | sol[1:-1, 1:-1, 2:] | sol[1:-1, 1:-1, :-2])
    cells = (mat == 1) & ~near
    assert cells.sum() > 1000000
    dmax = np.abs(div[cells]).max() / bn
    say(f"[C3] max |div u'| on {int(cells.sum())} liquid cells: {dmax:.2e} of |b|")
    assert dmax <= 3.0 * tol + 1e-13
