"""GPU parity: every stage of the substep on the HIP library against the FP64
oracle on identical inputs.

Bit-exact where the arithmetic order is fixed (material, forces, right-hand
side, MIC(0) pivots, SpMV, preconditioner sweeps, apply-pressure,
extrapolation, G2P + advection, surface classification).  Tolerances where the
GPU sums in another order: P2G and density (relative 1e-12), the PCG solution
(relative residual of the solve, and pressure within 1e-5 of the FP64 oracle's).
"""
import sys

import numpy as np
import pytest

from oracle_lib import FIELD, STAGE, oracle
from scenes import copy_state, gpu_like, oracle_dam_break, small_params

pytestmark = pytest.mark.gpu

CASES = {
    # (ni, nj, nk, dx, fluid fractions, seed)
    "c1_small": (24, 16, 12, 0.125, (0.35, 0.6, 1.0), 0),
    "c1_quarter": (32, 16, 16, 0.125, (0.35, 0.6, 1.0), 0),
    "odd": (19, 13, 11, 0.1, (0.5, 0.55, 1.0), 7),
    # several 8x8 (j,k) line tiles in both directions: exercises the tile-to-tile hand-off of the pipelined sweeps
    "tiles": (21, 40, 35, 0.05, (0.6, 0.7, 1.0), 9),
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


def _advance(o, steps, dt):
    for _ in range(steps):
        o.time_step(dt)


def _exact(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    if not np.array_equal(a, b):
        d = np.abs(a.astype(np.float64) - b.astype(np.float64))
        raise AssertionError(f"{what}: {np.count_nonzero(d)} entries differ, max {d.max():.3e}")


def _close(a, b, rtol, what):
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max() / scale
    assert err <= rtol, f"{what}: relative error {err:.3e} > {rtol:.1e}"


@pytest.mark.parametrize("case", list(CASES))
def test_stages_bitexact(xf, case):
    ni, nj, nk, dx, frac, seed = CASES[case]
    say(f"[parity] case={case} grid={ni}x{nj}x{nk}")
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, small_params())
    g = gpu_like(o)
    dt = 1.0 / 60
    for step in range(3):
        # material
        copy_state(o, g)
        o.stage(STAGE["MATERIAL"]); g.stage(STAGE["MATERIAL"])
        _exact(g.get_field(FIELD["MATERIAL"]), o.get(FIELD["MATERIAL"]), f"{case} step{step} material")
        o.stage(STAGE["SAVE"]); g.stage(STAGE["SAVE"])
        for f in ("USAVED", "VSAVED", "WSAVED"):
            _exact(g.get_field(FIELD[f]), o.get(FIELD[f]), f"{case} step{step} {f}")
        o.stage(STAGE["FORCES"], dt); g.stage(STAGE["FORCES"], dt)
        for f in ("U", "V", "W"):
            _exact(g.get_field(FIELD[f]), o.get(FIELD[f]), f"{case} step{step} forces {f}")
        # pressure solve: RHS and pivots exact, solution to tolerance
        so = o.stage(STAGE["SOLVE"]); sg = g.stage(STAGE["SOLVE"])
        _exact(g.get_field(FIELD["RHS"]), o.get(FIELD["RHS"]), f"{case} step{step} rhs")
        _exact(g.get_field(FIELD["MIC_INV"]), o.get(FIELD["MIC_INV"]), f"{case} step{step} mic0")
        assert so == 0 and sg == 0
        st_o, st_g = o.last_solve(), g.last_solve()
        assert abs(st_o.iterations - st_g.iterations) <= 2, (st_o.iterations, st_g.iterations)
        assert st_g.residual <= 1e-6 * st_g.rhs_norm
        _close(g.get_field(FIELD["PRESSURE"]), o.get(FIELD["PRESSURE"]), 1e-5, f"{case} step{step} pressure")
        # apply with the oracle's pressure: exact
        g.set_field(FIELD["PRESSURE"], o.get(FIELD["PRESSURE"]))
        o.stage(STAGE["APPLY"]); g.stage(STAGE["APPLY"])
        for f in ("U", "V", "W"):
            _exact(g.get_field(FIELD[f]), o.get(FIELD[f]), f"{case} step{step} apply {f}")
        o.stage(STAGE["EXTRAPOLATE"]); g.stage(STAGE["EXTRAPOLATE"])
        for f in ("U", "V", "W"):
            _exact(g.get_field(FIELD[f]), o.get(FIELD[f]), f"{case} step{step} extrapolate {f}")
        o.stage(STAGE["ADVECT"], dt); g.stage(STAGE["ADVECT"], dt)
        Po, Vo = o.particles()
        Pg, Vg = g.particles()
        _exact(Pg, Po, f"{case} step{step} advect positions")
        _exact(Vg, Vo, f"{case} step{step} advect velocities")
        o.stage(STAGE["P2G"]); g.stage(STAGE["P2G"])
        for f in ("U", "V", "W"):
            _close(g.get_field(FIELD[f]), o.get(FIELD[f]), 1e-12, f"{case} step{step} p2g {f}")
        o.stage(STAGE["DENSITY"]); g.stage(STAGE["DENSITY"])
        _close(g.get_field(FIELD["PHI"]), o.get(FIELD["PHI"]), 1e-12, f"{case} step{step} density")


@pytest.mark.parametrize("case", ["c1_small", "odd", "tiles"])
def test_rk2_trilinear_stages(xf, case):
    """The switchable forms: RK2 (midpoint) advection bit-exact and the trilinear
    P2G splat within 1e-12 of the oracle's same forms, on identical inputs, then
    three whole substeps with both switched on to the trajectory tolerance."""
    ni, nj, nk, dx, frac, seed = CASES[case]
    say(f"[variants] case={case} integrator=RK2 p2g=trilinear")
    prm = small_params(integrator=1, p2g_kernel=1)
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, prm)
    g = gpu_like(o)
    assert g.params.integrator == 1 and g.params.p2g_kernel == 1
    dt = 1.0 / 60
    for step in range(3):
        copy_state(o, g)
        for s in ("MATERIAL", "SAVE"):
            o.stage(STAGE[s]); g.stage(STAGE[s])
        o.stage(STAGE["FORCES"], dt); g.stage(STAGE["FORCES"], dt)
        assert o.stage(STAGE["SOLVE"]) == 0
        for s in ("APPLY", "EXTRAPOLATE"):
            o.stage(STAGE[s])
        copy_state(o, g)  # the oracle's projected, extrapolated velocities
        o.stage(STAGE["ADVECT"], dt); g.stage(STAGE["ADVECT"], dt)
        Po, Vo = o.particles()
        Pg, Vg = g.particles()
        _exact(Pg, Po, f"{case} step{step} RK2 positions")
        _exact(Vg, Vo, f"{case} step{step} G2P velocities")
        o.stage(STAGE["P2G"]); g.stage(STAGE["P2G"])
        for f in ("U", "V", "W"):
            _close(g.get_field(FIELD[f]), o.get(FIELD[f]), 1e-12, f"{case} step{step} trilinear p2g {f}")
    o2 = oracle_dam_break(ni, nj, nk, dx, frac, seed, prm)
    g2 = gpu_like(o2)
    for _ in range(3):
        assert o2.time_step(dt) == 0
        assert g2.timeStep(dt) == 0
    _close(g2.particles()[0], o2.particles()[0], 1e-9, f"{case} positions after 3 RK2/trilinear substeps")


@pytest.mark.parametrize("case", ["c1_small", "odd"])
def test_operators_bitexact(xf, case):
    """SpMV and the MIC(0) preconditioner (both sweeps) on random vectors."""
    ni, nj, nk, dx, frac, seed = CASES[case]
    say(f"[operators] case={case}")
    O = oracle()
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, small_params())
    o.stage(STAGE["MATERIAL"])
    o.stage(STAGE["FORCES"], 1.0 / 60)
    g = gpu_like(o)
    copy_state(o, g)
    g.stage(STAGE["SOLVE"])
    N = ni * nj * nk
    mat = o.get(FIELD["MATERIAL"]).ravel()
    A7 = np.empty(7 * N)
    b = np.empty(N)
    O.xo_assemble(ni, nj, nk, mat, o.get(FIELD["U"]).ravel(), o.get(FIELD["V"]).ravel(),
                  o.get(FIELD["W"]).ravel(), A7, b)
    inv = np.empty(N)
    O.xo_mic0(ni, nj, nk, A7, 0.97, 0.25, inv)
    _exact(g.get_field(FIELD["MIC_INV"]).ravel(), inv, "mic0")
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(N)
    x[inv == 0] = 0.0
    y_o = np.empty(N)
    O.xo_spmv(ni, nj, nk, A7, x, y_o)
    _exact(g.apply_matrix(x), y_o, "spmv")
    q = np.empty(N)
    z = np.empty(N)
    O.xo_sweep_fwd(ni, nj, nk, A7, inv, x, q)
    O.xo_sweep_bwd(ni, nj, nk, A7, inv, q, z)
    _exact(g.apply_precond(x), z, "precond (8x8-tile pipeline)")
    p = g.params
    p.sweep_kernel = 1
    g.set_params(p)
    _exact(g.apply_precond(x), z, "precond (per-plane sweeps)")


def test_surface_classification_exact(xf):
    """Same density field -> same surface sub-cell list (reference order) and cube cases."""
    ni, nj, nk, dx, frac, seed = CASES["c1_quarter"]
    say("[surface] c1_quarter")
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, small_params(iso=0.5, blob_radius=1.0 * dx))
    o.stage(STAGE["DENSITY"])
    g = gpu_like(o)
    g.set_field(FIELD["PHI"], o.get(FIELD["PHI"]))
    co, qo = o.surface_cells()
    cg, qg = g.surface_cells()
    assert len(co) > 100
    _exact(cg, co, "surface cells")
    _exact(qg, qo, "cube index")


@pytest.mark.parametrize("smooth", [False, True])
def test_mesh_exact(xf, smooth):
    """Same density field -> the same mesh as the oracle: vertices in creation
    order, triangles, smoothed positions and normals, bit for bit."""
    ni, nj, nk, dx, frac, seed = CASES["c1_quarter"]
    say(f"[mesh] c1_quarter smooth={smooth}")
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, small_params())
    _advance(o, 2, 1.0 / 60)
    o.stage(STAGE["DENSITY"])
    g = gpu_like(o)
    g.set_field(FIELD["PHI"], o.get(FIELD["PHI"]))
    vo, to, no = o.mesh(smooth)
    vg, tg, ng = g.generate_mesh(smooth)
    assert len(to) > 1000
    _exact(tg, to, "triangles")
    _exact(vg, vo, "vertices")
    _exact(ng, no, "normals")


def test_mesh_golden_ply(xf, tmp_path):
    """The reference's own mesh and PLY bytes (golden fixture) from its density field."""
    import os
    g0 = np.load(os.path.join(os.path.dirname(__file__), "golden", "dam14x12x10_mesh.npz"))
    ni, nj, nk = (int(x) for x in g0["dims"])
    say("[mesh] golden 14x12x10")
    p = xf.default_params(**{k: getattr(small_params(), k) for k in ("cfl", "pcg_tol", "pcg_max_iter")})
    g = xf.FluidSimulator(ni, nj, nk, p)
    g.setCellWidth(float(g0["dx"]))
    g.set_field(FIELD["PHI"], g0["phi"])
    v, t, _ = g.generate_mesh(True)
    _exact(t, g0["tris"], "triangles")
    _exact(v, g0["verts"], "vertices")
    path = tmp_path / "m.ply"
    g.write_mesh_ply(str(path))
    assert path.read_bytes() == g0["ply"].tobytes()


def test_mesh_frame_output(xf, tmp_path):
    """frameStep writes <prefix><frame>.ply on its last substep (FluidSimulator.cpp:290-305)."""
    ni, nj, nk, dx, frac, seed = CASES["c1_small"]
    say("[mesh] frame output c1_small")
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, small_params())
    g = gpu_like(o)
    prefix = str(tmp_path) + "/frame_"
    g.set_mesh_output(prefix)
    assert g.frameStep(1.0 / 30) == 0
    assert g.frameStep(1.0 / 30) == 0
    for f in (0, 1):
        data = (tmp_path / f"frame_{f}.ply").read_bytes()
        head = data[:data.index(b"end_header\n") + 11].decode()
        nv = int(head.split("element vertex ")[1].split()[0])
        nt = int(head.split("element face ")[1].split()[0])
        assert nv > 100 and nt > 100
        assert len(data) == len(head) + 12 * nv + 13 * nt
    # the second file is the mesh of the density field of frame 1's last substep,
    # which is still the current one
    g.generate_mesh(True)
    g.write_mesh_ply(str(tmp_path / "again.ply"))
    assert (tmp_path / "again.ply").read_bytes() == (tmp_path / "frame_1.ply").read_bytes()


def test_trajectory_tolerance(xf):
    """Five full substeps run independently on GPU and oracle stay within tolerance."""
    ni, nj, nk, dx, frac, seed = CASES["c1_small"]
    say("[trajectory] c1_small 5 substeps")
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, small_params())
    g = gpu_like(o)
    dt = 1.0 / 60
    for _ in range(5):
        assert o.time_step(dt) == 0
        assert g.timeStep(dt) == 0
    Po, Vo = o.particles()
    Pg, Vg = g.particles()
    _close(Pg, Po, 1e-9, "positions after 5 substeps")
    _close(Vg, Vo, 1e-6, "velocities after 5 substeps")
    _exact(g.get_field(FIELD["MATERIAL"]), o.get(FIELD["MATERIAL"]), "material after 5 substeps")


def test_mixed_precision_solve(xf):
    """FP32-bulk PCG with FP64 refinement reaches the FP64 tolerance on the same system."""
    ni, nj, nk, dx, frac, seed = CASES["c1_quarter"]
    say("[mixed] c1_quarter")
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, small_params())
    for _ in range(2):
        o.time_step(1.0 / 60)
    o.stage(STAGE["MATERIAL"]); o.stage(STAGE["SAVE"]); o.stage(STAGE["FORCES"], 1.0 / 60)
    g = gpu_like(o)
    copy_state(o, g)
    o.stage(STAGE["SOLVE"])
    p = g.params
    p.precision = xf.PREC_MIXED
    g.set_params(p)
    assert g.stage(STAGE["SOLVE"]) == 0
    st = g.last_solve()
    assert st.residual <= 1e-6 * st.rhs_norm
    assert st.outer >= 1
    _close(g.get_field(FIELD["PRESSURE"]), o.get(FIELD["PRESSURE"]), 1e-4, "mixed pressure vs FP64 oracle")


def test_frame_step_runs(xf):
    ni, nj, nk, dx, frac, seed = CASES["c1_small"]
    say("[frame] c1_small frameStep(1/30)")
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, small_params())
    g = gpu_like(o)
    assert o.frame_step(1.0 / 30) == 0
    assert g.frameStep(1.0 / 30) == 0
    Po, _ = o.particles()
    Pg, _ = g.particles()
    _close(Pg, Po, 1e-8, "positions after one frame")


@pytest.mark.parametrize("case,kind", [("tiles", 9), ("odd", 9), ("c1_quarter", 9), ("c1_small", 9)])
def test_pipelined_sweeps_bitexact(xf, case, kind):
    """The pipelined line-tile sweeps give the same bits as the serial sweeps,
    and a whole solve with them equals the per-hyperplane solve exactly."""
    ni, nj, nk, dx, frac, seed = CASES[case]
    say(f"[pipelined] case={case} schedule={kind}")
    O = oracle()
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, small_params())
    o.time_step(1.0 / 60)
    o.stage(STAGE["MATERIAL"]); o.stage(STAGE["SAVE"]); o.stage(STAGE["FORCES"], 1.0 / 60)
    g = gpu_like(o)
    copy_state(o, g)
    p = g.params
    p.sweep_kernel = 1
    g.set_params(p)
    assert g.stage(STAGE["SOLVE"]) == 0
    p1 = g.get_field(FIELD["PRESSURE"])
    it1 = g.last_solve().iterations
    N = ni * nj * nk
    mat = o.get(FIELD["MATERIAL"]).ravel()
    A7, b = np.empty(7 * N), np.empty(N)
    O.xo_assemble(ni, nj, nk, mat, o.get(FIELD["U"]).ravel(), o.get(FIELD["V"]).ravel(),
                  o.get(FIELD["W"]).ravel(), A7, b)
    inv = g.get_field(FIELD["MIC_INV"]).ravel()
    x = np.random.default_rng(seed + 1).standard_normal(N)
    x[inv == 0] = 0.0
    q, z = np.empty(N), np.empty(N)
    O.xo_sweep_fwd(ni, nj, nk, A7, inv, x, q)
    O.xo_sweep_bwd(ni, nj, nk, A7, inv, q, z)
    p.sweep_kernel = kind
    g.set_params(p)
    _exact(g.apply_precond(x), z, "precond (pipelined sweeps)")
    copy_state(o, g)
    assert g.stage(STAGE["SOLVE"]) == 0
    st = g.last_solve()
    assert st.fallback == 0, "pipelined wait timed out"
    assert st.iterations == it1
    _exact(g.get_field(FIELD["PRESSURE"]), p1, "pressure, pipelined vs per-plane sweeps")


@pytest.mark.parametrize("case,kind", [("tiles", 9), ("odd", 9), ("c1_quarter", 9)])
def test_pipelined_sweeps_fp32_bitexact(xf, case, kind):
    """FP32 sweeps (mixed-precision solve): a pipelined schedule gives the same
    bits as the per-hyperplane FP32 sweeps, so the whole mixed solve (iteration
    counts, outer refinements, pressure) is identical."""
    ni, nj, nk, dx, frac, seed = CASES[case]
    say(f"[pipelined fp32] case={case} schedule={kind}")
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, small_params())
    o.time_step(1.0 / 60)
    o.stage(STAGE["MATERIAL"]); o.stage(STAGE["SAVE"]); o.stage(STAGE["FORCES"], 1.0 / 60)
    g = gpu_like(o)
    res = []
    for k in (1, kind):
        copy_state(o, g)
        p = g.params
        p.precision = xf.PREC_MIXED
        p.sweep_kernel = k
        g.set_params(p)
        assert g.stage(STAGE["SOLVE"]) == 0
        st = g.last_solve()
        assert st.fallback == 0, "pipelined wait timed out"
        res.append((st.iterations, st.outer, g.get_field(FIELD["PRESSURE"])))
    assert res[0][:2] == res[1][:2]
    _exact(res[1][2], res[0][2], "mixed pressure, pipelined vs per-plane FP32 sweeps")
