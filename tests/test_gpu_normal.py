"""GPU parity of the NormalPressureSolver operator (pressuresolver.cpp:50-347)
against the oracle: stencil codes give the same right-hand side and MIC(0)
pivots bit for bit, the PCG pressure within 1e-5 of the FP64 oracle's, and
applyPressure bit for bit from the same pressure; full substeps (FP64 and
mixed precision) stay within the dam break's trajectory tolerances.
"""
import sys

import numpy as np
import pytest

from oracle_lib import FIELD, STAGE
from scenes import copy_state, gpu_like, oracle_dam_break, small_params

pytestmark = pytest.mark.gpu

CASES = {
    "c1_small": (24, 16, 12, 0.125, (0.35, 0.6, 1.0), 0),
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


def _exact(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    if not np.array_equal(a, b):
        d = np.abs(a.astype(np.float64) - b.astype(np.float64))
        raise AssertionError(f"{what}: {np.count_nonzero(d)} entries differ, max {d.max():.3e}")


def _close(a, b, rtol, what):
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max() / scale
    assert err <= rtol, f"{what}: relative error {err:.3e} > {rtol:.1e}"


@pytest.mark.parametrize("case", list(CASES))
def test_normal_stages(xf, case):
    ni, nj, nk, dx, frac, seed = CASES[case]
    say(f"[normal] case={case} grid={ni}x{nj}x{nk}")
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, small_params(pressure_solver=1))
    g = gpu_like(o)
    dt = 1.0 / 60
    for step in range(3):
        copy_state(o, g)
        for s in ("MATERIAL", "SAVE"):
            o.stage(STAGE[s]); g.stage(STAGE[s])
        o.stage(STAGE["FORCES"], dt); g.stage(STAGE["FORCES"], dt)
        so = o.stage(STAGE["SOLVE"]); sg = g.stage(STAGE["SOLVE"])
        _exact(g.get_field(FIELD["RHS"]), o.get(FIELD["RHS"]), f"{case} step{step} rhs")
        _exact(g.get_field(FIELD["MIC_INV"]), o.get(FIELD["MIC_INV"]), f"{case} step{step} mic0")
        assert so == 0 and sg == 0
        st_o, st_g = o.last_solve(), g.last_solve()
        assert abs(st_o.iterations - st_g.iterations) <= 2, (st_o.iterations, st_g.iterations)
        assert st_g.residual <= 1e-6 * st_g.rhs_norm
        _close(g.get_field(FIELD["PRESSURE"]), o.get(FIELD["PRESSURE"]), 1e-5, f"{case} step{step} pressure")
        g.set_field(FIELD["PRESSURE"], o.get(FIELD["PRESSURE"]))
        o.stage(STAGE["APPLY"]); g.stage(STAGE["APPLY"])
        for f in ("U", "V", "W"):
            _exact(g.get_field(FIELD[f]), o.get(FIELD[f]), f"{case} step{step} apply {f}")
        for s in ("EXTRAPOLATE",):
            o.stage(STAGE[s]); g.stage(STAGE[s])
        o.stage(STAGE["ADVECT"], dt); g.stage(STAGE["ADVECT"], dt)
        o.stage(STAGE["P2G"]); g.stage(STAGE["P2G"])


def test_normal_trajectory(xf):
    ni, nj, nk, dx, frac, seed = CASES["c1_small"]
    say("[normal] trajectory 5 substeps")
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, small_params(pressure_solver=1))
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


def test_normal_mixed_solve(xf):
    """FP32-bulk PCG + FP64 refinement on the Normal operator reaches the FP64 tolerance."""
    ni, nj, nk, dx, frac, seed = CASES["tiles"]
    say("[normal] mixed solve, tiles")
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, small_params(pressure_solver=1))
    for _ in range(2):
        o.time_step(1.0 / 60)
    o.stage(STAGE["MATERIAL"]); o.stage(STAGE["SAVE"]); o.stage(STAGE["FORCES"], 1.0 / 60)
    g = gpu_like(o)
    copy_state(o, g)
    assert o.stage(STAGE["SOLVE"]) == 0
    p = g.params
    p.precision = xf.PREC_MIXED
    g.set_params(p)
    assert g.stage(STAGE["SOLVE"]) == 0
    st = g.last_solve()
    assert st.residual <= 1e-6 * st.rhs_norm and st.outer >= 1
    _close(g.get_field(FIELD["PRESSURE"]), o.get(FIELD["PRESSURE"]), 1e-4, "mixed pressure vs FP64 oracle")
