"""The pipelined sweeps' timeout fallback (xf_capi.hip st_solve_from), forced
from the host: xf_sim_debug_fail_sweeps sets the error word that a timed-out
wait sets, every pipelined wait of the next solve then gives up at once (all
kernels exit normally), and the solve is redone with per-hyperplane sweeps.
Both schedules evaluate the same MIC(0) pivots, sweeps and dot(z, r) partials
in the same order, so the substep's fields equal those of a run without the
failure, bit for bit; the solve after that runs pipelined again."""
import sys

import numpy as np
import pytest

from scenes import gpu_like, oracle_dam_break, small_params

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_forced_sweep_timeout_falls_back_to_the_same_substep(prec):
    from xfluid_amd import FIELD
    say(f"[fallback] 32x16x16 dam break, precision {prec}: forced sweep timeout")
    o = oracle_dam_break(32, 16, 16, 0.1, (0.5, 0.6, 1.0), 5, small_params(precision=prec))
    a, b = gpu_like(o), gpu_like(o)
    dt = 1.0 / 240
    a.timeStep(dt)
    b.debug_fail_sweeps()
    b.timeStep(dt)
    sa, sb = a.last_solve(), b.last_solve()
    assert sa.fallback == 0 and sb.fallback == 1
    assert sb.status == 0 and sb.iterations == sa.iterations
    for f in ("PRESSURE", "U", "V", "W"):
        assert np.array_equal(a.get_field(FIELD[f]), b.get_field(FIELD[f])), f
    a.timeStep(dt)
    b.timeStep(dt)
    assert a.last_solve().fallback == 0 and b.last_solve().fallback == 0
    for f in ("PRESSURE", "U", "V", "W"):
        assert np.array_equal(a.get_field(FIELD[f]), b.get_field(FIELD[f])), f
    Pa, Va = a.particles()
    Pb, Vb = b.particles()
    assert np.array_equal(Pa, Pb) and np.array_equal(Va, Vb)
