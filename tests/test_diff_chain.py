"""The FP64 CPU adjoint of whole substeps (tests/diff_chain.py on the oracle)
against central finite differences of the oracle's own forward: two substeps
of a small dam break with FLIP 0.95 (so the saved-field sampling matters), a
per-face force, walls the particles hit, and a loss that is linear in the final
positions, velocities and face velocities.  The PCG runs to 1e-14 so the
solve's stopping error stays far below the differences.
"""
import numpy as np
import pytest

from diff_chain import rollout_grad
from oracle_lib import FIELD
from scenes import oracle_dam_break, small_params

DT = 1.0 / 30


def _scene(flip, integrator=0):
    p = small_params(flip_ratio=flip, pcg_tol=1e-14, pcg_max_iter=2000, tail_skip=0, integrator=integrator)
    o = oracle_dam_break(14, 12, 10, 0.1, (0.5, 0.6, 1.0), 5, p)
    for _ in range(2):
        o.time_step(DT)
    X, V = o.particles()
    U = [o.get(FIELD[n]) for n in ("U", "V", "W")]
    rng = np.random.default_rng(40)
    force = [rng.standard_normal(a.shape) for a in U]
    return o, X, V, U, force, rng


@pytest.mark.parametrize("flip,integrator", [(0.95, 0), (0.0, 1)])
def test_rollout_adjoint_finite_differences(flip, integrator):
    o, X, V, U, force, rng = _scene(flip, integrator)
    cX, cV = rng.standard_normal(X.shape), rng.standard_normal(V.shape)
    cU = [rng.standard_normal(a.shape) for a in U]

    def loss(Xn, Vn, Un):
        return float(np.sum(cX * Xn) + np.sum(cV * Vn) + sum(np.sum(c * u) for c, u in zip(cU, Un)))

    (Xn, Vn, Un), (gX, gV, gU, gF) = rollout_grad(o, DT, X, V, U, 2, lambda *s: (cX, cV, cU), force)
    L0 = loss(Xn, Vn, Un)
    assert np.isfinite(L0)

    def run(X1, V1, U1, F1):
        (a, b, c), _ = rollout_grad(o, DT, X1, V1, U1, 2, lambda *s: (cX, cV, cU), F1)
        return loss(a, b, c)

    eps = 1e-6
    for what in ("U", "X", "V", "F"):
        for _ in range(2):
            if what == "U":
                d = [rng.standard_normal(a.shape) for a in U]
                lp = run(X, V, [a + eps * b for a, b in zip(U, d)], force)
                lm = run(X, V, [a - eps * b for a, b in zip(U, d)], force)
                an = sum(float(np.sum(g * b)) for g, b in zip(gU, d))
            elif what == "F":
                d = [rng.standard_normal(a.shape) for a in force]
                lp = run(X, V, U, [a + eps * b for a, b in zip(force, d)])
                lm = run(X, V, U, [a - eps * b for a, b in zip(force, d)])
                an = sum(float(np.sum(g * b)) for g, b in zip(gF, d))
            else:
                d = rng.standard_normal(X.shape) * (0.01 if what == "X" else 1.0)
                if what == "X":
                    lp, lm = run(X + eps * d, V, U, force), run(X - eps * d, V, U, force)
                    an = float(np.sum(gX * d))
                else:
                    lp, lm = run(X, V + eps * d, U, force), run(X, V - eps * d, U, force)
                    an = float(np.sum(gV * d))
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - an) <= 1e-6 * abs(an) + 1e-9, (what, fd, an)
