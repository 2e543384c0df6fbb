"""The differentiable-projection workload at its full size (96^3, 4.67M
particles, 16 substeps, checkpoint every 4): the GPU gradient with respect to
the initial face velocity against the FP64 CPU adjoint of the same chain on
the oracle (tests/diff_chain.py).  The CPU side takes several minutes, so the
test runs only with XFLUID_FULL_SIZE=1."""
import os
import sys

import numpy as np
import pytest

import diff_chain
from scenes import gpu_like, oracle_dam_break, small_params

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("XFLUID_FULL_SIZE") != "1",
                                 reason="full-size CPU adjoint: set XFLUID_FULL_SIZE=1")]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_full_size_gradient_matches_cpu_adjoint():
    import time
    import torch
    import xfluid_amd.autograd as ag
    from xfluid_amd.scenes import curl_noise_faces
    n, steps, every, dt = 96, 16, 4, 1.0 / 240
    dx = 1.0 / n
    o = oracle_dam_break(n, n, n, dx, (1.0, 0.7, 1.0), 3,
                         small_params(flip_ratio=0.0, pcg_tol=1e-12, pcg_max_iter=4000, tail_skip=0))
    g = gpu_like(o)
    U0h = list(curl_noise_faces(n, n, n, dx, seed=3))
    X0h, V0h = o.particles()
    say(f"[full] {n}^3, {len(X0h)} particles")
    T = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")  # noqa: E731
    boxes = []
    for comp, shape in enumerate([(n, n, n + 1), (n, n + 1, n), (n + 1, n, n)]):
        k, j, i = np.meshgrid(*(np.arange(s) for s in shape), indexing="ij")
        off = [0.5, 0.5, 0.5]
        off[comp] = 0.0
        x, y, z = (i + off[0]) * dx, (j + off[1]) * dx, (k + off[2]) * dx
        boxes.append(((x >= 0.4) & (x <= 0.6) & (y >= 0.4) & (y <= 0.6) & (z >= 0.4) & (z <= 0.6)).astype(np.float64))
    leaves = [T(a).requires_grad_(True) for a in U0h]
    step = lambda X, V, u, v, w: ag.substep(g, dt, X, V, u, v, w)  # noqa: E731
    _, _, u, v, w = ag.rollout(step, (T(X0h), T(V0h), *leaves), steps, every)
    loss = 0.5 * sum((T(b) * t * t).sum() for b, t in zip(boxes, (u, v, w)))
    loss.backward()
    grad = [t.grad.cpu().numpy() for t in leaves]
    t0 = time.time()
    (_, _, Uo), (_, _, gU, _) = diff_chain.rollout_grad(
        o, dt, X0h, V0h, U0h, steps, lambda X, V, U: (np.zeros_like(X), np.zeros_like(V), [b * q for b, q in zip(boxes, U)]),
        progress=lambda m: say(f"[full] CPU adjoint: {m} ({time.time() - t0:.0f} s)"))
    Lo = 0.5 * sum(float(np.sum(b * q * q)) for b, q in zip(boxes, Uo))
    num = max(float(np.abs(a - b).max()) for a, b in zip(grad, gU))
    den = max(float(np.abs(b).max()) for b in gU)
    say(f"[full] loss GPU {loss.item():.12e} CPU {Lo:.12e}; gradient vs FP64 CPU adjoint {num / den:.2e} "
        f"(CPU {time.time() - t0:.0f} s)")
    assert abs(loss.item() - Lo) <= 1e-10 * abs(Lo)
    assert num / den <= 1e-8
