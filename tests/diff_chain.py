"""The FP64 CPU adjoint of the differentiable substep (xfluid_amd.autograd.substep),
run on the oracle: the same chain of stages forward and the oracle's adjoints
of P2G, the advection move (with its collision branch), G2P sampling,
extrapolation and projection backward.  Test infrastructure only.
"""
import numpy as np

from oracle_lib import FIELD, STAGE, oracle

NAMES = ("U", "V", "W")


def face_masks(mat):
    """Faces with a fluid cell on either side, for a (nk, nj, ni) material."""
    f = mat == 1
    mu = np.pad(f, ((0, 0), (0, 0), (1, 0))) | np.pad(f, ((0, 0), (0, 0), (0, 1)))
    mv = np.pad(f, ((0, 0), (1, 0), (0, 0))) | np.pad(f, ((0, 0), (0, 1), (0, 0)))
    mw = np.pad(f, ((1, 0), (0, 0), (0, 0))) | np.pad(f, ((0, 1), (0, 0), (0, 0)))
    return [m.astype(np.float64) for m in (mu, mv, mw)]


def _to_extrapolation(o, dt, X, V, U, force):
    o.set_particles(X, V)
    for name, a in zip(NAMES, U):
        o.set(FIELD[name], a)
    o.stage(STAGE["MATERIAL"])
    o.stage(STAGE["SAVE"])
    o.stage(STAGE["FORCES"], dt)
    mat = o.get(FIELD["MATERIAL"])
    masks = None
    if force is not None:
        masks = face_masks(mat)
        for name, m, f in zip(NAMES, masks, force):
            o.set(FIELD[name], o.get(FIELD[name]) + m * f * dt)
    assert o.stage(STAGE["SOLVE"]) == 0
    o.stage(STAGE["APPLY"])
    o.stage(STAGE["EXTRAPOLATE"])
    return mat, masks


def substep(o, dt, X, V, U, force=None):
    _to_extrapolation(o, dt, X, V, U, force)
    o.stage(STAGE["ADVECT"], dt)
    o.stage(STAGE["P2G"])
    Xn, Vn = o.particles()
    return Xn, Vn, [o.get(FIELD[name]) for name in NAMES]


def substep_vjp(o, dt, X, V, U, force, gXn, gVn, gUn):
    """Returns (gX, gV, gU, gF) for gradients (gXn, gVn, gUn) on the outputs."""
    L = oracle()
    p = o.params
    mat, masks = _to_extrapolation(o, dt, X, V, U, force)
    o.stage(STAGE["ADVECT"], dt)
    Xa, Va = o.particles()
    n = len(X)
    gpos = np.ascontiguousarray(gXn, np.float64).copy()
    gvel = np.zeros((n, 3))
    gU = [np.ascontiguousarray(g, np.float64) for g in gUn]
    L.xo_p2g_adjoint(o.h, n, np.ascontiguousarray(Xa), np.ascontiguousarray(Va), *gU, gvel, gpos)
    gV_new = gVn + gvel
    gX = np.zeros((n, 3))
    gU3 = [np.zeros_like(a) for a in U]
    X = np.ascontiguousarray(X)
    L.xo_advect_adjoint(o.h, n, X, dt, gpos, gX, *gU3)
    r = p.flip_ratio
    L.xo_sample_adjoint(o.h, n, X, np.ascontiguousarray(gV_new * ((1.0 - r) + r)), gX, *gU3, 0)
    gUs = [np.zeros_like(a) for a in U]
    if r != 0.0:
        L.xo_sample_adjoint(o.h, n, X, np.ascontiguousarray(-r * gV_new), gX, *gUs, 1)
    L.xo_extrapolate_adjoint(o.h, *gU3)
    nk, nj, ni = mat.shape
    assert L.xo_project_adjoint(ni, nj, nk, np.ascontiguousarray(mat).ravel(), *gU3, p.pcg_tol, p.pcg_max_iter,
                                p.mic_tau, p.mic_sigma) == 0
    gF = None if force is None else [dt * m * g for m, g in zip(masks, gU3)]
    return gX, gV_new * r, [a + b for a, b in zip(gU3, gUs)], gF


def rollout_grad(o, dt, X, V, U, steps, loss_grad, force=None, progress=None):
    """Runs `steps` substeps from (X, V, U), then the VJP chain backwards with
    the gradient loss_grad(Xn, Vn, Un) -> (gX, gV, gU) of the loss on the final
    state.  Returns the final state and (gX0, gV0, gU0, gF).  progress(msg), if
    given, is called after every substep of either pass."""
    states = [(X, V, U)]
    for t in range(steps):
        states.append(substep(o, dt, *states[-1], force))
        if progress:
            progress(f"forward substep {t + 1}/{steps}")
    g = loss_grad(*states[-1])
    gF = None if force is None else [np.zeros_like(a) for a in force]
    for t in range(steps - 1, -1, -1):
        gX, gV, gU, gf = substep_vjp(o, dt, *states[t], force, *g)
        g = (gX, gV, gU)
        if progress:
            progress(f"backward substep {t + 1}")
        if gF is not None:
            gF = [a + b for a, b in zip(gF, gf)]
    return states[-1], (g[0], g[1], g[2], gF)
