"""torch.autograd bindings of the adjoint kernels.

All tensors are float64 on the GPU; the simulator acts as the context that
holds the grid geometry, the solid cells, the constant forces and the
parameters, and its fields are overwritten by every call.

* `project(sim, u, v, w)`: the pressure projection of a fixed material (the
  sim's material when called) applied to three MAC face arrays; the material
  is snapshot in the forward and restored for the backward (a transposed-
  operator PCG solve, xf_sim_project_adjoint).
* `substep(sim, dt, X, V, u, v, w, force)`: one substep of FluidSimulator::
  timeStep (FluidSimulator.cpp:265-335) without the density field, on
  particles X, V (n x 3, id order) and face velocities u, v, w: material from
  the particles, save, forces (the sim's constant forces, plus `force`, an
  optional per-face acceleration on the faces next to fluid), projection,
  extrapolation, G2P FLIP/PIC blend, RK4/RK2 advection with collisions, P2G.
  The forward runs the library's own stages and keeps the substep's material,
  extrapolated field and moved particles; the backward puts that state back
  into the sim and runs the adjoint kernels in reverse (P2G, advection with
  its collision branch, G2P sampling of the new and the saved field,
  extrapolation, projection).  Material changes and the P2G normalisation's
  cut-off are piecewise constant and not differentiated.
* `density(sim, X)`: the particle-blob density on the sub-vertex lattice
  (DirectedDistanceField::updateFeild) and its VJP.
* `rollout(step, state, steps, every)`: repeated substeps keeping only every
  `every`-th state for the backward pass and recomputing the others.
"""
import torch
import torch.utils.checkpoint as _ckpt

from . import FIELD, STAGE

_F64 = torch.float64
NOALLOC_MAX_FACES = 2162688  # XF_P2G_ADJOINT_NOALLOC_MAX_FACES (include/xfluid.h)


def _sync():
    # torch work on its stream must be done before the library reads it, and the
    # library's entry points return with their own stream drained
    torch.cuda.synchronize()


def _check(status, what):
    if status != 0:
        raise RuntimeError(f"{what}: pressure solve did not converge (status {status})")


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sim, u, v, w):
        out = [t.detach().clone().contiguous() for t in (u, v, w)]
        mat = torch.empty((sim.nk, sim.nj, sim.ni), dtype=torch.uint8, device=u.device)
        _sync()
        sim.get_field_device(FIELD["MATERIAL"], mat)
        _check(sim.project_fields(*out), "projection")
        ctx.sim = sim
        ctx.save_for_backward(mat)
        return tuple(out)

    @staticmethod
    def backward(ctx, gu, gv, gw):
        (mat,) = ctx.saved_tensors
        g = [t.detach().clone().contiguous() for t in (gu, gv, gw)]
        _sync()
        ctx.sim.set_field_device(FIELD["MATERIAL"], mat)  # the forward's material, whatever ran since
        _check(ctx.sim.project_adjoint(*g), "projection adjoint")
        return None, g[0], g[1], g[2]


def project(sim, u, v, w):
    """Differentiable pressure projection of (u, v, w) on the sim's material."""
    return _Project.apply(sim, u, v, w)


def face_masks(mat):
    """Faces with a fluid cell on either side (where FluidSimulator::_applyForces
    acts, FluidSimulator.cpp:429-499), for a (nk, nj, ni) material tensor."""
    f = mat == 1
    pad = torch.nn.functional.pad
    mu = pad(f, (1, 0)) | pad(f, (0, 1))
    mv = pad(f, (0, 0, 1, 0)) | pad(f, (0, 0, 0, 1))
    mw = pad(f, (0, 0, 0, 0, 1, 0)) | pad(f, (0, 0, 0, 0, 0, 1))
    return [m.to(_F64) for m in (mu, mv, mw)]


def _faces_like(sim, device):
    ni, nj, nk = sim.ni, sim.nj, sim.nk
    return [torch.zeros(s, dtype=_F64, device=device) for s in ((nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni))]


def _forward_to_extrapolation(sim, dt, X, V, U, force, material=None):
    """Runs material .. extrapolation on the sim from the given state (the
    material from the particles, or `material` when given); returns the
    material and the face masks (masks only when a force is given)."""
    _sync()
    sim.set_particles_device(X, V)
    for name, t in zip(("U", "V", "W"), U):
        sim.set_field_device(FIELD[name], t)
    if material is None:
        sim.stage(STAGE["MATERIAL"])
        mat = torch.empty((sim.nk, sim.nj, sim.ni), dtype=torch.uint8, device=X.device)
        sim.get_field_device(FIELD["MATERIAL"], mat)
    else:
        mat = material.contiguous()
        sim.set_field_device(FIELD["MATERIAL"], mat)
    sim.stage(STAGE["SAVE"])
    sim.stage(STAGE["FORCES"], dt)
    masks = None
    if force is not None:
        masks = face_masks(mat)
        cur = _faces_like(sim, X.device)
        for name, t in zip(("U", "V", "W"), cur):
            sim.get_field_device(FIELD[name], t)
        for t, m, f in zip(cur, masks, force):
            t.add_(m * f * dt)
        _sync()
        for name, t in zip(("U", "V", "W"), cur):
            sim.set_field_device(FIELD[name], t)
    _check(sim.stage(STAGE["SOLVE"]), "substep")
    sim.stage(STAGE["APPLY"])
    sim.stage(STAGE["EXTRAPOLATE"])
    return mat, masks


class _Substep(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sim, dt, log, material, X, V, u, v, w, fu, fv, fw):
        p = sim.params
        if p.tail_skip:
            raise ValueError("substep: set tail_skip = 0 (every particle moves)")
        force = None if fu is None else (fu, fv, fw)
        X, V = X.detach().contiguous(), V.detach().contiguous()
        U = [t.detach().contiguous() for t in (u, v, w)]
        mat, masks = _forward_to_extrapolation(sim, dt, X, V, U, force, material)
        if log is not None:
            log.append(mat)
        U3 = _faces_like(sim, X.device)  # the extrapolated field the particles move in
        for name, t in zip(("U", "V", "W"), U3):
            sim.get_field_device(FIELD[name], t)
        sim.stage(STAGE["ADVECT"], dt)
        sim.stage(STAGE["P2G"])
        Xn, Vn = torch.empty_like(X), torch.empty_like(V)
        sim.get_particles_device(Xn, Vn)
        Un = _faces_like(sim, X.device)
        for name, t in zip(("U", "V", "W"), Un):
            sim.get_field_device(FIELD[name], t)
        ctx.sim, ctx.dt, ctx.has_force = sim, dt, force is not None
        # what the backward needs besides the inputs: the material, the
        # extrapolated field and the moved particles (P2G does not move them)
        ctx.save_for_backward(X, V, *U, mat, *U3, Xn, Vn, *(masks if masks is not None else ()))
        return Xn, Vn, Un[0], Un[1], Un[2]

    @staticmethod
    def backward(ctx, gXn, gVn, gun, gvn, gwn):
        sim, dt = ctx.sim, ctx.dt
        saved = ctx.saved_tensors
        X, V, U, mat = saved[0], saved[1], list(saved[2:5]), saved[5]
        U3, Xa, Va = list(saved[6:9]), saved[9], saved[10]
        masks = list(saved[11:14]) if ctx.has_force else None
        dev = X.device
        z = lambda t, like: torch.zeros_like(like) if t is None else t.detach().contiguous()  # noqa: E731
        gXn, gVn = z(gXn, X), z(gVn, V)
        gU = [z(g, t) for g, t in zip((gun, gvn, gwn), U)]
        # the sim's state of this substep: material, extrapolated field, saved field
        _sync()
        sim.set_field_device(FIELD["MATERIAL"], mat)
        for name, t in zip(("U", "V", "W"), U3):
            sim.set_field_device(FIELD[name], t)
        for name, t in zip(("USAVED", "VSAVED", "WSAVED"), U):
            sim.set_field_device(FIELD[name], t)
        # P2G: U' = P2G(X', V')
        gpos = gXn.clone()
        gvel = torch.zeros_like(V)
        ni, nj, nk = sim.ni, sim.nj, sim.nk
        faces = max((ni + 1) * nj * nk, ni * (nj + 1) * nk, ni * nj * (nk + 1))
        p2g_adj = sim.p2g_adjoint_noalloc if faces <= NOALLOC_MAX_FACES else sim.p2g_adjoint
        _sync()
        p2g_adj(Xa, Va, *gU, gvel, gpos)
        gV_new = gVn + gvel
        # advection: X' = move(X; U3)
        gX = torch.zeros_like(X)
        gU3 = _faces_like(sim, dev)
        _sync()
        sim.advect_adjoint(X, dt, gpos, gX, *gU3)
        # G2P: V' = Vn (1 - r) + (V + Vn - Vo) r, Vn = S(U3; X), Vo = S(Us; X)
        r = sim.params.flip_ratio
        g_new_field = gV_new * ((1.0 - r) + r)
        gV = gV_new * r
        _sync()
        sim.sample_adjoint(X, g_new_field.contiguous(), gX, *gU3, saved=False)
        gUs = _faces_like(sim, dev)
        if r != 0.0:
            g_old = (-r * gV_new).contiguous()
            _sync()
            sim.sample_adjoint(X, g_old, gX, *gUs, saved=True)
        # extrapolation and projection of this substep's material (set above)
        sim.extrapolate_adjoint(*gU3)
        _check(sim.project_adjoint(*gU3), "substep adjoint")
        gF = [None, None, None]
        if masks is not None:
            gF = [dt * m * g for m, g in zip(masks, gU3)]
        gUin = [a + b for a, b in zip(gU3, gUs)]
        return (None, None, None, None, gX, gV, gUin[0], gUin[1], gUin[2], gF[0], gF[1], gF[2])


def substep(sim, dt, X, V, u, v, w, force=None, material_log=None, material=None):
    """One differentiable substep; returns (X', V', u', v', w').  `force` is an
    optional triple of per-face accelerations (added times dt on the faces next
    to fluid, after the sim's constant forces).  `material_log`, a list, gets
    the substep's material (uint8 device tensor) appended: the piecewise-
    constant part of the step that the gradient holds fixed.  `material`, when
    given, is used as this substep's material instead of the one the particles
    make (the branch of the step the gradient differentiates)."""
    f = (None, None, None) if force is None else force
    return _Substep.apply(sim, dt, material_log, material, X, V, u, v, w, *f)


class _Density(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sim, X):
        X = X.detach().contiguous()
        d = sim.params.sub_division
        phi = torch.empty((d * sim.nk + 1, d * sim.nj + 1, d * sim.ni + 1), dtype=_F64, device=X.device)
        _sync()
        sim.set_particles_device(X, torch.zeros_like(X))
        sim.stage(STAGE["DENSITY"])
        sim.get_field_device(FIELD["PHI"], phi)
        ctx.sim = sim
        ctx.save_for_backward(X)
        return phi

    @staticmethod
    def backward(ctx, gphi):
        (X,) = ctx.saved_tensors
        gX = torch.zeros_like(X)
        g = gphi.detach().contiguous()
        _sync()
        ctx.sim.density_adjoint(X, g, gX)
        return None, gX


def density(sim, X):
    """Differentiable particle-blob density of the particles X on the sim's lattice."""
    return _Density.apply(sim, X)


def rollout(step, state, steps, every=4):
    """Applies state = step(*state) `steps` times; the backward pass keeps only
    the state at the start of each block of `every` steps and recomputes the
    block (substep checkpointing)."""
    def block(n):
        def run(*s):
            for _ in range(n):
                s = step(*s)
            return s
        return run

    done = 0
    while done < steps:
        n = min(every, steps - done)
        state = _ckpt.checkpoint(block(n), *state, use_reentrant=False)
        done += n
    return state


