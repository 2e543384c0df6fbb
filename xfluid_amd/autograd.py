"""torch.autograd bindings of the adjoint kernels.

`project(sim, u, v, w)` applies the pressure projection of the sim's current
material to three MAC face arrays (float64 tensors on the GPU) and
differentiates through it with the hand-written adjoint
(xf_sim_project_adjoint: a transposed-operator PCG solve).  The sim acts as the
projection context: each call overwrites its velocity and pressure fields.

`checkpointed(step, state, steps, every)` runs `step` repeatedly keeping only
every `every`-th state for the backward pass and recomputing the others.
"""
import torch
import torch.utils.checkpoint as _ckpt


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sim, u, v, w):
        out = [t.detach().clone().contiguous() for t in (u, v, w)]
        torch.cuda.synchronize()  # the library runs on its own HIP stream
        st = sim.project_fields(*out)
        if st != 0:
            raise RuntimeError(f"projection solve did not converge (status {st})")
        ctx.sim = sim
        return tuple(out)

    @staticmethod
    def backward(ctx, gu, gv, gw):
        g = [t.detach().clone().contiguous() for t in (gu, gv, gw)]
        torch.cuda.synchronize()
        st = ctx.sim.project_adjoint(*g)
        if st != 0:
            raise RuntimeError(f"adjoint solve did not converge (status {st})")
        return None, g[0], g[1], g[2]


def project(sim, u, v, w):
    """Differentiable pressure projection of (u, v, w) on the sim's material."""
    return _Project.apply(sim, u, v, w)


def checkpointed(step, state, steps, every=4):
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
