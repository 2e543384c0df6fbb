"""Differentiable substeps on the GPU (xfluid_amd.autograd) against the FP64
CPU adjoint of the same chain on the oracle (tests/diff_chain.py, itself
pinned by central differences) and against central differences of the GPU's
own forward:

* the new adjoint kernels one by one: extrapolation (bit-exact with the
  oracle's), the advection move with its collision branch (1e-12);
* device-memory particle and field transfer round trips;
* the projection binding restores its forward's material for the backward;
* three substeps (FLIP 0.95, a per-face force, checkpoint every 2): final
  state and gradients with respect to the initial positions, velocities,
  face velocities and the force equal the CPU adjoint's; a rollout without
  checkpoints gives the same gradient; central differences agree;
* the particle-blob density binding against the oracle's density adjoint;
* the differentiable-projection workload at reduced size (32^3, liquid below
  y = 0.7, curl-noise initial velocity, FLIP 0, 16 substeps, checkpoint every
  4, loss 1/2 |u|^2 over [0.4, 0.6]^3): gradient with respect to the initial
  velocity against the CPU adjoint and central differences (eps = 1e-4).
"""
import sys

import numpy as np
import pytest

import diff_chain
from oracle_lib import FIELD, STAGE, oracle
from scenes import copy_state, gpu_like, oracle_dam_break, small_params

pytestmark = pytest.mark.gpu
NAMES = ("U", "V", "W")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    import torch
    import xfluid_amd as xf
    if xf.device_count() < 1 or not torch.cuda.is_available():
        pytest.fail("no HIP device")
    import xfluid_amd.autograd as ag
    return xf, torch, ag


def _rel(a, b):
    a = [np.asarray(x) for x in (a if isinstance(a, (list, tuple)) else [a])]
    b = [np.asarray(x) for x in (b if isinstance(b, (list, tuple)) else [b])]
    num = max(float(np.abs(x - y).max()) for x, y in zip(a, b))
    den = max(float(np.abs(y).max()) for y in b)
    return num / max(den, 1e-300)


def test_extrapolation_adjoint_matches_oracle(env):
    xf, torch, ag = env
    say("[diff] extrapolation adjoint vs oracle")
    o = oracle_dam_break(21, 40, 35, 0.05, (0.6, 0.7, 1.0), 9, small_params(cfl=5.0))
    o.time_step(1.0 / 60)
    o.stage(STAGE["MATERIAL"])
    g = gpu_like(o)
    copy_state(o, g)
    rng = np.random.default_rng(50)
    y = [rng.standard_normal(o.get(FIELD[n]).shape) for n in NAMES]
    yo = [a.copy() for a in y]
    oracle().xo_extrapolate_adjoint(o.h, *yo)
    yg = [a.copy() for a in y]
    g.extrapolate_adjoint(*yg)
    for a, b, n in zip(yg, yo, NAMES):
        assert np.array_equal(a, b), f"extrapolation adjoint {n}"


def test_advect_adjoint_collisions_match_oracle(env):
    xf, torch, ag = env
    say("[diff] advection move adjoint with collisions vs oracle")
    for integ in (0, 1):
        o = oracle_dam_break(12, 10, 9, 0.1, (0.5, 0.6, 1.0), 7, small_params(integrator=integ))
        for _ in range(2):
            o.time_step(1.0 / 60)
        for n, push in zip(NAMES, (-1.5, -2.0, 1.0)):
            o.set(FIELD[n], 8.0 * o.get(FIELD[n]) + push)
        g = gpu_like(o)
        copy_state(o, g)
        P, _ = o.particles()
        P = np.ascontiguousarray(P)
        n = len(P)
        dt = 0.05
        moved, target = np.empty((n, 3)), np.empty((n, 3))
        oracle().xo_move_points(o.h, n, P, dt, moved)
        oracle().xo_advect_points(o.h, n, P, dt, target)
        assert (np.any(moved != target, 1) & np.any(moved != P, 1)).sum() > 50
        c = np.random.default_rng(51).standard_normal((n, 3))
        F = [o.get(FIELD[x]) for x in NAMES]
        go, gfo = np.zeros((n, 3)), [np.zeros_like(a) for a in F]
        oracle().xo_advect_adjoint(o.h, n, P, dt, c, go, *gfo)
        gg, gfg = np.zeros((n, 3)), [np.zeros_like(a) for a in F]
        g.advect_adjoint(P, dt, c, gg, *gfg)
        assert _rel(gg, go) <= 1e-12
        assert _rel(gfg, gfo) <= 1e-12


def test_device_transfer_roundtrip(env):
    xf, torch, ag = env
    say("[diff] device particle / field transfer")
    o = oracle_dam_break(14, 12, 10, 0.1, (0.5, 0.6, 1.0), 5, small_params())
    g = gpu_like(o)
    g.timeStep(1.0 / 60)  # particles now sorted by cell inside the library
    P, V = g.particles()
    dP = torch.empty((len(P), 3), dtype=torch.float64, device="cuda")
    dV = torch.empty_like(dP)
    g.get_particles_device(dP, dV)
    assert np.array_equal(dP.cpu().numpy(), P) and np.array_equal(dV.cpu().numpy(), V)
    g.set_particles_device(dP * 1.0, dV * 2.0)
    P2, V2 = g.particles()
    assert np.array_equal(P2, P) and np.array_equal(V2, 2.0 * V)
    u = torch.tensor(g.get_field(FIELD["U"]), device="cuda")
    g.set_field_device(FIELD["USAVED"], u * 3.0)
    assert np.array_equal(g.get_field(FIELD["USAVED"]), 3.0 * g.get_field(FIELD["U"]))
    m = torch.zeros((10, 12, 14), dtype=torch.uint8, device="cuda")
    g.get_field_device(FIELD["MATERIAL"], m)
    assert np.array_equal(m.cpu().numpy(), g.get_field(FIELD["MATERIAL"]))


def test_projection_restores_its_material(env):
    """A material change between forward and backward does not reach the VJP."""
    xf, torch, ag = env
    from test_adjoint import TOL, MAXIT, dam_material, project_adjoint, rand_faces
    say("[diff] projection binding: material snapshot")
    mat = dam_material()
    nk, nj, ni = mat.shape
    s = xf.FluidSimulator(ni, nj, nk, xf.default_params(pcg_tol=TOL, pcg_max_iter=MAXIT))
    s.setCellWidth(0.1)
    s.set_field(FIELD["MATERIAL"], mat)
    rng = np.random.default_rng(52)
    u0 = [torch.tensor(a, device="cuda", requires_grad=True) for a in rand_faces(rng, mat)]
    out = ag.project(s, *u0)
    other = mat.copy()
    other[other == 1] = 0  # every fluid cell turned to air before the backward
    s.set_field(FIELD["MATERIAL"], other)
    y = rand_faces(rng, mat)
    sum((o_ * torch.tensor(b, device="cuda")).sum() for o_, b in zip(out, y)).backward()
    ref = project_adjoint(mat, *y)
    assert _rel([t.grad.cpu().numpy() for t in u0], ref) <= 1e-10


def _torch_state(torch, o):
    X, V = o.particles()
    T = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")  # noqa: E731
    return T(X), T(V), [T(o.get(FIELD[n])) for n in NAMES]


def _diff_scene(xf, dims, dx, frac, seed, flip):
    p = small_params(flip_ratio=flip, pcg_tol=1e-13, pcg_max_iter=2000, tail_skip=0)
    o = oracle_dam_break(*dims, dx, frac, seed, p)
    for _ in range(2):
        o.time_step(1.0 / 30)
    g = gpu_like(o)
    copy_state(o, g)
    return o, g


@pytest.mark.parametrize("dims,dx,frac,seed", [((14, 12, 10), 0.1, (0.5, 0.6, 1.0), 5),
                                               ((21, 40, 35), 0.05, (0.6, 0.7, 1.0), 9)])
def test_substep_rollout_matches_cpu_adjoint(env, dims, dx, frac, seed):
    xf, torch, ag = env
    say(f"[diff] 3 substeps {dims}: GPU autograd vs the FP64 CPU adjoint")
    dt = 1.0 / 30
    o, g = _diff_scene(xf, dims, dx, frac, seed, 0.95)
    X0, V0, U0 = _torch_state(torch, o)
    rng = np.random.default_rng(53)
    Xh, Vh = X0.cpu().numpy(), V0.cpu().numpy()
    Uh = [t.cpu().numpy() for t in U0]
    Fh = [rng.standard_normal(a.shape) for a in Uh]
    cX, cV = rng.standard_normal(Xh.shape), rng.standard_normal(Vh.shape)
    cU = [rng.standard_normal(a.shape) for a in Uh]
    (Xo, Vo, Uo), (gX, gV, gU, gF) = diff_chain.rollout_grad(o, dt, Xh, Vh, Uh, 3, lambda *s: (cX, cV, cU), Fh)
    T = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731

    def run(every):
        leaves = [X0.clone().requires_grad_(True), V0.clone().requires_grad_(True)] + \
                 [t.clone().requires_grad_(True) for t in U0] + [T(a).requires_grad_(True) for a in Fh]
        F = tuple(leaves[5:8])
        step = lambda X, V, u, v, w: ag.substep(g, dt, X, V, u, v, w, F)  # noqa: E731
        X, V, u, v, w = ag.rollout(step, tuple(leaves[:5]), 3, every)
        loss = (T(cX) * X).sum() + (T(cV) * V).sum() + sum((T(c) * t).sum() for c, t in zip(cU, (u, v, w)))
        loss.backward()
        return (X, V, (u, v, w)), [t.grad.detach().cpu().numpy() for t in leaves]

    (X, V, U), grads = run(2)
    assert _rel(X.detach().cpu().numpy(), Xo) <= 1e-12
    assert _rel([t.detach().cpu().numpy() for t in U], Uo) <= 1e-10
    errs = dict(X=_rel(grads[0], gX), V=_rel(grads[1], gV), U=_rel(grads[2:5], gU), F=_rel(grads[5:8], gF))
    say("[diff] gradient vs CPU adjoint: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    for k, v in errs.items():
        assert v <= 1e-8, (k, v)
    _, grads2 = run(3)  # no intermediate checkpoint
    assert _rel(grads2, grads) <= 1e-12


def test_substep_finite_differences_gpu(env):
    xf, torch, ag = env
    say("[diff] 3 substeps: GPU gradient vs central differences of the GPU forward")
    dt = 1.0 / 30
    o, g = _diff_scene(xf, (14, 12, 10), 0.1, (0.5, 0.6, 1.0), 5, 0.95)
    X0, V0, U0 = _torch_state(torch, o)
    rng = np.random.default_rng(54)
    cU = [torch.tensor(rng.standard_normal(t.shape), device="cuda") for t in U0]

    def loss_of(U, grad=False):
        leaves = [t.clone().requires_grad_(grad) for t in U]
        step = lambda X, V, u, v, w: ag.substep(g, dt, X, V, u, v, w)  # noqa: E731
        _, _, u, v, w = ag.rollout(step, (X0, V0, *leaves), 3, 2)
        loss = sum((c * t * t).sum() for c, t in zip(cU, (u, v, w)))
        if grad:
            loss.backward()
            return loss.item(), [t.grad for t in leaves]
        return loss.item()

    _, grad = loss_of(U0, True)
    eps = 1e-6
    for _ in range(3):
        d = [torch.randn_like(t) for t in U0]
        with torch.no_grad():
            fd = (loss_of([a + eps * b for a, b in zip(U0, d)]) - loss_of([a - eps * b for a, b in zip(U0, d)])) / (2 * eps)
        an = sum(float((a * b).sum()) for a, b in zip(grad, d))
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_density_binding_matches_oracle(env):
    xf, torch, ag = env
    say("[diff] density binding vs oracle density adjoint")
    o = oracle_dam_break(14, 12, 10, 0.1, (0.5, 0.6, 1.0), 5, small_params())
    g = gpu_like(o)
    P, _ = o.particles()
    X = torch.tensor(P, device="cuda", requires_grad=True)
    phi = ag.density(g, X)
    o.stage(STAGE["DENSITY"])
    assert _rel(phi.detach().cpu().numpy(), o.get(FIELD["PHI"])) <= 1e-12
    c = np.random.default_rng(55).standard_normal(tuple(phi.shape))
    (phi * torch.tensor(c, device="cuda")).sum().backward()
    ref = np.zeros_like(P)
    oracle().xo_density_adjoint(o.h, len(P), np.ascontiguousarray(P), np.ascontiguousarray(c), ref)
    assert _rel(X.grad.cpu().numpy(), ref) <= 1e-12


def test_differentiable_projection_workload_small(env):
    """The differentiable-projection workload at 32^3."""
    xf, torch, ag = env
    from xfluid_amd.scenes import curl_noise_faces
    n, steps, every, dt = 32, 16, 4, 1.0 / 240
    dx = 1.0 / n
    say(f"[diff] differentiable projection {n}^3, {steps} substeps, checkpoint every {every}")
    prm = dict(flip_ratio=0.0, pcg_tol=1e-13, pcg_max_iter=2000, tail_skip=0)
    o = oracle_dam_break(n, n, n, dx, (1.0, 0.7, 1.0), 3, small_params(**prm))
    g = gpu_like(o)
    U0h = list(curl_noise_faces(n, n, n, dx, seed=3))
    X0h, V0h = o.particles()
    T = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")  # noqa: E731
    X0, V0 = T(X0h), T(V0h)
    # faces inside [0.4, 0.6]^3
    boxes = []
    for comp, shape in enumerate([(n, n, n + 1), (n, n + 1, n), (n + 1, n, n)]):
        k, j, i = np.meshgrid(*(np.arange(s) for s in shape), indexing="ij")
        off = [0.5, 0.5, 0.5]
        off[comp] = 0.0
        x, y, z = (i + off[0]) * dx, (j + off[1]) * dx, (k + off[2]) * dx
        boxes.append(((x >= 0.4) & (x <= 0.6) & (y >= 0.4) & (y <= 0.6) & (z >= 0.4) & (z <= 0.6)).astype(np.float64))
    Bd = [T(b) for b in boxes]

    def loss_of(U, grad=False, log=None, frozen=None):
        leaves = [T(a).requires_grad_(grad) for a in U]
        mats = iter(frozen) if frozen is not None else None
        step = lambda X, V, u, v, w: ag.substep(g, dt, X, V, u, v, w, material_log=log,  # noqa: E731
                                                material=next(mats) if mats is not None else None)
        _, _, u, v, w = ag.rollout(step, (X0, V0, *leaves), steps, every)
        loss = 0.5 * sum((b * t * t).sum() for b, t in zip(Bd, (u, v, w)))
        if grad:
            loss.backward()
            return loss.item(), [t.grad.cpu().numpy() for t in leaves]
        return loss.item()

    L0, grad = loss_of(U0h, True)
    with torch.no_grad():
        mats0 = []
        loss_of(U0h, log=mats0)
    (_, _, Uo), (_, _, gU, _) = diff_chain.rollout_grad(
        o, dt, X0h, V0h, U0h, steps, lambda X, V, U: (np.zeros_like(X), np.zeros_like(V), [b * u for b, u in zip(boxes, U)]))
    err_cpu = _rel(grad, gU)
    say(f"[diff] loss {L0:.6e}; gradient vs FP64 CPU adjoint {err_cpu:.2e}")
    assert err_cpu <= 1e-8
    # central differences along seeded directions.  At eps = 1e-4 with each
    # substep's material held at the unperturbed run's (the branch of the step
    # the gradient differentiates); and at eps = 1e-4 and 1e-6 with the
    # material the particles make, where a direction whose +eps or -eps run
    # changes some substep's material (a cell gains its first or loses its last
    # particle: a jump of the loss) is an outlier, not compared.  The rest differ
    # by the kinks of trilinear sampling (a particle crossing a cell face between
    # the two runs), which shrink with eps.
    rng = np.random.default_rng(3)
    gnorm = float(np.sqrt(sum(np.sum(a * a) for a in grad)))
    res = {"frozen 1e-4": [], "1e-4": [], "1e-6": []}
    outliers = 0
    for _ in range(8):
        d = [rng.standard_normal(a.shape) for a in U0h]
        an = sum(float(np.sum(a * b)) for a, b in zip(grad, d))
        row, flips = {}, 0
        for key, eps in (("frozen 1e-4", 1e-4), ("1e-4", 1e-4), ("1e-6", 1e-6)):
            mp, mm = [], []
            frozen = mats0 if key.startswith("frozen") else None
            with torch.no_grad():
                lp = loss_of([a + eps * b for a, b in zip(U0h, d)], log=mp, frozen=frozen)
                lm = loss_of([a - eps * b for a, b in zip(U0h, d)], log=mm, frozen=frozen)
            if frozen is None:
                flips += sum(int((a != b).sum()) for a, b in zip(mp + mm, mats0 + mats0))
            row[key] = (lp - lm) / (2 * eps)
        res["frozen 1e-4"].append((abs(row["frozen 1e-4"] - an) / abs(an), abs(row["frozen 1e-4"] - an) / gnorm))
        if flips:
            outliers += 1
            say(f"[diff]   direction with {flips} material changes: fd {row['1e-4']:.6e}, adjoint {an:.6e} (outlier)")
            continue
        for key in ("1e-4", "1e-6"):
            res[key].append((abs(row[key] - an) / abs(an), abs(row[key] - an) / gnorm))
    for key, r in res.items():
        say(f"[diff] central differences {key}: {len(r)} directions, worst relative error "
            f"{max(x[0] for x in r):.2e}, worst error / |grad| {max(x[1] for x in r):.2e}")
    say(f"[diff] {outliers} outlier directions")
    assert len(res["1e-6"]) >= 5
    assert max(x[0] for x in res["1e-6"]) <= 1e-4
    assert max(x[0] for x in res["frozen 1e-4"]) <= 1e-3
