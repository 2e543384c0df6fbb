"""GPU adjoint kernels against the FP64 CPU adjoint, the dot-product identity
and central finite differences; torch.autograd binding and substep
checkpointing."""
import sys

import numpy as np
import pytest

from oracle_lib import FIELD, oracle
from test_adjoint import (MAXIT, SIGMA, TAU, TOL, dam_material, dot, project, project_adjoint,
                          rand_faces)

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    import xfluid_amd as xf
    if xf.device_count() < 1:
        pytest.fail("no HIP device visible to libxfluid_hip.so")
    mat = dam_material()
    nk, nj, ni = mat.shape
    p = xf.default_params(pcg_tol=TOL, pcg_max_iter=MAXIT, mic_tau=TAU, mic_sigma=SIGMA, precision=0)
    s = xf.FluidSimulator(ni, nj, nk, p)
    s.setCellWidth(0.1)
    s.set_field(FIELD["MATERIAL"], mat)
    return xf, s, mat


def _rel(a, b):
    num = max(float(np.abs(x - y).max()) for x, y in zip(a, b))
    den = max(float(np.abs(y).max()) for y in b)
    return num / den


def test_projection_matches_oracle(setup):
    xf, s, mat = setup
    say("[adjoint] projection forward vs oracle")
    rng = np.random.default_rng(11)
    x = rand_faces(rng, mat)
    g = [a.copy() for a in x]
    assert s.project_fields(*g) == 0
    assert _rel(g, project(mat, *x)) <= 1e-10


def test_projection_adjoint_matches_oracle(setup):
    xf, s, mat = setup
    say("[adjoint] projection adjoint vs FP64 CPU adjoint")
    rng = np.random.default_rng(12)
    y = rand_faces(rng, mat)
    g = [a.copy() for a in y]
    assert s.project_adjoint(*g) == 0
    assert _rel(g, project_adjoint(mat, *y)) <= 1e-10
    # dot-product identity with the GPU forward
    x = rand_faces(rng, mat)
    px = [a.copy() for a in x]
    assert s.project_fields(*px) == 0
    lhs, rhs = dot(px, y), dot(x, g)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_torch_autograd_and_checkpointing(setup):
    xf, s, mat = setup
    import torch
    from xfluid_amd.autograd import project as tproject, rollout
    say("[adjoint] torch autograd: finite differences + checkpointed rollout")
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(13)
    u0 = [torch.tensor(a, device=dev) for a in rand_faces(rng, mat)]
    masks = []
    for t in u0:
        m = torch.zeros_like(t)
        a, b, c = t.shape
        m[a // 4: 3 * a // 4, b // 4: 3 * b // 4, c // 4: 3 * c // 4] = 1.0
        masks.append(m)
    gravity = -9.81 / 60

    def step(u, v, w):
        return tproject(s, u, v + gravity, w)

    def loss_fn(state, steps, every):
        out = rollout(step, state, steps, every)
        return 0.5 * sum((m * o * o).sum() for m, o in zip(masks, out))

    leaves = [t.clone().requires_grad_(True) for t in u0]
    loss = loss_fn(leaves, 6, 2)
    loss.backward()
    grad = [t.grad.detach().cpu().numpy() for t in leaves]
    # the same rollout without checkpointing gives the same gradient
    leaves2 = [t.clone().requires_grad_(True) for t in u0]
    loss2 = loss_fn(leaves2, 6, 6)
    loss2.backward()
    assert abs(loss.item() - loss2.item()) <= 1e-12 * abs(loss.item())
    assert _rel([t.grad.detach().cpu().numpy() for t in leaves2], grad) <= 1e-10
    # central finite differences along random directions (FP64 forward on the GPU)
    eps = 1e-4
    with torch.no_grad():
        for _ in range(4):
            d = [torch.tensor(a, device=dev) for a in rand_faces(rng, mat)]
            lp = loss_fn([a + eps * b for a, b in zip(u0, d)], 6, 6).item()
            lm = loss_fn([a - eps * b for a, b in zip(u0, d)], 6, 6).item()
            fd = (lp - lm) / (2 * eps)
            an = dot(grad, [t.cpu().numpy() for t in d])
            assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


@pytest.mark.parametrize("integrator,kernel", [(0, 0), (1, 1)])
def test_particle_adjoints_match_oracle(integrator, kernel):
    """G2P sampling, RK advection and P2G adjoints equal the FP64 CPU adjoints
    (which are pinned by finite differences) up to the atomic-add summation order."""
    import xfluid_amd as xf
    from test_adjoint import faces_of, particle_scene
    say(f"[adjoint] particle transfers integrator={integrator} kernel={kernel}")
    o, P, V = particle_scene(integrator=integrator, kernel=kernel)
    L = oracle()
    n = len(P)
    p = xf.default_params(integrator=integrator, p2g_kernel=kernel)
    s = xf.FluidSimulator(o.ni, o.nj, o.nk, p)
    s.setCellWidth(o.dx)
    s.set_field(FIELD["MATERIAL"], o.get(FIELD["MATERIAL"]))  # (the advection's collision branch reads it)
    F = faces_of(o)
    for name, a in zip(("U", "V", "W"), F):
        s.set_field(FIELD[name], a)
    rng = np.random.default_rng(31)
    c = rng.standard_normal((n, 3))
    dt = 1.0 / 60

    def zeros():
        return [np.zeros_like(a) for a in F]

    # sampling
    go, gg = np.zeros((n, 3)), zeros()
    L.xo_sample_adjoint(o.h, n, P, c, go, *gg, 0)
    gd, gdg = np.zeros((n, 3)), zeros()
    s.sample_adjoint(P, c, gd, *gdg)
    assert _rel([gd], [go]) <= 1e-12 and _rel(gdg, gg) <= 1e-12
    # advection
    go, gg = np.zeros((n, 3)), zeros()
    L.xo_advect_adjoint(o.h, n, P, dt, c, go, *gg)
    gd, gdg = np.zeros((n, 3)), zeros()
    s.advect_adjoint(P, dt, c, gd, *gdg)
    assert _rel([gd], [go]) <= 1e-12 and _rel(gdg, gg) <= 1e-12
    # P2G
    cf = [rng.standard_normal(a.shape) for a in F]
    gvo, gpo = np.zeros((n, 3)), np.zeros((n, 3))
    L.xo_p2g_adjoint(o.h, n, P, V, *cf, gvo, gpo)
    gvd, gpd = np.zeros((n, 3)), np.zeros((n, 3))
    s.p2g_adjoint(P, V, *cf, gvd, gpd)
    assert _rel([gvd], [gvo]) <= 1e-12 and _rel([gpd], [gpo]) <= 1e-11


def test_sphere_scene_projection_matches_oracle():
    """The projection benchmark's scene (union of seeded spheres in a walled
    cube, float32 face velocities) at 24^3: FP64 and mixed-precision GPU
    projection against the oracle's FP64 projection."""
    import xfluid_amd as xf
    from xfluid_amd.scenes import sphere_scene
    mat, u, v, w = sphere_scene(24, 6, 2)
    say(f"[projection] 24^3 sphere scene, {int((mat == 1).sum())} liquid cells")
    ref = project(mat, u, v, w)
    # (FP64 to the oracle's tolerance; mixed precision to 1e-9 relative residual)
    for prec, ptol, tol in ((0, TOL, 1e-10), (1, 1e-9, 1e-6)):
        p = xf.default_params(pcg_tol=ptol, pcg_max_iter=MAXIT, mic_tau=TAU, mic_sigma=SIGMA, precision=prec)
        s = xf.FluidSimulator(24, 24, 24, p)
        s.setCellWidth(1.0 / 24)
        s.set_field(FIELD["MATERIAL"], mat)
        g = [a.copy() for a in (u, v, w)]
        assert s.project_fields(*g) == 0
        assert _rel(g, ref) <= tol, (prec, _rel(g, ref))
        s.close()


def test_density_adjoint_matches_oracle():
    """GPU density VJP against the FP64 CPU adjoint (same vertex order)."""
    import xfluid_amd as xf
    from test_adjoint import particle_scene
    o, P, V = particle_scene()
    say(f"[adjoint] density VJP, {len(P)} particles")
    from scenes import gpu_like
    g = gpu_like(o)
    o.stage(0)
    rng = np.random.default_rng(31)
    c = np.ascontiguousarray(rng.standard_normal(o.get(FIELD["PHI"]).shape))
    ref = np.zeros((len(P), 3))
    oracle().xo_density_adjoint(o.h, len(P), P, c, ref)
    got = np.zeros((len(P), 3))
    assert g.density_adjoint(P, c, got) == 0
    assert np.array_equal(got, ref), float(np.abs(got - ref).max())
