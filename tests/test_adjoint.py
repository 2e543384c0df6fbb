"""Adjoint (vector-Jacobian) kernels of the substep, checked on the CPU oracle.

The reference has no gradients; these operators are new.  Each adjoint is
pinned by properties that a sign, index or transposition mistake breaks:
  * the dot-product identity <J x, y> = <x, J^T y> on random x, y;
  * central finite differences (eps = 1e-4) of a nonlinear loss along random
    directions (FP64 forward, relative error <= 1e-6).
"""
import numpy as np
import pytest

from oracle_lib import FIELD, oracle
from scenes import oracle_dam_break, small_params

TOL, MAXIT, TAU, SIGMA = 1e-14, 2000, 0.97, 0.25


def dam_material(dims=(14, 12, 10), steps=2, seed=3):
    o = oracle_dam_break(*dims, 0.1, (0.5, 0.6, 1.0), seed, small_params())
    for _ in range(steps):
        o.time_step(1.0 / 60)
    o.stage(1)  # material from the current particles
    return o.get(FIELD["MATERIAL"])


def face_shapes(ni, nj, nk):
    return [(nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)]


def project(mat, u, v, w):
    nk, nj, ni = mat.shape
    out = [a.copy() for a in (u, v, w)]
    st = oracle().xo_project(ni, nj, nk, mat.ravel(), *out, TOL, MAXIT, TAU, SIGMA)
    assert st == 0
    return out


def project_adjoint(mat, gu, gv, gw):
    nk, nj, ni = mat.shape
    out = [a.copy() for a in (gu, gv, gw)]
    st = oracle().xo_project_adjoint(ni, nj, nk, mat.ravel(), *out, TOL, MAXIT, TAU, SIGMA)
    assert st == 0
    return out


def rand_faces(rng, mat):
    nk, nj, ni = mat.shape
    return [rng.standard_normal(s) for s in face_shapes(ni, nj, nk)]


def dot(a, b):
    return sum(float(np.dot(x.ravel(), y.ravel())) for x, y in zip(a, b))


def test_projection_adjoint_dot_identity():
    mat = dam_material()
    rng = np.random.default_rng(1)
    for _ in range(3):
        x, y = rand_faces(rng, mat), rand_faces(rng, mat)
        lhs = dot(project(mat, *x), y)
        rhs = dot(x, project_adjoint(mat, *y))
        assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1.0), (lhs, rhs)


def test_projection_adjoint_finite_differences():
    """L(u) = 1/2 |P u|^2 over a box of faces; dL/du = P^T (mask * P u)."""
    mat = dam_material()
    nk, nj, ni = mat.shape
    rng = np.random.default_rng(2)
    u0 = rand_faces(rng, mat)
    masks = []
    for s in face_shapes(ni, nj, nk):
        m = np.zeros(s)
        m[s[0] // 4: 3 * s[0] // 4, s[1] // 4: 3 * s[1] // 4, s[2] // 4: 3 * s[2] // 4] = 1.0
        masks.append(m)

    def loss(u):
        pu = project(mat, *u)
        return 0.5 * sum(float(np.sum(m * a * a)) for m, a in zip(masks, pu))

    pu = project(mat, *u0)
    grad = project_adjoint(mat, *[m * a for m, a in zip(masks, pu)])
    eps = 1e-4
    for _ in range(8):
        d = rand_faces(rng, mat)
        fd = (loss([a + eps * b for a, b in zip(u0, d)]) - loss([a - eps * b for a, b in zip(u0, d)])) / (2 * eps)
        an = dot(grad, d)
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_projection_adjoint_solid_and_air_faces():
    """A single fluid cell between a wall and air exercises every apply/RHS
    case (solid row, air ghost); the dot identity holds exactly up to the solve."""
    ni = nj = nk = 5
    mat = np.zeros((nk, nj, ni), np.uint8)
    mat[0, :, :] = mat[-1, :, :] = mat[:, 0, :] = mat[:, -1, :] = mat[:, :, 0] = mat[:, :, -1] = 2
    mat[2, 1, 2] = 1          # fluid on the floor
    mat[2, 1, 3] = 1          # and its +i neighbour
    rng = np.random.default_rng(4)
    x, y = rand_faces(rng, mat), rand_faces(rng, mat)
    lhs = dot(project(mat, *x), y)
    rhs = dot(x, project_adjoint(mat, *y))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


# ---------------------------------------------------------------------------
# particle transfers: G2P sampling, RK advection, P2G
def particle_scene(integrator=0, kernel=0):
    p = small_params(integrator=integrator, p2g_kernel=kernel)
    o = oracle_dam_break(12, 10, 9, 0.1, (0.5, 0.6, 1.0), 7, p)
    for _ in range(2):
        o.time_step(1.0 / 60)
    P, V = o.particles()
    return o, np.ascontiguousarray(P), np.ascontiguousarray(V)


def faces_of(o):
    return [o.get(FIELD[f]).copy() for f in ("U", "V", "W")]


def _fd_check(loss, x0, grad, rng, eps, n_dirs=6, rtol=1e-5):
    for _ in range(n_dirs):
        d = rng.standard_normal(x0.shape)
        fd = (loss(x0 + eps * d) - loss(x0 - eps * d)) / (2 * eps)
        an = float(np.sum(grad * d))
        assert abs(fd - an) <= rtol * max(abs(an), 1e-12), (fd, an)


def test_sample_adjoint():
    o, P, _ = particle_scene()
    L = oracle()
    n = len(P)
    rng = np.random.default_rng(21)
    c = rng.standard_normal((n, 3))
    F = faces_of(o)

    def loss_pos(pos):
        out = np.empty((n, 3))
        L.xo_sample_points(o.h, n, np.ascontiguousarray(pos), out, 0)
        return float(np.sum(c * out))

    gpos = np.zeros((n, 3))
    gF = [np.zeros_like(a) for a in F]
    L.xo_sample_adjoint(o.h, n, P, c, gpos, *gF, 0)
    _fd_check(loss_pos, P, gpos, rng, 1e-8)

    def loss_u(fields):
        for name, a in zip(("U", "V", "W"), fields):
            o.set(FIELD[name], a)
        return loss_pos(P)

    for a in range(3):  # linear in the grid: dot identity per face array
        d = rng.standard_normal(F[a].shape)
        fields = [x.copy() for x in F]
        fields[a] = F[a] + d
        fd = loss_u(fields) - loss_u(F)
        assert abs(fd - float(np.sum(gF[a] * d))) <= 1e-10 * max(abs(fd), 1.0)


@pytest.mark.parametrize("integrator", [0, 1])
def test_advect_adjoint(integrator):
    o, P, _ = particle_scene(integrator=integrator)
    L = oracle()
    n = len(P)
    dt = 1.0 / 60
    rng = np.random.default_rng(22)
    c = rng.standard_normal((n, 3))

    def loss_pos(pos):
        out = np.empty((n, 3))
        L.xo_move_points(o.h, n, np.ascontiguousarray(pos), dt, out)
        return float(np.sum(c * out))

    F = faces_of(o)
    gpos = np.zeros((n, 3))
    gF = [np.zeros_like(a) for a in F]
    L.xo_advect_adjoint(o.h, n, P, dt, c, gpos, *gF)
    _fd_check(loss_pos, P, gpos, rng, 1e-8)

    def loss_f(x):
        fields = np.split(x, np.cumsum([a.size for a in F])[:-1])
        for name, a, ref in zip(("U", "V", "W"), fields, F):
            o.set(FIELD[name], a.reshape(ref.shape))
        return loss_pos(P)

    x0 = np.concatenate([a.ravel() for a in F])
    g0 = np.concatenate([a.ravel() for a in gF])
    _fd_check(loss_f, x0, g0, rng, 1e-7)


@pytest.mark.parametrize("integrator", [0, 1])
def test_advect_adjoint_collisions(integrator):
    """The collision branch of the move (FluidSimulator.cpp:686-722): fast
    particles driven into the walls are pulled back from the first solid voxel
    (X' = p + rn (t - back)) or keep their position; the VJP of the whole move
    against central differences in positions and grid values."""
    o, P, _ = particle_scene(integrator=integrator)
    L = oracle()
    n = len(P)
    dt = 0.05
    F = faces_of(o)
    push = (-1.5, -2.0, 1.0)  # towards the -x, -y and +z walls
    for name, a, extra in zip(("U", "V", "W"), F, push):
        o.set(FIELD[name], 8.0 * a + extra)
    F = faces_of(o)
    target, moved = np.empty((n, 3)), np.empty((n, 3))
    L.xo_advect_points(o.h, n, P, dt, target)
    L.xo_move_points(o.h, n, P, dt, moved)
    pulled = np.any(moved != target, 1) & np.any(moved != P, 1)
    kept = np.all(moved == P, 1)
    assert pulled.sum() > 50 and kept.sum() >= 0
    rng = np.random.default_rng(24)
    c = rng.standard_normal((n, 3))

    def loss_pos(pos):
        out = np.empty((n, 3))
        L.xo_move_points(o.h, n, np.ascontiguousarray(pos), dt, out)
        return float(np.sum(c * out))

    gpos = np.zeros((n, 3))
    gF = [np.zeros_like(a) for a in F]
    L.xo_advect_adjoint(o.h, n, P, dt, c, gpos, *gF)
    # the pulled-back particles' own gradients (a gather of the whole set would hide them)
    cp = np.where(pulled[:, None], c, 0.0)
    gp2 = np.zeros((n, 3))
    L.xo_advect_adjoint(o.h, n, P, dt, cp, gp2, *[np.zeros_like(a) for a in F])

    def loss_pulled(pos):
        out = np.empty((n, 3))
        L.xo_move_points(o.h, n, np.ascontiguousarray(pos), dt, out)
        return float(np.sum(cp * out))

    _fd_check(loss_pulled, P, gp2, rng, 1e-8)
    _fd_check(loss_pos, P, gpos, rng, 1e-8)

    def loss_f(x):
        fields = np.split(x, np.cumsum([a.size for a in F])[:-1])
        for name, a, ref in zip(("U", "V", "W"), fields, F):
            o.set(FIELD[name], a.reshape(ref.shape))
        return loss_pos(P)

    x0 = np.concatenate([a.ravel() for a in F])
    g0 = np.concatenate([a.ravel() for a in gF])
    _fd_check(loss_f, x0, g0, rng, 1e-7)


@pytest.mark.parametrize("kernel", [0, 1])
def test_p2g_adjoint(kernel):
    o, P, V = particle_scene(kernel=kernel)
    L = oracle()
    n = len(P)
    rng = np.random.default_rng(23)
    F = faces_of(o)
    c = [rng.standard_normal(a.shape) for a in F]

    def loss(pos, vel):
        out = [np.empty_like(a) for a in F]
        L.xo_p2g_points(o.h, n, np.ascontiguousarray(pos), np.ascontiguousarray(vel), *out)
        return sum(float(np.sum(a * b)) for a, b in zip(c, out))

    gvel = np.zeros((n, 3))
    gpos = np.zeros((n, 3))
    L.xo_p2g_adjoint(o.h, n, P, V, *c, gvel, gpos)
    _fd_check(lambda v: loss(P, v), V, gvel, rng, 1e-3, rtol=1e-9)   # linear in the velocities
    _fd_check(lambda x: loss(x, V), P, gpos, rng, 1e-8)
    # P2G of the oracle's particles equals the substep's own P2G stage
    out = [np.empty_like(a) for a in F]
    L.xo_p2g_points(o.h, n, P, V, *out)
    o.stage(8)
    for name, a in zip(("U", "V", "W"), out):
        assert np.array_equal(a, o.get(FIELD[name]))


def test_density_adjoint():
    """Particle-blob density VJP: central differences of <c, phi(P)> in the
    positions (the Wyvill polynomial and its derivative vanish at R^2, so phi
    is C1 in P)."""
    o, P, V = particle_scene()
    L = oracle()
    n = len(P)
    rng = np.random.default_rng(24)
    o.stage(0)
    c = rng.standard_normal(o.get(FIELD["PHI"]).shape)

    def loss(pos):
        o.set_particles(np.ascontiguousarray(pos), V)
        o.stage(0)
        return float(np.sum(c * o.get(FIELD["PHI"])))

    gpos = np.zeros((n, 3))
    L.xo_density_adjoint(o.h, n, P, np.ascontiguousarray(c), gpos)
    _fd_check(loss, P, gpos, rng, 1e-8, n_dirs=4)
    # linear in gphi: the dot identity with a second cotangent
    c2 = rng.standard_normal(c.shape)
    g2 = np.zeros((n, 3))
    L.xo_density_adjoint(o.h, n, P, np.ascontiguousarray(c2), g2)
    g12 = np.zeros((n, 3))
    L.xo_density_adjoint(o.h, n, P, np.ascontiguousarray(c + c2), g12)
    assert np.allclose(g12, gpos + g2, rtol=1e-12, atol=1e-12)


def _extrapolated(o, fields):
    for name, a in zip(("U", "V", "W"), fields):
        o.set(FIELD[name], a)
    o.stage(6)  # EXTRAPOLATE
    return faces_of(o)


@pytest.mark.parametrize("cfl", [1.0, 5.0])
def test_extrapolation_adjoint_dot_identity(cfl):
    """The extrapolation is linear for a fixed material: <E x, y> = <x, E^T y>."""
    o, _, _ = particle_scene()
    p = o.params
    p.cfl = cfl
    o.set_params(p)
    o.stage(1)  # material from the particles
    rng = np.random.default_rng(31)
    F = faces_of(o)
    for _ in range(3):
        x = [rng.standard_normal(a.shape) for a in F]
        y = [rng.standard_normal(a.shape) for a in F]
        Ex = _extrapolated(o, [a.copy() for a in x])
        ETy = [np.ascontiguousarray(a.copy()) for a in y]
        oracle().xo_extrapolate_adjoint(o.h, *ETy)
        lhs, rhs = dot(Ex, y), dot(x, ETy)
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0), (lhs, rhs)


def test_extrapolation_adjoint_single_cell():
    """One fluid cell: a layer-1 face's gradient goes to its layer-0 neighbours,
    split by their count; faces beyond the layers keep theirs, faces of layers >= 1 lose theirs."""
    o = oracle_dam_break(7, 7, 7, 0.1, (0.0, 0.0, 0.0), 0, small_params(cfl=1.0))  # 6 layers
    mat = o.get(FIELD["MATERIAL"])
    mat[3, 3, 3] = 1
    o.set(FIELD["MATERIAL"], mat)
    shapes = [a.shape for a in faces_of(o)]
    gu = np.zeros(shapes[0])
    gu[3, 3, 2] = 1.0   # layer 1: its only layer-0 neighbour is u[3, 3, 3] (the cell's -x face)
    gu[3, 4, 3] = 2.0   # layer 1: neighbours u[3, 3, 3] (-j)
    gv, gw = np.zeros(shapes[1]), np.zeros(shapes[2])
    oracle().xo_extrapolate_adjoint(o.h, gu, gv, gw)
    assert gu[3, 3, 2] == 0.0 and gu[3, 4, 3] == 0.0
    assert gu[3, 3, 3] == 3.0 and gu[3, 3, 4] == 0.0
    assert not gv.any() and not gw.any()
