"""Closed-form pins of the oracle's switchable forms: the RK2 (midpoint)
advection and the trilinear P2G splat, beside the reference's RK4
(FluidSimulator.cpp:896-907) and normalised Wyvill P2G (FluidSimulator.cpp:
909-1085).

Each check is one a plausible mistake fails:
  * advection of an affine field v(p) = A p + b, whose trilinear samples are
    exact at interior points, so the integrators reduce to their Taylor forms:
    RK2  p + dt f + dt^2/2 A f,  RK4  p + dt f + dt^2/2 A f + dt^3/6 A^2 f +
    dt^4/24 A^3 f  (f = A p + b); the two differ by ~1e-5 here, far above
    rounding, so RK2 computed as RK4 (or Euler) fails;
  * a constant field moves every particle by b dt;
  * a lone particle's P2G weights are the tent products
    (1 - |e_x|/dx)(1 - |e_y|/dx)(1 - |e_z|/dx) on the 8 faces around it
    (trilinear) or W(|e|^2) with R = 1.5 dx (Wyvill), e = face - (p + dx/2 e_axis);
  * uniform particle velocity gives a uniform face field;
  * a regular lattice of particles (spacing dx/2) carrying an affine field
    reproduces it at interior faces (the tent of half-width dx is the linear
    B-spline of two interleaved dx-lattices: partition of unity and exact first
    moment).
"""
import numpy as np
import pytest

from oracle_lib import FIELD, OracleSim, default_params, oracle

RK4, RK2 = 0, 1
WYVILL, TRILINEAR = 0, 1

A = np.array([[0.3, -1.2, 0.5], [0.9, 0.2, -0.7], [-0.4, 0.8, -0.1]])
B = np.array([0.25, -0.5, 0.125])


def _sim(n=16, dx=0.1, **kw):
    return OracleSim(n, n, n, dx, default_params(tail_skip=0, **kw))


def _face_positions(o, comp):
    """(x, y, z) of every face of component comp, arrays in (k, j, i) order."""
    ni, nj, nk, dx = o.ni, o.nj, o.nk, o.dx
    shape = [(nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)][comp]
    off = [(0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0)][comp]
    k, j, i = np.meshgrid(*(np.arange(s) for s in shape), indexing="ij")
    return (i + off[0]) * dx, (j + off[1]) * dx, (k + off[2]) * dx


def _set_affine(o, Am, b):
    for comp, name in enumerate(("U", "V", "W")):
        x, y, z = _face_positions(o, comp)
        o.set(FIELD[name], Am[comp, 0] * x + Am[comp, 1] * y + Am[comp, 2] * z + b[comp])


def _advect(o, P, dt):
    out = np.empty_like(P)
    oracle().xo_advect_points(o.h, len(P), np.ascontiguousarray(P), dt, out)
    return out


def _taylor(P, dt, order):
    f = P @ A.T + B
    out = P + dt * f
    term = f
    for m in range(2, order + 1):
        term = term @ A.T
        out = out + dt ** m / np.prod(np.arange(1, m + 1)) * term
    return out


def test_rk2_and_rk4_affine_field():
    rng = np.random.default_rng(11)
    P = rng.uniform(0.45, 1.15, (300, 3))
    dt = 0.05
    rk2_exact, rk4_exact = _taylor(P, dt, 2), _taylor(P, dt, 4)
    assert np.abs(rk4_exact - rk2_exact).max() > 1e-6  # the check separates the integrators
    for integ, exact in ((RK2, rk2_exact), (RK4, rk4_exact)):
        o = _sim(integrator=integ)
        _set_affine(o, A, B)
        got = _advect(o, P, dt)
        assert np.abs(got - exact).max() <= 1e-14, (integ, np.abs(got - exact).max())


@pytest.mark.parametrize("integ", [RK2, RK4])
def test_constant_field_moves_by_v_dt(integ):
    rng = np.random.default_rng(12)
    P = rng.uniform(0.2, 1.4, (300, 3))
    dt = 0.03
    o = _sim(integrator=integ)
    _set_affine(o, np.zeros((3, 3)), B)
    got = _advect(o, P, dt)
    assert np.abs(got - (P + B * dt)).max() <= 1e-15


def _p2g_weights(o, axis, P, V):
    shape = [(o.nk, o.nj, o.ni + 1), (o.nk, o.nj + 1, o.ni), (o.nk + 1, o.nj, o.ni)][axis]
    out, wt = np.empty(shape), np.empty(shape)
    oracle().xo_p2g_weights(o.h, axis, len(P), np.ascontiguousarray(P), np.ascontiguousarray(V), out, wt)
    return out, wt


@pytest.mark.parametrize("kernel", [TRILINEAR, WYVILL])
def test_lone_particle_weights(kernel):
    dx = 0.1
    o = _sim(n=12, dx=dx, p2g_kernel=kernel)
    P = np.array([[0.437, 0.523, 0.611]])
    V = np.array([[1.0, -2.0, 3.0]])
    for axis in range(3):
        out, wt = _p2g_weights(o, axis, P, V)
        q = P[0] + 0.5 * dx * np.eye(3)[axis]
        # the faces of this axis sit at idx2pos = (idx + 1/2) dx in the shifted frame
        shape = wt.shape
        k, j, i = np.meshgrid(*(np.arange(s) for s in shape), indexing="ij")
        e = [q[0] - (i + 0.5) * dx, q[1] - (j + 0.5) * dx, q[2] - (k + 0.5) * dx]
        if kernel == TRILINEAR:
            t = [np.clip(1.0 - np.abs(x) / dx, 0.0, None) for x in e]
            exp = t[0] * t[1] * t[2]
            assert np.count_nonzero(exp) == 8
            assert abs(wt.sum() - 1.0) <= 1e-14  # partition of unity
        else:
            R2 = (1.5 * dx) ** 2
            s = e[0] ** 2 + e[1] ** 2 + e[2] ** 2
            c0, c1, c2 = 4.0 / 9.0 / R2 ** 3, 17.0 / 9.0 / R2 ** 2, 22.0 / 9.0 / R2
            exp = np.where(s < R2, 1.0 - c0 * s ** 3 + c1 * s ** 2 - c2 * s, 0.0)
            assert np.count_nonzero(exp) > 8  # radius 1.5 dx reaches past the 8 nearest faces
        assert np.abs(wt - exp).max() <= 1e-14, (axis, np.abs(wt - exp).max())
        on = wt > 1e-7
        assert np.abs(out[on] - V[0, axis]).max() <= 1e-15 * abs(V[0, axis]) and np.all(out[~on] == 0.0)


@pytest.mark.parametrize("kernel", [TRILINEAR, WYVILL])
def test_uniform_velocity_gives_uniform_field(kernel):
    rng = np.random.default_rng(5)
    o = _sim(n=12, dx=0.1, p2g_kernel=kernel)
    P = rng.uniform(0.1, 1.1, (3000, 3))
    V = np.tile([0.25, -1.5, 3.0], (len(P), 1))
    for axis in range(3):
        out, wt = _p2g_weights(o, axis, P, V)
        on = wt > 1e-7
        assert on.sum() > 500
        assert np.abs(out[on] - V[0, axis]).max() <= 1e-13 * abs(V[0, axis])


def test_trilinear_lattice_reproduces_affine_field():
    n, dx = 12, 0.1
    h = 0.5 * dx
    o = _sim(n=n, dx=dx, p2g_kernel=TRILINEAR)
    c = (np.arange(2 * n) + 0.5) * h
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    P = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    V = P @ A.T + B
    L = n * dx
    for axis in range(3):
        out, wt = _p2g_weights(o, axis, P, V)
        fx, fy, fz = _face_positions(o, axis)  # MAC face positions = particle-space positions
        exp = A[axis, 0] * fx + A[axis, 1] * fy + A[axis, 2] * fz + B[axis]
        inner = np.ones(out.shape, bool)
        for f in (fx, fy, fz):
            inner &= (f >= dx + h) & (f <= L - dx - h)
        assert inner.sum() > 300
        assert np.abs(wt[inner] - 8.0).max() <= 1e-13  # two interleaved lattices per direction
        assert np.abs(out[inner] - exp[inner]).max() <= 1e-13, (axis, np.abs(out[inner] - exp[inner]).max())
