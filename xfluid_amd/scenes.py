"""Synthetic workloads for the projection benchmark.

sphere_scene builds the 256^3 projection workload: a unit cube with a 1-cell
solid shell, liquid = the union of seeded spheres (radius U[0.05, 0.2], centre
U[0.2, 0.8]^3, cell centres inside a sphere are liquid), air elsewhere; face
velocities drawn as float32 U[-1, 1] on faces with a liquid cell on at least
one side, 0 on faces touching solid (or the domain boundary) and elsewhere.
The projection seam it exercises is VariationalPressureSolver::solve +
applyPressure (variational_pressuresolver.cpp:80-220).
"""
import numpy as np

AIR, FLUID, SOLID = 0, 1, 2


def sphere_scene(n=256, nspheres=40, seed=2):
    """Returns (material[nk, nj, ni] uint8, u, v, w) with float64 face arrays
    of shapes (n, n, n+1), (n, n+1, n), (n+1, n, n)."""
    rng = np.random.default_rng(seed)
    radii = rng.uniform(0.05, 0.2, nspheres)
    centres = rng.uniform(0.2, 0.8, (nspheres, 3))
    dx = 1.0 / n
    mat = np.zeros((n, n, n), np.uint8)
    c = (np.arange(n) + 0.5) * dx
    for r, (cx, cy, cz) in zip(radii, centres):
        lo = [max(int((q - r) / dx) - 1, 0) for q in (cx, cy, cz)]
        hi = [min(int((q + r) / dx) + 2, n) for q in (cx, cy, cz)]
        zz = (c[lo[2]:hi[2]] - cz)[:, None, None] ** 2
        yy = (c[lo[1]:hi[1]] - cy)[None, :, None] ** 2
        xx = (c[lo[0]:hi[0]] - cx)[None, None, :] ** 2
        inside = (zz + yy + xx) < r * r
        mat[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]][inside] = FLUID
    mat[0, :, :] = mat[-1, :, :] = SOLID
    mat[:, 0, :] = mat[:, -1, :] = SOLID
    mat[:, :, 0] = mat[:, :, -1] = SOLID
    faces = []
    for axis in range(3):  # u (i), v (j), w (k); array axes are (k, j, i)
        ax = 2 - axis
        pad = [(0, 0)] * 3
        pad[ax] = (1, 1)
        m = np.pad(mat, pad, constant_values=SOLID)  # outside the domain: solid
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[ax] = slice(0, -1)
        sl_hi[ax] = slice(1, None)
        a, b = m[tuple(sl_lo)], m[tuple(sl_hi)]  # the two cells beside each face
        liquid = (a == FLUID) | (b == FLUID)
        solid = (a == SOLID) | (b == SOLID)
        vals = rng.uniform(-1.0, 1.0, a.shape).astype(np.float32).astype(np.float64)
        faces.append(np.where(liquid & ~solid, vals, 0.0))
    return mat, faces[0], faces[1], faces[2]


def _perlin3(x, y, z, perm, grad):
    """Gradient noise (Perlin's improved noise: fade 6t^5 - 15t^4 + 10t^3,
    pseudo-random unit gradients at the integer lattice points, a permutation
    table for hashing), evaluated at the points (x, y, z)."""
    xi, yi, zi = np.floor(x).astype(np.int64), np.floor(y).astype(np.int64), np.floor(z).astype(np.int64)
    xf, yf, zf = x - xi, y - yi, z - zi
    fade = lambda t: t * t * t * (t * (t * 6.0 - 15.0) + 10.0)  # noqa: E731
    u, v, w = fade(xf), fade(yf), fade(zf)
    m = len(perm) - 1

    def corner(dx, dy, dz):
        h = perm[(perm[(perm[(xi + dx) & m] + yi + dy) & m] + zi + dz) & m]
        gx, gy, gz = grad[h, 0], grad[h, 1], grad[h, 2]
        return gx * (xf - dx) + gy * (yf - dy) + gz * (zf - dz)

    lerp = lambda a, b, t: a + t * (b - a)  # noqa: E731
    x00 = lerp(corner(0, 0, 0), corner(1, 0, 0), u)
    x10 = lerp(corner(0, 1, 0), corner(1, 1, 0), u)
    x01 = lerp(corner(0, 0, 1), corner(1, 0, 1), u)
    x11 = lerp(corner(0, 1, 1), corner(1, 1, 1), u)
    return lerp(lerp(x00, x10, v), lerp(x01, x11, v), w)


def curl_noise_faces(ni, nj, nk, dx, seed=3, periods=4.0, noise=0.1):
    """MAC face velocities (u, v, w) = curl of a seeded Perlin vector potential
    psi, scaled to unit RMS, plus N(0, noise) per face, rounded to float32 and
    returned as float64 arrays of shapes (nk, nj, ni+1), (nk, nj+1, ni),
    (nk+1, nj, ni).  psi_x, psi_y, psi_z live on the cell edges along x, y, z,
    so the curl's discrete divergence on every cell is zero up to rounding
    before the noise is added.  `periods` noise periods span the longest side."""
    rng = np.random.default_rng(seed)
    size = 256
    perm = np.tile(rng.permutation(size), 2)
    g = rng.standard_normal((size, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    freq = periods / (max(ni, nj, nk) * dx)
    off = rng.uniform(0, 64, (3, 3))  # a different region of the noise per component

    def field(comp, shape, cx, cy, cz):
        k, j, i = np.meshgrid(*(np.arange(s) for s in shape), indexing="ij")
        x, y, z = (i + cx) * dx * freq, (j + cy) * dx * freq, (k + cz) * dx * freq
        return _perlin3(x + off[comp, 0], y + off[comp, 1], z + off[comp, 2], perm, g)

    px = field(0, (nk + 1, nj + 1, ni), 0.5, 0.0, 0.0)   # x-edges ((i+1/2) dx, j dx, k dx)
    py = field(1, (nk + 1, nj, ni + 1), 0.0, 0.5, 0.0)   # y-edges (i dx, (j+1/2) dx, k dx)
    pz = field(2, (nk, nj + 1, ni + 1), 0.0, 0.0, 0.5)   # z-edges (i dx, j dx, (k+1/2) dx)
    u = (pz[:, 1:, :] - pz[:, :-1, :]) / dx - (py[1:, :, :] - py[:-1, :, :]) / dx
    v = (px[1:, :, :] - px[:-1, :, :]) / dx - (pz[:, :, 1:] - pz[:, :, :-1]) / dx
    w = (py[:, :, 1:] - py[:, :, :-1]) / dx - (px[:, 1:, :] - px[:, :-1, :]) / dx
    rms = np.sqrt((np.sum(u * u) + np.sum(v * v) + np.sum(w * w)) / (u.size + v.size + w.size))
    out = []
    for a in (u, v, w):
        a = a / rms + rng.normal(0.0, noise, a.shape)
        out.append(a.astype(np.float32).astype(np.float64))
    return tuple(out)
