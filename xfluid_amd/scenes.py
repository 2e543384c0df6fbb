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
