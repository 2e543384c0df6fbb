"""Writes the golden fixtures of tests/golden/ by running the reference's own
code (oracle/_ref/libxfref.so, built from the reference sources by
oracle/Makefile `ref`).  Run in the build container:  python tests/golden/make_golden.py
Only inputs and outputs are stored (npz, no pickles)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import FIELD  # noqa: E402
from scenes import oracle_dam_break, ref_like, small_params  # noqa: E402


def substeps():
    dims, dx, frac, seed, steps, dt = (16, 12, 10), 0.125, (0.35, 0.6, 1.0), 0, 3, 1.0 / 60
    o = oracle_dam_break(*dims, dx, frac, seed, small_params())  # only for the seeded particles + scene
    r = ref_like(o)
    p0, _ = r.particles()
    for _ in range(steps):
        r.time_step(dt)
    P, V = r.particles()
    out = dict(dims=np.array(dims), dx=dx, frac=np.array(frac), seed=seed, steps=steps, dt=dt, p0=p0,
               p_final=P, v_final=V)
    for f in ("U", "V", "W", "MATERIAL", "PHI"):
        out[f] = r.get(FIELD[f])
    np.savez_compressed(os.path.join(HERE, "dam16x12x10_substeps.npz"), **out)


def pressure():
    dims, dx, frac, seed = (14, 11, 9), 0.1, (0.5, 0.6, 1.0), 4
    o = oracle_dam_break(*dims, dx, frac, seed, small_params())
    r = ref_like(o)
    for _ in range(2):
        r.time_step(1.0 / 60)
    r.stage(1, 0.0)
    r.stage(3, 1.0 / 60)
    N = int(np.prod(dims))
    A, b, inv, z = np.empty(7 * N), np.empty(N), np.empty(N), np.empty(N)
    x = np.random.default_rng(0).standard_normal(N)
    assert r.lib.xr_pressure_pieces(r.h, A, b, inv, x, z) == 0
    np.savez_compressed(os.path.join(HERE, "dam14x11x9_pressure.npz"), dims=np.array(dims),
                        material=r.get(FIELD["MATERIAL"]), u=r.get(FIELD["U"]), v=r.get(FIELD["V"]),
                        w=r.get(FIELD["W"]), A7=A, b=b, inv=inv, r=x, z=z)


def mesh():
    """FluidMesher::generateMesh + TriangleMesh::writeMeshToPLY on a dam-break
    density field: phi, the smoothed vertices, triangles and the PLY bytes."""
    import tempfile
    dims, dx, frac, seed = (14, 12, 10), 0.1, (0.5, 0.6, 1.0), 6
    o = oracle_dam_break(*dims, dx, frac, seed, small_params())
    r = ref_like(o)
    for _ in range(3):
        r.time_step(1.0 / 60)
    r.stage(0, 0.0)
    verts, tris = r.mesh(True)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "mesh.ply")
        r.write_ply(path)
        with open(path, "rb") as f:
            ply = np.frombuffer(f.read(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "dam14x12x10_mesh.npz"), dims=np.array(dims), dx=dx,
                        phi=r.get(FIELD["PHI"]), verts=verts, tris=tris, ply=ply)


if __name__ == "__main__":
    substeps()
    pressure()
    mesh()
    print("golden fixtures written")
