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


SOURCE_SCENE = dict(
    dims=(16, 14, 12), dx=0.125, dt=0.0125, seed=2024,
    centers=((0.5, 1.0, 1.0), (0.0625, 0.125, 0.0), (1.9, 0.75, 0.625)),
    halves=((0.25, 0.375, 0.5), (0.125, 0.3125, 0.0625), (0.125, 0.125, 0.125)),
    vels=((10.0, 0.0, 0.0), (3.0, 4.0, 0.0), (0.0, -2.0, 1.0)))


def sources():
    """CubicFluidSource::update (fluidsource.cpp:75-92) of the reference after
    srand(seed), walls first: the raw rand() values it consumed, the order in
    which this build drew a particle's x, y, z, and the particles it appended."""
    from oracle_lib import RefSim, default_params, ref_rand
    sc = SOURCE_SCENE
    r = RefSim(*sc["dims"], sc["dx"], default_params())
    r.lib.xr_set_solid_boundary(r.h)
    for c, w, v in zip(sc["centers"], sc["halves"], sc["vels"]):
        r.add_fluid_source(c, *w, v)
    r.apply_sources(sc["dt"], sc["seed"])
    P, V = r.particles()
    raw, rmax = ref_rand(sc["seed"], 3 * len(P))
    assert rmax == 2147483647
    # the argument order of Vec3d(randomDouble x, y, z): the one whose jitter,
    # randomDouble = lo + r / (RAND_MAX / (hi - lo)) (xfluidutil.cpp:67-70),
    # reproduces the positions' offsets from their cell centres
    dx = sc["dx"]
    lo, hi = -0.4999 * dx, 0.4999 * dx
    jit = lo + raw.reshape(-1, 3).astype(np.float64) / (2147483647.0 / (hi - lo))
    centre = (np.floor(P / dx) + 0.5) * dx
    order = None
    for cand in ((0, 1, 2), (2, 1, 0)):
        if np.array_equal(centre + jit[:, cand], P):
            order = cand
    assert order is not None
    np.savez_compressed(os.path.join(HERE, "sources16x14x12.npz"), dims=np.array(sc["dims"]), dx=dx,
                        dt=sc["dt"], seed=sc["seed"], centers=np.array(sc["centers"]),
                        halves=np.array(sc["halves"]), vels=np.array(sc["vels"]), raw=raw,
                        order=np.array(order), P=P, V=V, material=r.get(FIELD["MATERIAL"]))


def normal():
    """NormalPressureSolver (pressuresolver.cpp:50-347) on a dam-break state:
    inputs, matrix, RHS, the pressure of its solve and the velocities after its
    applyPressure."""
    dims, dx, frac, seed = (14, 11, 9), 0.1, (0.5, 0.6, 1.0), 4
    p = small_params(pressure_solver=1)
    o = oracle_dam_break(*dims, dx, frac, seed, p)  # the scene's state (inputs only)
    for _ in range(2):
        o.time_step(1.0 / 60)
    o.stage(1, 0.0)
    o.stage(3, 1.0 / 60)
    r = ref_like(o)
    for f in ("U", "V", "W", "MATERIAL"):
        r.set(FIELD[f], o.get(FIELD[f]))
    inputs = {f.lower(): o.get(FIELD[f]) for f in ("U", "V", "W")}
    N = int(np.prod(dims))
    A, b, pr = np.empty(7 * N), np.empty(N), np.empty(N)
    assert r.lib.xr_normal_pieces(r.h, A, b, pr) == 0
    np.savez_compressed(os.path.join(HERE, "normal14x11x9.npz"), dims=np.array(dims),
                        material=o.get(FIELD["MATERIAL"]), u=inputs["u"], v=inputs["v"], w=inputs["w"],
                        A7=A, b=b, pressure=pr, u_out=r.get(FIELD["U"]), v_out=r.get(FIELD["V"]),
                        w_out=r.get(FIELD["W"]))


if __name__ == "__main__":
    todo = sys.argv[1:] or ["substeps", "pressure", "mesh", "sources", "normal"]
    for name in todo:
        globals()[name]()
    print("golden fixtures written:", ", ".join(todo))
