"""Benchmark: substeps/s of the FLIP/PIC substep on one MI355X.

Workload (default): dam break on a 128^3 MAC grid, dx = 1/128, 1-cell solid
walls, fluid x <= 0.4, y <= 0.5 (cell centres) -> 396,900 fluid cells and
3,175,200 particles (8 per cell, one per 2x2x2 sub-cell, SplitMix64 seed 1),
gravity -9.81, FLIP 0.95, fixed dt = 1/240, PCG to 1e-6 relative (inf-norm)
residual.  One step = one substep (density, material, save, forces, pressure
solve, apply, extrapolation, G2P + RK4 advection, P2G).  The timed run uses the
mixed-precision PCG; the plain FP64 PCG runs the same substeps beside it.

Prints one JSON line.  `--gpus N` with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # name: (ni, nj, nk, dx, rd fractions, seed, pcg_max_iter)
    "dam128": (128, 128, 128, 1.0 / 128, (0.4, 0.5, 1.0), 1, 200),
    "dam64x32x32": (64, 32, 32, 0.125, (0.35, 0.6, 1.0), 0, 200),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_sim(xf, cfg, precision, sweep_kernel=0, pressure=0):
    ni, nj, nk, dx, frac, seed, maxit = CONFIGS[cfg]
    p = xf.default_params(flip_ratio=0.95, pcg_tol=1e-6, pcg_max_iter=maxit, precision=precision,
                          sweep_kernel=sweep_kernel, mixed_inner_tol=1e-4, pressure_solver=pressure)
    s = xf.FluidSimulator(ni, nj, nk, p)
    s.setCellWidth(dx)
    s.setSolidBoundary()
    L = (ni * dx, nj * dx, nk * dx)
    s.addCubicFluid((0.0, 0.0, 0.0), (frac[0] * L[0], frac[1] * L[1], frac[2] * L[2]))
    s.initParticles(xf.SEED_SUBCELL, seed)
    s.addConstantForce(0.0, -9.81, 0.0)
    return s


def run_gpu(xf, cfg, precision, steps, warmup, dt, sweep_kernel=0, pressure=0, timing=False):
    """`steps` substeps after `warmup`.  timing=False: the production path, timed
    by the host clock only; timing=True: per-stage and sweep-dispatch events
    (the stage split and the sweep roofline), which add a wait per substep."""
    s = make_sim(xf, cfg, precision, sweep_kernel, pressure)
    for _ in range(warmup):
        s.timeStep(dt)
    s.synchronize()
    s.set_timing(timing)
    stage_ms = [0.0] * 9
    solve_ms, iters, outer, worst = 0.0, 0, 0, 0.0
    sweep_ms, sweeps, fallbacks, box_cells = 0.0, 0, 0, 0
    t0 = time.perf_counter()
    for _ in range(steps):
        rc = s.timeStep(dt)
        st = s.last_solve()
        solve_ms += st.ms
        sweep_ms += st.sweep_ms
        sweeps += st.sweeps
        fallbacks += st.fallback
        box_cells += st.box_cells * st.sweeps // 2  # weighted by the timed preconditioner applications
        iters += st.iterations
        outer += st.outer
        worst = max(worst, st.residual / max(st.rhs_norm, 1e-300))
        if timing:
            t = s.last_timing()
            for i in range(9):
                stage_ms[i] += t.ms[i]
        if rc != 0:
            log(f"warning: substep solve status {rc}")
    s.synchronize()
    elapsed = time.perf_counter() - t0
    n = s.num_particles()
    s.close()
    return dict(elapsed=elapsed, solve_ms=solve_ms, iters=iters, outer=outer, worst_rel_res=worst,
                stage_ms=stage_ms, particles=n, sweep_ms=sweep_ms, sweeps=sweeps, fallbacks=fallbacks,
                box_cells=box_cells / max(sweeps // 2, 1))


def measured_traffic(cfg):
    """(HBM bytes per active sweep launch, mean box cells of the counted run) from
    the committed counter summary (FETCH_SIZE and WRITE_SIZE of the 128^3 run,
    scaled by a plain copy), or (None, None)."""
    path = os.path.join(ROOT, "profiles", "sweep_traffic_dam128.md")
    if cfg != "dam128" or not os.path.exists(path):
        return None, None
    got = {}
    for line in open(path):
        for key in ("sweep_traffic_bytes_per_launch", "sweep_traffic_box_cells"):
            if line.startswith(key + ":"):
                got[key] = int(line.split(":")[1])
    return got.get("sweep_traffic_bytes_per_launch"), got.get("sweep_traffic_box_cells")


def host_cores():
    """CPU threads this process may run on (the box's share, not the machine's)."""
    return len(os.sched_getaffinity(0))


def oracle_run(cfg, dt, warmup):
    """The oracle (FP64 C++ restatement, reference-faithful threading: serial
    assembly, MIC(0), PCG and P2G; 6-chunk OpenMP density and G2P) on the same
    scene: `warmup` untimed substeps, then one timed substep.  Returns the
    baseline record and the oracle (its state after warmup + 1 substeps)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from scenes import oracle_dam_break
    from oracle_lib import default_params
    ni, nj, nk, dx, frac, seed, maxit = CONFIGS[cfg]
    p = default_params(flip_ratio=0.95, pcg_tol=1e-6, pcg_max_iter=maxit)
    o = oracle_dam_break(ni, nj, nk, dx, frac, seed, p)
    for _ in range(warmup):
        o.time_step(dt)
    t0 = time.perf_counter()
    o.time_step(dt)
    el = time.perf_counter() - t0
    st = o.last_solve()
    omp = int(os.environ.get("OMP_NUM_THREADS", host_cores()))
    rec = dict(value=1.0 / el, unit="substeps/s", cores=min(omp, 6), kind="port",
               host_cores=host_cores(), omp_threads=omp, pcg_iterations=int(st.iterations),
               sample=f"substep {warmup + 1} of {cfg} ({o.lib.xo_num_particles(o.h)} particles, the GPU run's "
                      f"scene and state), timed alone after {warmup} untimed substeps; the oracle in "
                      f"reference-faithful mode: serial assembly/MIC(0)/PCG/P2G, 6-chunk OpenMP density and "
                      f"G2P (so at most 6 of the {host_cores()} host threads work); FP64 PCG to 1e-6, "
                      f"{st.iterations} iterations; {el:.2f} s")
    return rec, o


def per_field_error(xf, cfg, o, substeps, dt):
    """Relative max-norm error of the GPU's pressure, u, v, w and particle
    positions against the FP64 oracle after the same `substeps` substeps from
    the same initial state, for the FP64 and the mixed-precision PCG."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import FIELD
    ref = {f: o.get(FIELD[f]) for f in ("PRESSURE", "U", "V", "W")}
    ref["positions"] = o.particles()[0]
    out = {}
    for name, prec in (("fp64", xf.PREC_FP64), ("mixed", xf.PREC_MIXED)):
        s = make_sim(xf, cfg, prec)
        for _ in range(substeps):
            s.timeStep(dt)
        got = {f: s.get_field(FIELD[f]) for f in ("PRESSURE", "U", "V", "W")}
        got["positions"] = s.particles()[0]
        s.close()
        out[name] = {k.lower(): float(np.abs(got[k] - ref[k]).max() / max(np.abs(ref[k]).max(), 1e-300))
                     for k in ref}
    out["substeps"] = substeps
    return out


def copy_calibration(mib):
    """Plain device-to-device copies of a known size: the reference point for
    scaling the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4}), flush=True)


def run_projection(xf, n, reps):
    """The projection benchmark: one pressure solve + apply on the n^3 sphere
    scene (xfluid_amd.scenes.sphere_scene), FP64 PCG and mixed precision, to
    1e-6 and 1e-9 relative residual (max 2000 iterations).  Algorithmic bytes
    per PCG iteration: 131 B/cell FP64, 67 B/cell FP32 bulk (N = n^3 cells)."""
    from xfluid_amd.scenes import sphere_scene
    mat, u, v, w = sphere_scene(n)
    N = n ** 3
    out = []
    for tol in (1e-6, 1e-9):
        for prec in (xf.PREC_FP64, xf.PREC_MIXED):
            p = xf.default_params(pcg_tol=tol, pcg_max_iter=2000, precision=prec)
            s = xf.FluidSimulator(n, n, n, p)
            s.set_timing(True)  # (the solve's device time, st.ms)
            s.setCellWidth(1.0 / n)
            s.set_field(xf.FIELD["MATERIAL"], mat)
            best = None
            for rep in range(reps + 1):  # (the first solve allocates the solver buffers)
                g = [a.copy() for a in (u, v, w)]
                rc = s.project_fields(*g)
                st = s.last_solve()
                log(f"[projection] n={n} tol={tol:g} prec={prec} rep={rep}: rc {rc}, {st.iterations} iterations, "
                    f"{st.ms:.2f} ms")
                if rep > 0 and (best is None or st.ms < best.ms):
                    best = st
            bpc = 131.0 if prec == xf.PREC_FP64 else 67.0
            out.append({"tol": tol, "precision": "fp64" if prec == xf.PREC_FP64 else "mixed",
                        "status": best.status, "iterations": best.iterations, "outer": best.outer,
                        "ms": round(best.ms, 3), "rel_residual": best.residual / max(best.rhs_norm, 1e-300),
                        "achieved_GBs": round(bpc * N * best.iterations / (best.ms * 1e-3) / 1e9, 2)})
            s.close()
    print(json.dumps({"metric": f"PCG solve ms ({n}^3 sphere-scene projection)", "unit": "ms",
                      "higher_is_better": False, "n_gpus": 1, "data": "synthetic",
                      "config": {"workload": f"{n}^3 unit cube, solid shell, liquid = union of 40 seeded spheres",
                                 "grid": [n, n, n], "liquid_cells": int((mat == 1).sum())},
                      "solves": out}), flush=True)


SOURCE_DEMO = {
    # the reference's demo_source (main.cpp:5-26): 64^3, dx 0.125, gravity -20,
    # a cubic source at (4, 30, 30) dx of half-widths (2, 5, 10) dx emitting at
    # (10, 0, 0), no initial fluid, 1/30 s frames (CFL 1, FLIP 0.95, PCG 1e-5)
    "grid": (64, 64, 64), "dx": 0.125, "force": (0.0, -20.0, 0.0),
    "center": (4.0, 30.0, 30.0), "half": (2.0, 5.0, 10.0), "velocity": (10.0, 0.0, 0.0), "frame": 1.0 / 30,
}


def run_source_demo(xf, frames, cpu_frames):
    """The reference's fluid-source demo: frames/s of frameStep(1/30) on the
    GPU (mesh built on each frame's last substep, as the reference does; no
    file output), and the oracle's time for its first `cpu_frames` frames."""
    d = SOURCE_DEMO
    dx = d["dx"]
    c = [x * dx for x in d["center"]]
    w = [x * dx for x in d["half"]]
    s = xf.FluidSimulator(*d["grid"], xf.default_params())
    s.setCellWidth(dx)
    s.addConstantForce(*d["force"])
    s.addFluidSource(c, *w, d["velocity"])
    s.init()
    s.synchronize()
    per_frame = []
    for f in range(frames):
        t0 = time.perf_counter()
        rc = s.frameStep(d["frame"])
        s.synchronize()
        per_frame.append(time.perf_counter() - t0)
        if rc < 0:
            raise SystemExit(f"bench.py: source demo frame {f}: rc {rc}")
    n_end = s.num_particles()
    s.close()
    total = sum(per_frame)
    cpu = None
    if cpu_frames:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import OracleSim, default_params
        o = OracleSim(*d["grid"], dx, default_params())
        o.lib.xo_add_constant_force(o.h, *d["force"])
        o.add_fluid_source(c, *w, d["velocity"])
        o.lib.xo_set_solid_boundary(o.h)
        t0 = time.perf_counter()
        for _ in range(cpu_frames):
            o.frame_step(d["frame"])
        cpu_s = time.perf_counter() - t0
        gpu_s = sum(per_frame[:cpu_frames])
        cpu = {"frames": cpu_frames, "seconds": round(cpu_s, 3), "gpu_seconds_same_frames": round(gpu_s, 4),
               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port"}
    print(json.dumps({"metric": "frames/s (64^3 fluid-source demo, frameStep(1/30) with mesh)",
                      "value": round(frames / total, 3), "unit": "frames/s", "higher_is_better": True,
                      "n_gpus": 1, "data": "synthetic", "dtype": "f64",
                      "config": {"workload": "demo_source: 64^3, dx 0.125, source (4,30,30) dx +- (2,5,10) dx "
                                             "at (10,0,0), gravity -20, no initial fluid",
                                 "frames": frames},
                      "particles_end": n_end, "ms_first_frame": round(per_frame[0] * 1e3, 2),
                      "ms_last_frame": round(per_frame[-1] * 1e3, 2), "cpu_baseline": cpu}), flush=True)


def _face_boxes(ni, nj, nk, dx, lo, hi, torch):
    """0/1 masks of the MAC faces whose positions lie in [lo, hi]^3 (device tensors)."""
    out = []
    for comp, shape in enumerate([(nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)]):
        k, j, i = np.meshgrid(*(np.arange(s_) for s_ in shape), indexing="ij")
        off = [0.5, 0.5, 0.5]
        off[comp] = 0.0
        x, y, z = (i + off[0]) * dx, (j + off[1]) * dx, (k + off[2]) * dx
        m = (x >= lo) & (x <= hi) & (y >= lo) & (y <= hi) & (z >= lo) & (z <= hi)
        out.append(torch.tensor(m.astype(np.float64), device="cuda"))
    return out


def run_diff_projection(xf, n, steps, every, dirs, eps):
    """The differentiable-projection workload: an n^3 unit cube with 1-cell walls,
    liquid below y = 0.7 (8 particles per cell, one per sub-cell, seed 3), initial
    face velocity = curl of a seeded Perlin vector potential (unit RMS) + N(0, 0.1),
    rounded to float32; gravity -9.81; `steps` substeps of dt = 1/240 with FLIP 0
    (the particles carry the grid velocity: advect -> P2G -> project), FP64 PCG to
    1e-12; loss 1/2 |u|^2 over the faces in [0.4, 0.6]^3 after the last substep.
    The gradient with respect to the initial face velocity comes from the adjoint
    kernels with a checkpoint every `every` substeps; it is checked by central
    differences along `dirs` seeded directions (step eps) with each substep's
    material held at the unperturbed run's (the branch the gradient
    differentiates), and along fewer directions with the material the
    particles make, at eps and 1e-6: there a direction whose +eps or -eps run
    changes some substep's material (a cell gaining its first or losing its last
    particle: a jump of the loss) is counted, not compared."""
    import torch
    import xfluid_amd.autograd as ag
    from xfluid_amd.scenes import curl_noise_faces
    dx, dt = 1.0 / n, 1.0 / 240
    p = xf.default_params(flip_ratio=0.0, pcg_tol=1e-12, pcg_max_iter=4000, tail_skip=0, density_every_substep=0)
    s = xf.FluidSimulator(n, n, n, p)
    s.setCellWidth(dx)
    s.setSolidBoundary()
    s.addCubicFluid((0.0, 0.0, 0.0), (1.0, 0.7, 1.0))
    s.initParticles(xf.SEED_SUBCELL, 3)
    s.addConstantForce(0.0, -9.81, 0.0)
    npart = s.num_particles()
    X0 = torch.empty((npart, 3), dtype=torch.float64, device="cuda")
    V0 = torch.empty_like(X0)
    s.get_particles_device(X0, V0)
    U0 = [torch.tensor(a, device="cuda") for a in curl_noise_faces(n, n, n, dx, seed=3)]
    boxes = _face_boxes(n, n, n, dx, 0.4, 0.6, torch)
    log(f"[diff-projection] {n}^3, {npart} particles, {steps} substeps, checkpoint every {every}")

    def loss_of(U, grad=False, mats=None, frozen=None):
        leaves = [t.clone().requires_grad_(grad) for t in U]
        it = iter(frozen) if frozen is not None else None
        step = lambda X, V, u, v, w: ag.substep(s, dt, X, V, u, v, w, material_log=mats,  # noqa: E731
                                                material=next(it) if it is not None else None)
        _, _, u, v, w = ag.rollout(step, (X0, V0, *leaves), steps, every)
        loss = 0.5 * sum((b * t * t).sum() for b, t in zip(boxes, (u, v, w)))
        if grad:
            loss.backward()
            return loss.item(), [t.grad for t in leaves]
        return loss.item()

    loss_of(U0)  # warm-up (first-use allocations)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        mats0 = []
        L0 = loss_of(U0, mats=mats0)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, grad = loss_of(U0, grad=True)
    torch.cuda.synchronize()
    t_fb = time.perf_counter() - t0
    gnorm = float(torch.sqrt(sum((g * g).sum() for g in grad)))
    gen = torch.Generator(device="cuda").manual_seed(3)

    def fd_pass(eps_, ndirs, frozen=False):
        """(relative errors, errors / |grad|) of the directions without material
        change, and the outliers; frozen: every substep keeps the unperturbed
        run's material"""
        rel_, nrm_, outl = [], [], []
        for q in range(ndirs):
            d = [torch.randn(t.shape, generator=gen, dtype=torch.float64, device="cuda") for t in U0]
            mp, mm = [], []
            fz = mats0 if frozen else None
            with torch.no_grad():
                lp = loss_of([a + eps_ * b for a, b in zip(U0, d)], mats=mp, frozen=fz)
                lm = loss_of([a - eps_ * b for a, b in zip(U0, d)], mats=mm, frozen=fz)
            flips = 0 if frozen else sum(int((a != b).sum()) for a, b in zip(mp + mm, mats0 + mats0))
            fd = (lp - lm) / (2 * eps_)
            an = float(sum((g * b).sum() for g, b in zip(grad, d)))
            if flips:
                outl.append({"direction": q, "material_changes": flips, "rel_err": abs(fd - an) / abs(an)})
            else:
                rel_.append(abs(fd - an) / abs(an))
                nrm_.append(abs(fd - an) / gnorm)
            if q % 8 == 7:
                log(f"[diff-projection] eps {eps_:g}: {q + 1} directions, worst {max(rel_) if rel_ else 0:.2e}, "
                    f"{len(outl)} outliers")
        return rel_, nrm_, outl

    errs, norm_errs, _ = fd_pass(eps, dirs, frozen=True)
    live = fd_pass(eps, max(dirs // 8, 1))
    small = fd_pass(1e-6, max(dirs // 2, 1))
    errs_sorted = sorted(errs)
    print(json.dumps({
        "metric": f"differentiable projection {n}^3: adjoint gradient vs central differences",
        "value": max(small[0]) if small[0] else None,
        "unit": "max relative error |fd - <g, d>| / |<g, d>| at eps 1e-6 (directions whose runs keep every "
                "substep's material)",
        "higher_is_better": False, "n_gpus": 1, "data": "synthetic", "dtype": "f64",
        "config": {"workload": f"{n}^3 unit cube, walls, liquid y < 0.7, curl-noise initial velocity (seed 3), "
                               f"gravity, FLIP 0, dt 1/240, FP64 PCG 1e-12",
                   "grid": [n, n, n], "particles": npart, "substeps": steps, "checkpoint_every": every},
        "loss": L0, "grad_norm": gnorm,
        "frozen_material_eps": {"eps": eps, "directions": len(errs),
                                "median_rel_err": errs_sorted[len(errs_sorted) // 2] if errs else None,
                                "max_rel_err": max(errs) if errs else None,
                                "max_err_over_grad_norm": max(norm_errs) if norm_errs else None},
        "material_from_particles_eps_1e-4": {"directions": max(dirs // 8, 1),
                                             "directions_with_material_change": len(live[2]),
                                             "max_rel_err_others": max(live[0]) if live[0] else None},
        "material_from_particles_eps_1e-6": {"directions_compared": len(small[0]),
                                             "max_rel_err": max(small[0]) if small[0] else None,
                                             "median_rel_err": sorted(small[0])[len(small[0]) // 2] if small[0] else None,
                                             "directions_with_material_change": len(small[2])},
        "forward_s": round(t_fwd, 3), "forward_backward_s": round(t_fb, 3),
        "substeps_per_s_forward": round(steps / t_fwd, 2),
        "substeps_per_s_forward_backward": round(steps / t_fb, 2)}), flush=True)
    s.close()


def _trilinear_upsample(coarse, pts, torch):
    """Trilinear interpolation of coarse (C, Z, Y, X) cell-centre values spread over
    the domain box at the points pts = (x, y, z) normalised to [0, 1]; values past
    the outer cell centres are clamped to them."""
    C, Z, Y, X = coarse.shape
    gx = (pts[0] * X - 0.5).clamp(0, X - 1)
    gy = (pts[1] * Y - 0.5).clamp(0, Y - 1)
    gz = (pts[2] * Z - 0.5).clamp(0, Z - 1)
    x0, y0, z0 = gx.floor().long().clamp(max=X - 2), gy.floor().long().clamp(max=Y - 2), gz.floor().long().clamp(max=Z - 2)
    fx, fy, fz = gx - x0, gy - y0, gz - z0
    out = 0.0
    for dz in (0, 1):
        for dy in (0, 1):
            for dx_ in (0, 1):
                wgt = (fx if dx_ else 1 - fx) * (fy if dy else 1 - fy) * (fz if dz else 1 - fz)
                out = out + wgt * coarse[:, z0 + dz, y0 + dy, x0 + dx_]
    return out


def run_force_control(xf, iters, steps, every, lr):
    """Gradient-based body-force control: a 128 x 64 x 64 tank (dx 1/64, walls),
    fluid y < 0.3 with 8 particles per cell (seed 4); a learnable float32 force
    on a 16 x 8 x 8 coarse grid (3 components) is trilinearly upsampled to the
    MAC faces and added each substep (on the faces next to fluid, beside gravity
    -9.81); `steps` substeps of dt 1/240, FLIP 0.95, mixed-precision PCG.  Loss:
    squared L2 distance between the final particle-blob density field and a
    target = the initial fluid's mean interior density below a Gaussian mound
    y_t(x) = 0.3 + 0.5 exp(-(x - 0.75)^2 / (2 * 0.15^2)), zero above.  Backward
    through P2G, advection, G2P, extrapolation, projection and the density field,
    checkpointed every `every` substeps; `iters` Adam steps."""
    import torch
    import xfluid_amd.autograd as ag
    ni, nj, nk = 128, 64, 64
    dx, dt = 1.0 / 64, 1.0 / 240
    p = xf.default_params(flip_ratio=0.95, pcg_tol=1e-6, pcg_max_iter=400, precision=xf.PREC_MIXED, tail_skip=0,
                          density_every_substep=0, mixed_inner_tol=1e-4)
    s = xf.FluidSimulator(ni, nj, nk, p)
    s.setCellWidth(dx)
    s.setSolidBoundary()
    s.addCubicFluid((0.0, 0.0, 0.0), (ni * dx, 0.3, nk * dx))
    s.initParticles(xf.SEED_SUBCELL, 4)
    s.addConstantForce(0.0, -9.81, 0.0)
    npart = s.num_particles()
    X0 = torch.empty((npart, 3), dtype=torch.float64, device="cuda")
    V0 = torch.empty_like(X0)
    s.get_particles_device(X0, V0)
    U0 = [torch.zeros(sh, dtype=torch.float64, device="cuda") for sh in ((nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni))]
    Lx, Ly, Lz = ni * dx, nj * dx, nk * dx
    face_pts = []
    for comp, shape in enumerate([(nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)]):
        k, j, i = torch.meshgrid(*(torch.arange(s_, device="cuda", dtype=torch.float64) for s_ in shape), indexing="ij")
        off = [0.5, 0.5, 0.5]
        off[comp] = 0.0
        face_pts.append(((i + off[0]) * dx / Lx, (j + off[1]) * dx / Ly, (k + off[2]) * dx / Lz))
    coarse = torch.zeros((3, 8, 8, 16), dtype=torch.float32, device="cuda", requires_grad=True)
    opt = torch.optim.Adam([coarse], lr=lr)
    # target density
    with torch.no_grad():
        phi0 = ag.density(s, X0)
        d = p.sub_division
        kk, jj, ii = torch.meshgrid(*(torch.arange(m, device="cuda", dtype=torch.float64) for m in phi0.shape),
                                    indexing="ij")
        sub = dx / d
        xv, yv = (ii + 0.5) * sub, (jj + 0.5) * sub
        inner = (yv < 0.2) & (yv > 0.05) & (xv > 0.1) & (xv < Lx - 0.1)
        level = float(phi0[inner].mean())
        yt = 0.3 + 0.5 * torch.exp(-((xv - 0.75) ** 2) / (2 * 0.15 ** 2))
        target = torch.where(yv < yt, level, 0.0)
    free0, total = torch.cuda.mem_get_info()
    peak = [total - free0]

    def probe():
        free, _ = torch.cuda.mem_get_info()
        peak[0] = max(peak[0], total - free)

    def step(X, V, u, v, w):
        out = ag.substep(s, dt, X, V, u, v, w, force)
        probe()
        return out

    log(f"[force-control] {ni}x{nj}x{nk}, {npart} particles, {steps} substeps, {iters} Adam steps")
    losses, times = [], []
    for it in range(iters):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        opt.zero_grad()
        force = tuple(_trilinear_upsample(coarse.double(), pts, torch)[comp] for comp, pts in enumerate(face_pts))
        X, V, u, v, w = ag.rollout(step, (X0, V0, *U0), steps, every)
        phi = ag.density(s, X)
        loss = ((phi - target) ** 2).sum()
        loss.backward()
        probe()
        opt.step()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        losses.append(loss.item())
        if it % 10 == 0 or it == iters - 1:
            log(f"[force-control] step {it}: loss {losses[-1]:.6e}, {times[-1]:.2f} s")
    t_iter = sorted(times[1:])[len(times[1:]) // 2] if len(times) > 1 else times[0]
    print(json.dumps({
        "metric": "body-force control 128x64x64: forward+backward substeps/s",
        "value": round(steps / t_iter, 3), "unit": "substeps/s (forward + backward, checkpointed)",
        "higher_is_better": True, "n_gpus": 1, "data": "synthetic",
        "dtype": "f64 (PCG bulk f32 + f64 dots/refinement; force parameters f32)",
        "config": {"workload": "128x64x64 tank, dx 1/64, fluid y < 0.3 (seed 4), 16x8x8 learnable force, "
                               "gravity, FLIP 0.95, dt 1/240, mixed PCG 1e-6",
                   "particles": npart, "substeps": steps, "checkpoint_every": every, "adam_steps": iters, "lr": lr},
        "loss_first": losses[0], "loss_last": losses[-1], "loss_min": min(losses),
        "median_iteration_s": round(t_iter, 3), "peak_hbm_bytes": int(peak[0]),
        "peak_hbm_GiB": round(peak[0] / 2 ** 30, 3)}), flush=True)
    s.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="dam128", choices=list(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--sweep", type=int, default=0)
    ap.add_argument("--pressure", default="variational", choices=["variational", "normal"],
                    help="pressure operator: the reference's VariationalPressureSolver (default) or "
                         "NormalPressureSolver")
    ap.add_argument("--projection", type=int, default=0, metavar="N",
                    help="only run the N^3 sphere-scene projection benchmark (N = 256 for the full size) and exit")
    ap.add_argument("--source-demo", type=int, default=0, metavar="FRAMES",
                    help="only run the reference's fluid-source demo for FRAMES frames and exit")
    ap.add_argument("--diff-projection", type=int, default=0, metavar="N",
                    help="only run the differentiable-projection workload on an N^3 grid (N = 96 for the full size)")
    ap.add_argument("--directions", type=int, default=64, help="finite-difference directions (--diff-projection)")
    ap.add_argument("--force-control", type=int, default=0, metavar="ITERS",
                    help="only run the body-force control workload for ITERS Adam steps (50 for the full run)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calib:
        copy_calibration(args.copy_calib)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the substep runs on one GPU; multi-GPU runs are not part of this library "
              "(the path would shard as z-slabs, with the MIC(0) sweeps pipelined across slabs).",
              file=sys.stderr)
        sys.exit(2)
    import xfluid_amd as xf
    if xf.device_count() < 1:
        raise SystemExit("bench.py: no HIP device")
    if args.projection:
        run_projection(xf, args.projection, 2)
        return
    if args.diff_projection:
        run_diff_projection(xf, args.diff_projection, 16, 4, args.directions, 1e-4)
        return
    if args.force_control:
        run_force_control(xf, args.force_control, 40, 4, 1.0)
        return
    if args.source_demo:
        run_source_demo(xf, args.source_demo, 0 if args.no_cpu else 2)
        return
    dt = 1.0 / 240
    ni, nj, nk = CONFIGS[args.config][:3]
    N = ni * nj * nk
    log(f"[bench] mixed precision, {args.config}, warmup {args.warmup}, steps {args.steps}")
    op = xf.PRESSURE_NORMAL if args.pressure == "normal" else xf.PRESSURE_VARIATIONAL
    mixed = run_gpu(xf, args.config, xf.PREC_MIXED, args.steps, args.warmup, dt, args.sweep, op)
    log(f"[bench] mixed: {mixed['elapsed']:.3f} s, iters {mixed['iters']}")
    # the same substeps again with events (solve times, the stage split and the
    # sweep roofline: the production pass records none)
    mixed_t = run_gpu(xf, args.config, xf.PREC_MIXED, args.steps, args.warmup, dt, args.sweep, op, timing=True)
    log(f"[bench] mixed (events): {mixed_t['elapsed']:.3f} s, solve {mixed_t['solve_ms']:.1f} ms")
    fp64 = fp64_t = None
    if not args.no_fp64:
        log("[bench] fp64 PCG")
        fp64 = run_gpu(xf, args.config, xf.PREC_FP64, args.steps, args.warmup, dt, args.sweep, op)
        fp64_t = run_gpu(xf, args.config, xf.PREC_FP64, args.steps, args.warmup, dt, args.sweep, op, timing=True)
        log(f"[bench] fp64: {fp64['elapsed']:.3f} s, solve (events) {fp64_t['solve_ms']:.1f} ms, "
            f"iters {fp64['iters']}")
    K = args.steps
    value = K / mixed["elapsed"]
    # Roofline of the dominant kernel, the pipelined MIC(0) sweep: one launch of
    # k_sweep_w8_fb = the forward and the backward sweep of one preconditioner
    # application.  Algorithmic bytes per cell and sweep = read input + pivot +
    # 2-byte stencil code, write output: FP32 (mixed run) 4+4+2+4 = 14 B, FP64
    # 26 B; 28 B per cell and launch in the mixed run.  The cells are those of
    # the active rows' bounding box (the library's box_cells: the only cells the
    # sweep and the PCG vector kernels visit); the full-grid figure (N cells)
    # is reported beside it.  Average launch duration from device events on the
    # library's stream around the launches of every 5th application.
    apps = mixed_t["sweeps"] // 2
    sweep_us = 1e3 * mixed_t["sweep_ms"] / max(apps, 1)
    box = mixed_t["box_cells"]
    ach_sweep = 28.0 * box / (sweep_us * 1e-6) / 1e9 if apps else 0.0
    ach_sweep_full = 28.0 * N / (sweep_us * 1e-6) / 1e9 if apps else 0.0
    # Counted bytes scale with the active box (tiles, slots and padding all do):
    # the counter run's bytes per box cell times this run's mean box.
    traffic_counted, traffic_box = measured_traffic(args.config) if args.sweep in (0, 9) else (None, None)
    traffic = int(traffic_counted * box / traffic_box) if traffic_counted and traffic_box and box else None
    # whole-solve rate (algorithmic bytes per PCG iteration: FP64 131 B/cell,
    # FP32 bulk 67 B/cell over the active box)
    it_m = max(mixed["iters"], 1)
    ach_m = 67.0 * box * it_m / (mixed_t["solve_ms"] * 1e-3) / 1e9
    out = {
        "metric": "substeps/s (128^3 dam break, mixed-precision PCG to 1e-6 rel. residual)",
        "value": round(value, 4),
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mixed["elapsed"] / K, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (PCG bulk f32 + f64 dots/refinement)",
        "data": "synthetic",
        "config": {"workload": f"{args.config} dam break, {mixed['particles']} particles, dt=1/240, FLIP 0.95",
                   "grid": [ni, nj, nk], "particles": mixed["particles"], "pressure_solver": args.pressure},
        "pcg_ms_per_solve_mixed": round(mixed_t["solve_ms"] / K, 3),
        "pcg_iters_per_solve_mixed": round(mixed["iters"] / K, 2),
        "pcg_outer_per_solve_mixed": round(mixed["outer"] / K, 2),
        "worst_rel_residual_mixed": mixed["worst_rel_res"],
        "particles_per_s": round(mixed["particles"] * value, 1),
        "stage_ms_per_step_mixed": [round(x / K, 3) for x in mixed_t["stage_ms"]],
        "ms_per_step_with_events": round(1e3 * mixed_t["elapsed"] / K, 3),
        "roofline": {"bound": "hbm",
                     "kernel": "k_sweep_w8_fb (forward + backward MIC(0) sweep, FP32 in the mixed solve)",
                     "achieved": round(ach_sweep, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach_sweep / HBM_PEAK_GBS, 5),
                     "traffic": traffic, "traffic_counted": traffic_counted, "traffic_counted_box_cells": traffic_box,
                     "bytes_per_launch": int(28 * box), "box_cells": int(box), "avg_launch_us": round(sweep_us, 2),
                     "launches_timed": apps,
                     "achieved_full_grid": round(ach_sweep_full, 2), "bytes_per_launch_full_grid": 28 * N},
        "sweep_fallbacks_mixed": mixed["fallbacks"],
        "pcg_mixed_achieved_GBs": round(ach_m, 2),
    }
    if fp64:
        it_f = max(fp64["iters"], 1)
        out["substeps_per_s_fp64"] = round(K / fp64["elapsed"], 4)
        out["pcg_ms_per_solve_fp64"] = round(fp64_t["solve_ms"] / K, 3)
        out["pcg_iters_per_solve_fp64"] = round(fp64["iters"] / K, 2)
        out["mixed_vs_fp64_solve_rate"] = round(fp64_t["solve_ms"] / max(mixed_t["solve_ms"], 1e-9), 3)
        out["pcg_fp64_achieved_GBs"] = round(131.0 * fp64_t["box_cells"] * it_f / (fp64_t["solve_ms"] * 1e-3) / 1e9,
                                             2)
        out["sweep_fallbacks_fp64"] = fp64["fallbacks"]
        if fp64_t["sweeps"]:
            out["sweep_fp64_avg_launch_us"] = round(1e3 * fp64_t["sweep_ms"] / (fp64_t["sweeps"] // 2), 2)
    if not args.no_cpu:
        log(f"[bench] cpu baseline (oracle: {args.warmup} untimed substeps + 1 timed)")
        out["cpu_baseline"], o = oracle_run(args.config, dt, args.warmup)
        log("[bench] per-field error of the GPU runs against the oracle after the same substeps")
        out["per_field_error_vs_fp64_oracle"] = per_field_error(xf, args.config, o, args.warmup + 1, dt)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
