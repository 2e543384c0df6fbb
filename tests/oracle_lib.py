"""ctypes bindings of the oracle (oracle/liboracle.so) and of the reference
harness (oracle/_ref/libxfref.so).  Test infrastructure: the product path never
imports this module.

Both expose the same handle shape as the HIP library so a test can drive the
reference, the oracle and the GPU through one scenario object (`Scene`).
"""
import atexit
import ctypes as C
import os
import shutil
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libxfref.so")

FIELD = dict(U=0, V=1, W=2, USAVED=3, VSAVED=4, WSAVED=5, MATERIAL=6, PHI=7, PRESSURE=8, RHS=9, MIC_INV=10)
STAGE = dict(DENSITY=0, MATERIAL=1, SAVE=2, FORCES=3, SOLVE=4, APPLY=5, EXTRAPOLATE=6, ADVECT=7, P2G=8)

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class XfParams(C.Structure):
    """Mirror of xf_params in include/xfluid.h."""
    _fields_ = [
        ("flip_ratio", C.c_double), ("cfl", C.c_double), ("max_dt", C.c_double), ("pcg_tol", C.c_double),
        ("pcg_max_iter", C.c_int), ("sub_division", C.c_int), ("mic_tau", C.c_double), ("mic_sigma", C.c_double),
        ("blob_radius", C.c_double), ("iso", C.c_double), ("precision", C.c_int), ("integrator", C.c_int),
        ("p2g_kernel", C.c_int), ("tail_skip", C.c_int), ("density_every_substep", C.c_int),
        ("sweep_kernel", C.c_int), ("mixed_inner_tol", C.c_double), ("mixed_max_outer", C.c_int),
        ("pressure_solver", C.c_int),
        ("reserved", C.c_int * 4),
    ]


class XfSolveStats(C.Structure):
    # layout of xf_solve_stats (include/xfluid.h)
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("outer", C.c_int), ("fallback", C.c_int),
                ("residual", C.c_double), ("rhs_norm", C.c_double), ("ms", C.c_double),
                ("sweep_ms", C.c_double), ("sweeps", C.c_int), ("sweep_fused", C.c_int),
                ("box_cells", C.c_int64)]


def _load_oracle():
    lib = C.CDLL(ORACLE_SO)
    lib.xf_params_default.argtypes = [C.POINTER(XfParams)]
    lib.xo_create.restype = C.c_void_p
    lib.xo_create.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.xo_destroy.argtypes = [C.c_void_p]
    lib.xo_set_cell_width.argtypes = [C.c_void_p, C.c_double]
    lib.xo_set_params.argtypes = [C.c_void_p, C.POINTER(XfParams)]
    lib.xo_set_solid_boundary.argtypes = [C.c_void_p]
    lib.xo_add_sphere_fluid.argtypes = [C.c_void_p] + [C.c_double] * 4
    lib.xo_add_cubic_fluid.argtypes = [C.c_void_p, _dp, _dp]
    lib.xo_add_constant_force.argtypes = [C.c_void_p] + [C.c_double] * 3
    lib.xo_add_fluid_source.argtypes = [C.c_void_p, _dp] + [C.c_double] * 3 + [_dp]
    lib.xo_set_source_seed.argtypes = [C.c_void_p, C.c_uint64]
    lib.xo_apply_sources.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
    lib.xo_seed_particles.argtypes = [C.c_void_p, C.c_int, C.c_uint64]
    lib.xo_seed_particles_draws.argtypes = [C.c_void_p, C.c_void_p]
    lib.xo_num_particles.restype = C.c_int64
    lib.xo_num_particles.argtypes = [C.c_void_p]
    lib.xo_set_particles.argtypes = [C.c_void_p, C.c_int64] + [_dp] * 6
    lib.xo_get_particles.argtypes = [C.c_void_p] + [_dp] * 6
    lib.xo_set_field.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
    lib.xo_get_field.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
    lib.xo_stage.argtypes = [C.c_void_p, C.c_int, C.c_double]
    lib.xo_time_step.argtypes = [C.c_void_p, C.c_double, C.c_int]
    lib.xo_frame_step.argtypes = [C.c_void_p, C.c_double]
    lib.xo_last_solve.argtypes = [C.c_void_p, C.POINTER(XfSolveStats)]
    lib.xo_surface_cells.restype = C.c_int64
    lib.xo_surface_cells.argtypes = [C.c_void_p, _i32, _u8, C.c_int64]
    lib.xo_mesh.restype = C.c_int64
    lib.xo_mesh.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64)]
    lib.xo_get_mesh.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.xo_particle_radius.restype = C.c_double
    lib.xo_particle_radius.argtypes = [C.c_void_p]
    g = [C.c_int] * 3
    lib.xo_assemble.argtypes = g + [_u8, _dp, _dp, _dp, _dp, _dp]
    lib.xo_spmv.argtypes = g + [_dp, _dp, _dp]
    lib.xo_mic0.argtypes = g + [_dp, C.c_double, C.c_double, _dp]
    lib.xo_sweep_fwd.argtypes = g + [_dp, _dp, _dp, _dp]
    lib.xo_sweep_bwd.argtypes = g + [_dp, _dp, _dp, _dp]
    lib.xo_pcg.argtypes = g + [_dp, _dp, _dp, C.c_double, C.c_int, C.c_double, C.c_double,
                               C.POINTER(C.c_int), C.POINTER(C.c_double)]
    lib.xo_apply_pressure.argtypes = g + [_u8, _dp, _dp, _dp, _dp]
    lib.xo_assemble_normal.argtypes = g + [_u8, _dp, _dp, _dp, _dp, _dp]
    lib.xo_apply_pressure_normal.argtypes = g + [_u8, _dp, _dp, _dp, _dp]
    P = C.c_void_p
    lib.xo_sample_points.argtypes = [P, C.c_int64, _dp, _dp, C.c_int]
    lib.xo_sample_adjoint.argtypes = [P, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int]
    lib.xo_advect_points.argtypes = [P, C.c_int64, _dp, C.c_double, _dp]
    lib.xo_move_points.argtypes = [P, C.c_int64, _dp, C.c_double, _dp]
    lib.xo_extrapolate_adjoint.argtypes = [P, _dp, _dp, _dp]
    lib.xo_advect_adjoint.argtypes = [P, C.c_int64, _dp, C.c_double, _dp, _dp, _dp, _dp, _dp]
    lib.xo_p2g_points.argtypes = [P, C.c_int64, _dp, _dp, _dp, _dp, _dp]
    lib.xo_p2g_adjoint.argtypes = [P, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.xo_p2g_weights.argtypes = [P, C.c_int, C.c_int64, _dp, _dp, _dp, _dp]
    lib.xo_density_adjoint.argtypes = [P, C.c_int64, _dp, _dp, _dp]
    lib.xo_project.argtypes = g + [_u8, _dp, _dp, _dp, C.c_double, C.c_int, C.c_double, C.c_double]
    lib.xo_project_adjoint.argtypes = g + [_u8, _dp, _dp, _dp, C.c_double, C.c_int, C.c_double, C.c_double]
    return lib


def _load_ref():
    lib = C.CDLL(REF_SO)
    lib.xr_create.restype = C.c_void_p
    lib.xr_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
    lib.xr_destroy.argtypes = [C.c_void_p]
    lib.xr_set_params.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_double, C.c_double]
    lib.xr_set_solid_boundary.argtypes = [C.c_void_p]
    lib.xr_add_cubic_fluid.argtypes = [C.c_void_p, _dp, _dp]
    lib.xr_add_sphere_fluid.argtypes = [C.c_void_p] + [C.c_double] * 4
    lib.xr_add_constant_force.argtypes = [C.c_void_p] + [C.c_double] * 3
    lib.xr_add_fluid_source.argtypes = [C.c_void_p, _dp] + [C.c_double] * 3 + [_dp]
    lib.xr_apply_sources.argtypes = [C.c_void_p, C.c_double, C.c_uint]
    lib.xr_rand.argtypes = [C.c_uint, C.c_int64, C.c_void_p]
    lib.xr_init_particles.argtypes = [C.c_void_p, C.c_uint]
    lib.xr_init.argtypes = [C.c_void_p]
    lib.xr_set_particles.argtypes = [C.c_void_p, C.c_int64] + [_dp] * 6
    lib.xr_num_particles.restype = C.c_int64
    lib.xr_num_particles.argtypes = [C.c_void_p]
    lib.xr_get_particles.argtypes = [C.c_void_p] + [_dp] * 6
    lib.xr_get_field.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    lib.xr_set_field.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    lib.xr_stage.argtypes = [C.c_void_p, C.c_int, C.c_double]
    lib.xr_last_solve.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int)]
    lib.xr_time_step.argtypes = [C.c_void_p, C.c_double]
    lib.xr_frame_step.argtypes = [C.c_void_p, C.c_double]
    lib.xr_surface_cells.restype = C.c_int64
    lib.xr_surface_cells.argtypes = [C.c_void_p, _i32, C.c_int64]
    lib.xr_pressure_pieces.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
    lib.xr_normal_pieces.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.xr_mesh.restype = C.c_int64
    lib.xr_mesh.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64)]
    lib.xr_get_mesh.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.xr_write_ply.argtypes = [C.c_void_p, C.c_char_p]
    lib.xr_set_output.argtypes = [C.c_char_p]
    # the reference writes each frame's mesh to <Config::outputFild><frame>.ply:
    # send those files to a scratch directory removed at exit
    out = tempfile.mkdtemp(prefix="xfref_frames_")
    atexit.register(shutil.rmtree, out, True)
    lib.xr_set_output(os.fsencode(out + os.sep))
    return lib


_ORACLE = None
_REF = None


def oracle():
    global _ORACLE
    if _ORACLE is None:
        _ORACLE = _load_oracle()
    return _ORACLE


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _REF
    if _REF is None:
        _REF = _load_ref()
    return _REF


def default_params(**kw):
    p = XfParams()
    oracle().xf_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def field_shape(ni, nj, nk, which, d=2):
    """numpy shape (k, j, i) and dtype of a field, i fastest as in array3d.h:143-146."""
    if which == FIELD["U"] or which == FIELD["USAVED"]:
        return (nk, nj, ni + 1), np.float64
    if which == FIELD["V"] or which == FIELD["VSAVED"]:
        return (nk, nj + 1, ni), np.float64
    if which == FIELD["W"] or which == FIELD["WSAVED"]:
        return (nk + 1, nj, ni), np.float64
    if which == FIELD["MATERIAL"]:
        return (nk, nj, ni), np.uint8
    if which == FIELD["PHI"]:
        return (d * nk + 1, d * nj + 1, d * ni + 1), np.float64
    return (nk, nj, ni), np.float64


class OracleSim:
    """The FP64 restatement, one handle."""

    def __init__(self, ni, nj, nk, dx, params=None):
        self.lib = oracle()
        self.ni, self.nj, self.nk, self.dx = ni, nj, nk, dx
        self.h = self.lib.xo_create(ni, nj, nk)
        self.lib.xo_set_cell_width(self.h, dx)
        self.params = params or default_params()
        self.lib.xo_set_params(self.h, C.byref(self.params))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.xo_destroy(self.h)
            self.h = None

    def set_params(self, p):
        self.params = p
        self.lib.xo_set_params(self.h, C.byref(p))

    def get(self, which):
        shape, dt = field_shape(self.ni, self.nj, self.nk, which, self.params.sub_division)
        a = np.empty(shape, dt)
        assert self.lib.xo_get_field(self.h, which, a.ctypes.data, a.nbytes) == 0
        return a

    def set(self, which, a):
        shape, dt = field_shape(self.ni, self.nj, self.nk, which, self.params.sub_division)
        a = np.ascontiguousarray(a, dtype=dt).reshape(shape)
        assert self.lib.xo_set_field(self.h, which, a.ctypes.data, a.nbytes) == 0

    def set_particles(self, P, Vv):
        P = np.ascontiguousarray(P, np.float64)
        Vv = np.ascontiguousarray(Vv, np.float64)
        cols = [np.ascontiguousarray(P[:, a]) for a in range(3)] + [np.ascontiguousarray(Vv[:, a]) for a in range(3)]
        self.lib.xo_set_particles(self.h, len(P), *cols)

    def particles(self):
        n = self.lib.xo_num_particles(self.h)
        cols = [np.empty(n) for _ in range(6)]
        self.lib.xo_get_particles(self.h, *cols)
        return np.stack(cols[:3], 1), np.stack(cols[3:], 1)

    def stage(self, s, dt=0.0):
        return self.lib.xo_stage(self.h, s, dt)

    def add_fluid_source(self, c, wx, wy, wz, v):
        self.lib.xo_add_fluid_source(self.h, np.asarray(c, np.float64), wx, wy, wz, np.asarray(v, np.float64))

    def set_source_seed(self, seed):
        self.lib.xo_set_source_seed(self.h, seed)

    def apply_sources(self, dt, draws=None):
        """One _applyFluidSource; `draws` (uint32, 3 per new particle) replaces the stream."""
        if draws is None:
            return self.lib.xo_apply_sources(self.h, dt, None)
        d = np.ascontiguousarray(draws, np.uint32)
        return self.lib.xo_apply_sources(self.h, dt, d.ctypes.data)

    def time_step(self, dt, mesh=False):
        return self.lib.xo_time_step(self.h, dt, int(mesh))

    def frame_step(self, dt):
        return self.lib.xo_frame_step(self.h, dt)

    def last_solve(self):
        st = XfSolveStats()
        self.lib.xo_last_solve(self.h, C.byref(st))
        return st

    def mesh(self, smooth=True):
        """Mesh of the current density field (surface list recomputed from phi)."""
        nt = C.c_int64()
        nv = self.lib.xo_mesh(self.h, int(smooth), C.byref(nt))
        verts = np.empty((nv, 3)); tris = np.empty((nt.value, 3), np.int32); normals = np.empty((nv, 3))
        self.lib.xo_get_mesh(self.h, verts.ctypes.data, tris.ctypes.data, normals.ctypes.data)
        return verts, tris, normals

    def surface_cells(self):
        n = self.lib.xo_surface_cells(self.h, np.empty(1, np.int32), np.empty(1, np.uint8), 0)
        cells = np.empty(max(n, 1), np.int32)
        cube = np.empty(max(n, 1), np.uint8)
        self.lib.xo_surface_cells(self.h, cells, cube, n)
        return cells[:n], cube[:n]


_REF_FIELDS = {FIELD["U"]: 0, FIELD["V"]: 1, FIELD["W"]: 2, FIELD["USAVED"]: 3, FIELD["VSAVED"]: 4,
               FIELD["WSAVED"]: 5, FIELD["MATERIAL"]: 6, FIELD["PHI"]: 7, FIELD["PRESSURE"]: 8, FIELD["RHS"]: 9,
               FIELD["MIC_INV"]: 10}


class RefSim:
    """The reference's own FluidSimulator, driven stage by stage."""

    def __init__(self, ni, nj, nk, dx, params=None):
        self.lib = ref()
        self.ni, self.nj, self.nk, self.dx = ni, nj, nk, dx
        self.h = self.lib.xr_create(ni, nj, nk, dx)
        p = params or default_params()
        self.params = p
        self.lib.xr_set_params(self.h, p.flip_ratio, p.cfl, p.max_dt, p.pcg_tol, p.pcg_max_iter, p.blob_radius, p.iso)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.xr_destroy(self.h)
            self.h = None

    def get(self, which):
        shape, dt = field_shape(self.ni, self.nj, self.nk, which, 2)
        a = np.empty(shape, dt)
        assert self.lib.xr_get_field(self.h, _REF_FIELDS[which], a.ctypes.data) == 0
        return a

    def set(self, which, a):
        shape, dt = field_shape(self.ni, self.nj, self.nk, which, 2)
        a = np.ascontiguousarray(a, dtype=dt).reshape(shape)
        assert self.lib.xr_set_field(self.h, _REF_FIELDS[which], a.ctypes.data) == 0

    def set_particles(self, P, Vv):
        cols = [np.ascontiguousarray(P[:, a], np.float64) for a in range(3)] + \
               [np.ascontiguousarray(Vv[:, a], np.float64) for a in range(3)]
        self.lib.xr_set_particles(self.h, len(P), *cols)

    def particles(self):
        n = self.lib.xr_num_particles(self.h)
        cols = [np.empty(n) for _ in range(6)]
        self.lib.xr_get_particles(self.h, *cols)
        return np.stack(cols[:3], 1), np.stack(cols[3:], 1)

    def stage(self, s, dt=0.0):
        return self.lib.xr_stage(self.h, s, dt)

    def time_step(self, dt):
        return self.lib.xr_time_step(self.h, dt)

    def add_fluid_source(self, c, wx, wy, wz, v):
        self.lib.xr_add_fluid_source(self.h, np.asarray(c, np.float64), wx, wy, wz, np.asarray(v, np.float64))

    def apply_sources(self, dt, seed):
        """One _applyFluidSource of the reference, rand() restarted at srand(seed)."""
        self.lib.xr_apply_sources(self.h, dt, seed)

    def frame_step(self, dt):
        return self.lib.xr_frame_step(self.h, dt)

    def last_solve(self):
        it, res, st = C.c_int(), C.c_double(), C.c_int()
        self.lib.xr_last_solve(self.h, C.byref(it), C.byref(res), C.byref(st))
        return it.value, res.value, st.value

    def mesh(self, smooth=True):
        nt = C.c_int64()
        nv = self.lib.xr_mesh(self.h, int(smooth), C.byref(nt))
        verts = np.empty((nv, 3)); tris = np.empty((nt.value, 3), np.int32)
        self.lib.xr_get_mesh(self.h, verts.ctypes.data, tris.ctypes.data)
        return verts, tris

    def write_ply(self, path):
        self.lib.xr_write_ply(self.h, os.fsencode(path))

    def surface_cells(self):
        n = self.lib.xr_surface_cells(self.h, np.empty(1, np.int32), 0)
        cells = np.empty(max(n, 1), np.int32)
        self.lib.xr_surface_cells(self.h, cells, n)
        return cells[:n]


def ref_rand(seed, n):
    """The first n values of the reference build's rand() after srand(seed), and RAND_MAX."""
    out = np.empty(max(n, 1), np.uint32)
    rmax = ref().xr_rand(seed, n, out.ctypes.data)
    return out[:n], rmax


def dam_break_scene(ni, nj, nk, dx, fx, fy, fz=1.0):
    """Material of a dam break: 1-cell walls + addCubicFluid((0,0,0),(fx*Lx, fy*Ly, fz*Lz))."""
    return (0.0, 0.0, 0.0), (fx * ni * dx, fy * nj * dx, fz * nk * dx)


def build_oracle_dam_break(ni, nj, nk, dx, frac, seed, mode=1, params=None, gravity=-9.81):
    o = OracleSim(ni, nj, nk, dx, params)
    lu, rd = dam_break_scene(ni, nj, nk, dx, *frac)
    o.lib.xo_set_solid_boundary(o.h)
    o.lib.xo_add_cubic_fluid(o.h, np.array(lu), np.array(rd))
    o.lib.xo_set_solid_boundary(o.h)
    o.lib.xo_seed_particles(o.h, mode, seed)
    o.lib.xo_add_constant_force(o.h, 0.0, gravity, 0.0)
    return o
