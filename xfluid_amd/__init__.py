"""xfluid_amd -- Python host binding of the MI355X (gfx950) FLIP/PIC substep library.

The compute path is libxfluid_hip.so (hand-written HIP kernels + C ABI, see
include/xfluid.h).  This module is a thin ctypes mirror of the reference's
FluidSimulator interface (xFluid/xFluid/FluidSimulator.h:19-88) with the same
method names and argument meaning; there is no CPU fallback: if the library or a
HIP device is missing, construction raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxfluid_hip.so")

FIELD = dict(U=0, V=1, W=2, USAVED=3, VSAVED=4, WSAVED=5, MATERIAL=6, PHI=7, PRESSURE=8, RHS=9, MIC_INV=10)
STAGE = dict(DENSITY=0, MATERIAL=1, SAVE=2, FORCES=3, SOLVE=4, APPLY=5, EXTRAPOLATE=6, ADVECT=7, P2G=8)
SEED_REFERENCE, SEED_SUBCELL = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
PRESSURE_VARIATIONAL, PRESSURE_NORMAL = 0, 1

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class XfParams(C.Structure):
    """Mirror of xf_params (include/xfluid.h)."""
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
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("outer", C.c_int), ("fallback", C.c_int),
                ("residual", C.c_double), ("rhs_norm", C.c_double), ("ms", C.c_double),
                ("sweep_ms", C.c_double), ("sweeps", C.c_int), ("sweep_fused", C.c_int),
                ("box_cells", C.c_int64)]


class XfStepTiming(C.Structure):
    _fields_ = [("ms", C.c_double * 9), ("total_ms", C.c_double)]


class XfError(RuntimeError):
    pass


_LIB = None


def _share_torch_hip_runtime():
    """PyTorch-ROCm ships its own libamdhip64.so with the same soname as the
    system one.  Two HIP runtimes in one process do not see the same devices
    (the second one finds no GPU), so when torch is installed it is imported
    first and the library binds to torch's runtime; torch device tensors can
    then be passed to the *_device entry points.  XFLUID_NO_TORCH=1 skips this."""
    if os.environ.get("XFLUID_NO_TORCH") == "1":
        return
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Load libxfluid_hip.so (raises if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise XfError(f"{LIB_PATH} is missing: build it with `make -C xfluid_amd` (hipcc, gfx950)")
    _share_torch_hip_runtime()
    L = C.CDLL(LIB_PATH)
    V = C.c_void_p
    L.xf_last_error.restype = C.c_char_p
    L.xf_params_default.argtypes = [C.POINTER(XfParams)]
    L.xf_device_count.restype = C.c_int
    L.xf_sim_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(V)]
    L.xf_sim_destroy.argtypes = [V]
    L.xf_sim_set_cell_width.argtypes = [V, C.c_double]
    L.xf_sim_set_params.argtypes = [V, C.POINTER(XfParams)]
    L.xf_sim_get_params.argtypes = [V, C.POINTER(XfParams)]
    L.xf_sim_set_solid_boundary.argtypes = [V]
    L.xf_sim_add_sphere_fluid.argtypes = [V] + [C.c_double] * 4
    L.xf_sim_add_cubic_fluid.argtypes = [V, _dp, _dp]
    L.xf_sim_add_constant_force.argtypes = [V] + [C.c_double] * 3
    L.xf_sim_add_fluid_source.argtypes = [V, _dp] + [C.c_double] * 3 + [_dp]
    L.xf_sim_set_source_seed.argtypes = [V, C.c_uint64]
    L.xf_sim_seed_particles.argtypes = [V, C.c_int, C.c_uint64]
    L.xf_sim_init.argtypes = [V]
    L.xf_sim_frame_step.argtypes = [V, C.c_double]
    L.xf_sim_time_step.argtypes = [V, C.c_double, C.c_int]
    L.xf_sim_stage.argtypes = [V, C.c_int, C.c_double]
    L.xf_sim_last_solve.argtypes = [V, C.POINTER(XfSolveStats)]
    L.xf_sim_set_timing.argtypes = [V, C.c_int]
    L.xf_sim_last_timing.argtypes = [V, C.POINTER(XfStepTiming)]
    L.xf_sim_synchronize.argtypes = [V]
    L.xf_sim_debug_fail_sweeps.argtypes = [V]
    L.xf_sim_num_particles.restype = C.c_int64
    L.xf_sim_num_particles.argtypes = [V]
    L.xf_sim_set_particles.argtypes = [V, C.c_int64] + [_dp] * 6
    L.xf_sim_get_particles.argtypes = [V] + [_dp] * 6
    L.xf_sim_field_bytes.restype = C.c_size_t
    L.xf_sim_field_bytes.argtypes = [V, C.c_int]
    L.xf_sim_set_field.argtypes = [V, C.c_int, V, C.c_size_t]
    L.xf_sim_get_field.argtypes = [V, C.c_int, V, C.c_size_t]
    L.xf_sim_frame_count.argtypes = [V]
    L.xf_sim_surface_cells.restype = C.c_int64
    L.xf_sim_surface_cells.argtypes = [V, V, V, C.c_int64]
    L.xf_sim_project_fields.argtypes = [V, V, V, V, C.c_int]
    L.xf_sim_project_adjoint.argtypes = [V, V, V, V, C.c_int]
    L.xf_sim_sample_adjoint.argtypes = [V, C.c_int64, V, V, V, V, V, V, C.c_int, C.c_int]
    L.xf_sim_advect_adjoint.argtypes = [V, C.c_int64, V, C.c_double, V, V, V, V, V, C.c_int]
    L.xf_sim_p2g_adjoint.argtypes = [V, C.c_int64, V, V, V, V, V, V, V, C.c_int]
    L.xf_sim_density_adjoint.argtypes = [V, C.c_int64, V, V, V, C.c_int]
    L.xf_sim_p2g_adjoint_noalloc.argtypes = [V, C.c_int64, V, V, V, V, V, V, V]
    L.xf_sim_extrapolate_adjoint.argtypes = [V, V, V, V, C.c_int]
    L.xf_sim_set_particles_device.argtypes = [V, C.c_int64, V, V]
    L.xf_sim_get_particles_device.argtypes = [V, V, V]
    L.xf_sim_set_field_device.argtypes = [V, C.c_int, V, C.c_size_t]
    L.xf_sim_get_field_device.argtypes = [V, C.c_int, V, C.c_size_t]
    L.xf_sim_generate_mesh.argtypes = [V, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.xf_sim_get_mesh.argtypes = [V, V, V, V]
    L.xf_sim_write_mesh_ply.argtypes = [V, C.c_char_p]
    L.xf_sim_set_mesh_output.argtypes = [V, C.c_char_p]
    L.xf_sim_apply_matrix.argtypes = [V, _dp, _dp]
    L.xf_sim_apply_precond.argtypes = [V, _dp, _dp]
    L.xf_project.argtypes = [C.c_int, C.c_int, C.c_int, _u8, _dp, _dp, _dp, V, C.POINTER(XfParams),
                             C.POINTER(XfSolveStats)]
    _LIB = L
    return L


def _chk(rc, what):
    if rc < 0:
        raise XfError(f"{what} failed ({rc}): {lib().xf_last_error().decode()}")
    return rc


def _torch_done(tensors):
    """Waits for torch's current stream on the tensors' device.  The library
    works on a non-blocking stream of its own, which does not wait for torch's
    streams: a tensor that torch is still filling (or has not yet zeroed)
    must be complete before the library reads or writes it.  Every entry point
    returns with the library's own stream drained, so nothing is needed after."""
    import torch
    torch.cuda.current_stream(tensors[0].device).synchronize()


def default_params(**kw):
    p = XfParams()
    lib().xf_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def device_count():
    return lib().xf_device_count()


class FluidSimulator:
    """Mirror of the reference's FluidSimulator (FluidSimulator.h:19-88) on one GPU.

    Method names follow the reference (setCellWidth, addCubicFluid, frameStep,
    timeStep, ...); state lives in HBM and is read back with get_field /
    particles.  Failed solves do not raise: timeStep returns the status (>0 =
    not converged, as the reference's solve() == false) and the state stays
    defined.
    """

    def __init__(self, ni, nj, nk, params=None):
        self.L = lib()
        self.ni, self.nj, self.nk = ni, nj, nk
        h = C.c_void_p()
        _chk(self.L.xf_sim_create(ni, nj, nk, C.byref(h)), "xf_sim_create")
        self.h = h
        self.params = params or default_params()
        _chk(self.L.xf_sim_set_params(self.h, C.byref(self.params)), "xf_sim_set_params")

    def close(self):
        if getattr(self, "h", None):
            self.L.xf_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # -- reference interface -------------------------------------------------
    def setCellWidth(self, dx):
        self.dx = dx
        _chk(self.L.xf_sim_set_cell_width(self.h, dx), "setCellWidth")

    def setSolidBoundary(self):
        _chk(self.L.xf_sim_set_solid_boundary(self.h), "setSolidBoundary")

    def addSpereFluid(self, x, y, z, r):
        _chk(self.L.xf_sim_add_sphere_fluid(self.h, x, y, z, r), "addSpereFluid")

    def addCubicFluid(self, lu, rd):
        _chk(self.L.xf_sim_add_cubic_fluid(self.h, np.asarray(lu, np.float64), np.asarray(rd, np.float64)),
             "addCubicFluid")

    def addConstantForce(self, fx, fy, fz):
        _chk(self.L.xf_sim_add_constant_force(self.h, fx, fy, fz), "addConstantForce")

    def addFluidSource(self, center, wx, wy, wz, velocity):
        """FluidSimulator::addFluidSource (FluidSimulator.cpp:152-166): a cubic source
        of half-widths (wx, wy, wz) around `center`, emitting with `velocity`."""
        _chk(self.L.xf_sim_add_fluid_source(self.h, np.asarray(center, np.float64), wx, wy, wz,
                                            np.asarray(velocity, np.float64)), "addFluidSource")

    def set_source_seed(self, seed):
        """Seed (and restart) the fluid sources' jitter stream."""
        _chk(self.L.xf_sim_set_source_seed(self.h, seed), "set_source_seed")

    def initParticles(self, mode=SEED_REFERENCE, seed=0):
        _chk(self.L.xf_sim_seed_particles(self.h, mode, seed), "initParticles")

    def init(self):
        _chk(self.L.xf_sim_init(self.h), "init")

    def frameStep(self, dt):
        return _chk(self.L.xf_sim_frame_step(self.h, dt), "frameStep")

    def timeStep(self, dt, outputMesh=False):
        return _chk(self.L.xf_sim_time_step(self.h, dt, int(outputMesh)), "timeStep")

    # -- extensions ------------------------------------------------------------
    def set_params(self, p):
        self.params = p
        _chk(self.L.xf_sim_set_params(self.h, C.byref(p)), "set_params")

    def stage(self, s, dt=0.0):
        return _chk(self.L.xf_sim_stage(self.h, s, dt), "stage")

    def last_solve(self):
        st = XfSolveStats()
        self.L.xf_sim_last_solve(self.h, C.byref(st))
        return st

    def set_timing(self, on):
        self.L.xf_sim_set_timing(self.h, int(on))

    def last_timing(self):
        t = XfStepTiming()
        self.L.xf_sim_last_timing(self.h, C.byref(t))
        return t

    def synchronize(self):
        _chk(self.L.xf_sim_synchronize(self.h), "synchronize")

    def debug_fail_sweeps(self):
        """Test hook: the next solve's pipelined sweeps give up at once (as after a
        timed-out wait) and the solve is redone with per-hyperplane sweeps."""
        _chk(self.L.xf_sim_debug_fail_sweeps(self.h), "debug_fail_sweeps")

    def num_particles(self):
        return self.L.xf_sim_num_particles(self.h)

    def set_particles(self, P, Vv):
        P = np.asarray(P, np.float64)
        Vv = np.asarray(Vv, np.float64)
        cols = [np.ascontiguousarray(P[:, a]) for a in range(3)] + [np.ascontiguousarray(Vv[:, a]) for a in range(3)]
        _chk(self.L.xf_sim_set_particles(self.h, len(P), *cols), "set_particles")

    def particles(self):
        n = self.num_particles()
        cols = [np.empty(n) for _ in range(6)]
        if n:
            _chk(self.L.xf_sim_get_particles(self.h, *cols), "get_particles")
        return np.stack(cols[:3], 1), np.stack(cols[3:], 1)

    def _shape(self, which):
        ni, nj, nk = self.ni, self.nj, self.nk
        d = self.params.sub_division
        if which in (0, 3):
            return (nk, nj, ni + 1), np.float64
        if which in (1, 4):
            return (nk, nj + 1, ni), np.float64
        if which in (2, 5):
            return (nk + 1, nj, ni), np.float64
        if which == 6:
            return (nk, nj, ni), np.uint8
        if which == 7:
            return (d * nk + 1, d * nj + 1, d * ni + 1), np.float64
        return (nk, nj, ni), np.float64

    def get_field(self, which):
        shape, dt = self._shape(which)
        a = np.empty(shape, dt)
        _chk(self.L.xf_sim_get_field(self.h, which, a.ctypes.data, a.nbytes), "get_field")
        return a

    def set_field(self, which, a):
        shape, dt = self._shape(which)
        a = np.ascontiguousarray(np.asarray(a, dt).reshape(shape))
        _chk(self.L.xf_sim_set_field(self.h, which, a.ctypes.data, a.nbytes), "set_field")

    def surface_cells(self):
        n = _chk(self.L.xf_sim_surface_cells(self.h, None, None, 0), "surface_cells")
        cells = np.empty(max(n, 1), np.int32)
        cube = np.empty(max(n, 1), np.uint8)
        if n:
            _chk(self.L.xf_sim_surface_cells(self.h, cells.ctypes.data, cube.ctypes.data, n), "surface_cells")
        return cells[:n], cube[:n]

    def _faces_call(self, fn, u, v, w, what):
        """Runs fn on three face arrays in place: numpy float64 (host) or
        contiguous float64 device tensors (anything with data_ptr() and is_cuda)."""
        arrs = (u, v, w)
        if all(getattr(a, "is_cuda", False) for a in arrs):
            for a in arrs:
                if not a.is_contiguous() or str(a.dtype) != "torch.float64":
                    raise ValueError(f"{what}: device arrays must be contiguous float64")
            _torch_done(arrs)
            return _chk(fn(self.h, arrs[0].data_ptr(), arrs[1].data_ptr(), arrs[2].data_ptr(), 1), what)
        for a in arrs:
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
                raise ValueError(f"{what}: host arrays must be C-contiguous float64 numpy arrays")
        return _chk(fn(self.h, u.ctypes.data, v.ctypes.data, w.ctypes.data, 0), what)

    def project_fields(self, u, v, w):
        """u' = u - G A^-1 B u in place on the sim's material; returns the PCG status."""
        return self._faces_call(self.L.xf_sim_project_fields, u, v, w, "project_fields")

    def extrapolate_adjoint(self, gu, gv, gw):
        """Vector-Jacobian product of the velocity extrapolation on the sim's material, in place."""
        return self._faces_call(self.L.xf_sim_extrapolate_adjoint, gu, gv, gw, "extrapolate_adjoint")

    # -- device (torch) state transfer -------------------------------------------
    @staticmethod
    def _dev(t, what, dtype="torch.float64"):
        if not getattr(t, "is_cuda", False) or not t.is_contiguous() or str(t.dtype) != dtype:
            raise ValueError(f"{what}: expected a contiguous {dtype} device tensor")
        _torch_done((t,))
        return t.data_ptr()

    def set_particles_device(self, pos, vel):
        """Positions and velocities from n x 3 float64 device tensors (ids 0..n-1 in row order)."""
        n = pos.shape[0]
        _chk(self.L.xf_sim_set_particles_device(self.h, n, self._dev(pos, "pos"), self._dev(vel, "vel")),
             "set_particles_device")

    def get_particles_device(self, pos, vel):
        """Positions and velocities into n x 3 float64 device tensors, in id order."""
        _chk(self.L.xf_sim_get_particles_device(self.h, self._dev(pos, "pos"), self._dev(vel, "vel")),
             "get_particles_device")

    def set_field_device(self, which, t):
        dt = "torch.uint8" if which == FIELD["MATERIAL"] else "torch.float64"
        _chk(self.L.xf_sim_set_field_device(self.h, which, self._dev(t, "field", dt), t.numel() * t.element_size()),
             "set_field_device")

    def get_field_device(self, which, t):
        dt = "torch.uint8" if which == FIELD["MATERIAL"] else "torch.float64"
        _chk(self.L.xf_sim_get_field_device(self.h, which, self._dev(t, "field", dt), t.numel() * t.element_size()),
             "get_field_device")

    def project_adjoint(self, gu, gv, gw):
        """Vector-Jacobian product of project_fields, in place; returns the PCG status."""
        return self._faces_call(self.L.xf_sim_project_adjoint, gu, gv, gw, "project_adjoint")

    @staticmethod
    def _ptrs(arrays, what):
        """data pointers of host float64 arrays or device tensors, and the device flag"""
        if all(getattr(a, "is_cuda", False) for a in arrays):
            for a in arrays:
                if not a.is_contiguous() or str(a.dtype) != "torch.float64":
                    raise ValueError(f"{what}: device arrays must be contiguous float64")
            _torch_done(arrays)
            return [a.data_ptr() for a in arrays], 1
        for a in arrays:
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
                raise ValueError(f"{what}: host arrays must be C-contiguous float64 numpy arrays")
        return [a.ctypes.data for a in arrays], 0

    def sample_adjoint(self, pos, gvel, gpos, gu, gv, gw, saved=False):
        """VJP of G2P sampling; gpos, gu, gv, gw are added to in place."""
        p, dev = self._ptrs((pos, gvel, gpos, gu, gv, gw), "sample_adjoint")
        return _chk(self.L.xf_sim_sample_adjoint(self.h, len(pos), *p, int(saved), dev), "sample_adjoint")

    def advect_adjoint(self, pos, dt, gnew, gpos, gu, gv, gw):
        """VJP of the RK position update; gpos is written, gu, gv, gw are added to."""
        p, dev = self._ptrs((pos, gnew, gpos, gu, gv, gw), "advect_adjoint")
        return _chk(self.L.xf_sim_advect_adjoint(self.h, len(pos), p[0], dt, *p[1:], dev), "advect_adjoint")

    def p2g_adjoint(self, pos, vel, gu, gv, gw, gvel, gpos):
        """VJP of P2G; gvel is written, gpos is added to."""
        p, dev = self._ptrs((pos, vel, gu, gv, gw, gvel, gpos), "p2g_adjoint")
        return _chk(self.L.xf_sim_p2g_adjoint(self.h, len(pos), *p, dev), "p2g_adjoint")

    def p2g_adjoint_noalloc(self, pos, vel, gu, gv, gw, gvel, gpos):
        """VJP of P2G with no device allocation (device tensors only; the sums use
        the library's per-GPU static scratch); gvel is written, gpos is added to."""
        p, dev = self._ptrs((pos, vel, gu, gv, gw, gvel, gpos), "p2g_adjoint_noalloc")
        if not dev:
            raise ValueError("p2g_adjoint_noalloc: arrays must be device tensors")
        return _chk(self.L.xf_sim_p2g_adjoint_noalloc(self.h, len(pos), *p), "p2g_adjoint_noalloc")

    def density_adjoint(self, pos, gphi, gpos):
        """VJP of the particle-blob density on the sub-vertex lattice; gpos is added to."""
        p, dev = self._ptrs((pos, gphi, gpos), "density_adjoint")
        return _chk(self.L.xf_sim_density_adjoint(self.h, len(pos), *p, dev), "density_adjoint")

    def generate_mesh(self, smooth=True):
        """FluidMesher::generateMesh on the current density field -> (verts, tris, normals)."""
        nv, nt = C.c_int64(), C.c_int64()
        _chk(self.L.xf_sim_generate_mesh(self.h, int(smooth), C.byref(nv), C.byref(nt)), "generate_mesh")
        verts = np.empty((nv.value, 3), np.float64)
        tris = np.empty((nt.value, 3), np.int32)
        normals = np.empty((nv.value, 3), np.float64)
        _chk(self.L.xf_sim_get_mesh(self.h, verts.ctypes.data, tris.ctypes.data, normals.ctypes.data), "get_mesh")
        return verts, tris, normals

    def write_mesh_ply(self, path):
        """TriangleMesh::writeMeshToPLY of the last generated mesh."""
        _chk(self.L.xf_sim_write_mesh_ply(self.h, os.fsencode(path)), "write_mesh_ply")

    def set_mesh_output(self, prefix):
        """Config::outputFild: the last substep of each frame writes <prefix><frame>.ply."""
        _chk(self.L.xf_sim_set_mesh_output(self.h, None if prefix is None else os.fsencode(prefix)),
             "set_mesh_output")

    def apply_matrix(self, x):
        x = np.ascontiguousarray(x, np.float64).ravel()
        y = np.empty_like(x)
        _chk(self.L.xf_sim_apply_matrix(self.h, x, y), "apply_matrix")
        return y

    def apply_precond(self, r):
        r = np.ascontiguousarray(r, np.float64).ravel()
        z = np.empty_like(r)
        _chk(self.L.xf_sim_apply_precond(self.h, r, z), "apply_precond")
        return z


def project(material, u, v, w, params=None):
    """Standalone projection (VariationalPressureSolver solve + applyPressure).

    material: (nk,nj,ni) uint8; u,v,w MAC arrays (updated copies are returned).
    Returns (u, v, w, pressure, stats).
    """
    L = lib()
    material = np.ascontiguousarray(material, np.uint8)
    nk, nj, ni = material.shape
    u = np.ascontiguousarray(u, np.float64).copy()
    v = np.ascontiguousarray(v, np.float64).copy()
    w = np.ascontiguousarray(w, np.float64).copy()
    p = np.empty((nk, nj, ni))
    st = XfSolveStats()
    prm = params or default_params()
    rc = L.xf_project(ni, nj, nk, material.ravel(), u.ravel(), v.ravel(), w.ravel(), p.ctypes.data, C.byref(prm),
                      C.byref(st))
    _chk(rc, "xf_project")
    return u, v, w, p, st
