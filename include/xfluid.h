/*
 * xfluid.h -- C ABI of the MI355X (gfx950) FLIP/PIC liquid substep library.
 *
 * One opaque handle (xf_sim) owns the device state of one simulator and one
 * HIP stream.  Every entry point replaces a public method of the reference's
 * FluidSimulator (xFluid/xFluid/FluidSimulator.h:19-88) or the pressure-solver
 * seam (variationalpressure/variational_pressuresolver.h:35-44); the comment on
 * each declaration names the reference interface it stands in for.
 *
 * Conventions
 *   - All host buffers are caller-owned; the library copies in and out.
 *   - Grids are flat i-fastest arrays, index = i + ni*(j + nj*k), exactly the
 *     reference's Array3d layout (array3d.h:143-146).  The MAC face arrays are
 *     u: (ni+1)*nj*nk, v: ni*(nj+1)*nk, w: ni*nj*(nk+1) doubles.
 *   - Material is one byte per cell: 0 air, 1 fluid, 2 solid (material.h:4-8).
 *   - Return status: 0 ok; >0 the pressure solve did not converge (the
 *     reference's solve()==false, but the state stays defined: the last PCG
 *     iterate is applied); <0 argument or HIP error, text in xf_last_error().
 *   - Calls on one handle are serialised by the caller; one handle per thread.
 */
#ifndef XFLUID_H
#define XFLUID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xf_sim xf_sim;

/* Simulation parameters.  Defaults (xf_params_default) are the reference's
 * hard-wired values; every one is a member or literal of the reference:
 *   flip_ratio   FluidSimulator.h:187          cfl          FluidSimulator.h:224
 *   max_dt       FluidSimulator.h:223          pcg_tol/iter pcg_solver.h:231
 *   mic_tau/sig  pcg_solver.h:87 (hard-wired, pcg_solver.h:302)
 *   blob_radius  3*r_p, FluidSimulator.cpp:210,380-385 (<=0 selects it)
 *   iso          FluidSimulator.h:202          sub_division FluidSimulator.h:203
 */
typedef struct xf_params {
    double flip_ratio;        /* FLIP fraction of the G2P blend                    */
    double cfl;               /* CFL number of frame_step; extrapolation layers = (int)(cfl+5) */
    double max_dt;            /* upper bound of a substep                          */
    double pcg_tol;           /* relative inf-norm tolerance of the pressure solve */
    int    pcg_max_iter;      /* PCG iteration cap (per solve; inner cap in mixed mode) */
    int    sub_division;      /* density/marching-cubes subdivision level (2)      */
    double mic_tau;           /* MIC(0) modification parameter                     */
    double mic_sigma;         /* MIC(0) pivot guard ratio                          */
    double blob_radius;       /* particle-blob radius (absolute); <=0: 3*r_p       */
    double iso;               /* density iso level for surface / marching cubes    */
    int    precision;         /* XF_PREC_FP64 or XF_PREC_MIXED                      */
    int    integrator;        /* XF_INTEGRATOR_RK4 (reference) or XF_INTEGRATOR_RK2 */
    int    p2g_kernel;        /* XF_P2G_WYVILL (reference) or XF_P2G_TRILINEAR      */
    int    tail_skip;         /* 1: the last n%6 particles are not advected (reference chunking) */
    int    density_every_substep; /* 1: density each substep (reference); 0: only when a mesh is output */
    int    sweep_kernel;      /* MIC(0) sweep schedule: 0 auto (= 9), 1 per-hyperplane launches,
                                 9 pipelined, one wavefront per 8x8 tile of i-lines */
    double mixed_inner_tol;   /* mixed mode: relative reduction asked of an FP32 inner solve; later
                                 inner solves stop at half of what is left to the outer tolerance */
    int    mixed_max_outer;   /* mixed mode: cap on FP64 refinement steps           */
    int    pressure_solver;   /* XF_PRESSURE_VARIATIONAL (the reference's, FluidSimulator.h:215) or
                                 XF_PRESSURE_NORMAL (NormalPressureSolver, FluidSimulator.h:216) */
    int    reserved[4];
} xf_params;

/* Pressure operators of the projection seam (FluidSimulator.h:214-216). */
enum {
    XF_PRESSURE_VARIATIONAL = 0, /* VariationalPressureSolver: fluid + fluid-adjacent solid rows,
                                    weights fluid 1, air 2, solid 0.5 (variational_pressuresolver.cpp:231-326) */
    XF_PRESSURE_NORMAL      = 1  /* NormalPressureSolver: fluid rows, diagonal = non-solid neighbours,
                                    -1 per fluid neighbour; solid faces dropped from the divergence;
                                    apply with air ghost 0 and solid faces set to zero normal
                                    velocity (pressuresolver.cpp:109-212, 220-347) */
};

/* Particle seeding modes of xf_sim_seed_particles. */
enum {
    XF_SEED_REFERENCE = 0,    /* 8 points at c +- dx/4 with +-0.01 absolute jitter (FluidSimulator.cpp:347-370) */
    XF_SEED_SUBCELL   = 1     /* one point uniform in each 2x2x2 sub-cell, same point order */
};

enum { XF_PREC_FP64 = 0, XF_PREC_MIXED = 1 };
enum { XF_INTEGRATOR_RK4 = 0, XF_INTEGRATOR_RK2 = 1 };
enum { XF_P2G_WYVILL = 0, XF_P2G_TRILINEAR = 1 };

/* Field identifiers for xf_sim_get_field / xf_sim_set_field. */
enum {
    XF_FIELD_U = 0, XF_FIELD_V = 1, XF_FIELD_W = 2,
    XF_FIELD_USAVED = 3, XF_FIELD_VSAVED = 4, XF_FIELD_WSAVED = 5,
    XF_FIELD_MATERIAL = 6,    /* uint8 per cell                                    */
    XF_FIELD_PHI = 7,         /* (d*ni+1)(d*nj+1)(d*nk+1) doubles                  */
    XF_FIELD_PRESSURE = 8,    /* ni*nj*nk doubles, last solve                      */
    XF_FIELD_RHS = 9,         /* ni*nj*nk doubles, last assembled right-hand side  */
    XF_FIELD_MIC_INV = 10,    /* ni*nj*nk doubles, MIC(0) reciprocal pivots        */
    XF_FIELD_COUNT = 11
};

/* Stage identifiers for xf_sim_stage: the stages of FluidSimulator::timeStep
 * (FluidSimulator.cpp:265-335), runnable one at a time for parity checks. */
enum {
    XF_STAGE_DENSITY = 0,     /* updateFeild + surface cells  directeddistancefield.cpp:122-244 */
    XF_STAGE_MATERIAL = 1,    /* _updateMaterial               FluidSimulator.cpp:402-427 */
    XF_STAGE_SAVE = 2,        /* _saveVelocity                 FluidSimulator.cpp:1095-1100 */
    XF_STAGE_FORCES = 3,      /* _applyForces                  FluidSimulator.cpp:429-499 */
    XF_STAGE_SOLVE = 4,       /* VariationalPressureSolver::solve variational_pressuresolver.cpp:50-118 */
    XF_STAGE_APPLY = 5,       /* applyPressure                 variational_pressuresolver.cpp:120-220 */
    XF_STAGE_EXTRAPOLATE = 6, /* _expolateVelocity             FluidSimulator.cpp:511-650 */
    XF_STAGE_ADVECT = 7,      /* G2P blend + RK4 + collision   FluidSimulator.cpp:652-723 */
    XF_STAGE_P2G = 8,         /* _updateVelocityByParticles    FluidSimulator.cpp:909-1085 */
    XF_STAGE_COUNT = 9
};

typedef struct xf_solve_stats {
    int    status;            /* 0 converged, 1 not converged, 2 rho invalid       */
    int    iterations;        /* PCG iterations (mixed: total inner iterations)    */
    int    outer;             /* mixed: FP64 refinement steps; FP64: 1             */
    int    fallback;          /* 1: a pipelined sweep wait timed out and the solve was
                                 redone with per-hyperplane sweeps                  */
    double residual;          /* final inf-norm residual (FP64)                    */
    double rhs_norm;          /* inf-norm of the right-hand side                   */
    double ms;                /* timing on: wall time of the solve on the device,
                                 ms (events around it); 0 with timing off          */
    double sweep_ms;          /* timing on, pipelined sweeps: summed device time of
                                 the MIC(0) sweep launches of every 5th
                                 preconditioner application, ms                    */
    int    sweeps;            /* number of sweep launches summed in sweep_ms       */
    int    sweep_fused;       /* pipelined sweeps: 1 forward + backward in one launch
                                 (all tiles resident at once), 0 two launches;
                                 -1 per-hyperplane sweeps                           */
    int64_t box_cells;        /* cells in the bounding box of the active rows: the
                                 cells the PCG iteration kernels and the pipelined
                                 sweeps visit (the other rows are empty)           */
} xf_solve_stats;

/* Per-substep stage timings (device events, ms), filled when timing is on. */
typedef struct xf_step_timing {
    double ms[XF_STAGE_COUNT];
    double total_ms;
} xf_step_timing;

/* ---- library ---------------------------------------------------------- */
const char* xf_last_error(void);
void xf_params_default(xf_params* p);
int  xf_device_count(void);

/* ---- construction: FluidSimulator(int,int,int) / setSize / setCellWidth
 *      (FluidSimulator.h:23-30; FluidSimulator.cpp:26-62) ------------------ */
int  xf_sim_create(int ni, int nj, int nk, xf_sim** out);
void xf_sim_destroy(xf_sim* s);
int  xf_sim_set_cell_width(xf_sim* s, double dx);
int  xf_sim_set_params(xf_sim* s, const xf_params* p);
int  xf_sim_get_params(const xf_sim* s, xf_params* p);

/* ---- scene: setSolidBoundary / addSpereFluid / addCubicFluid /
 *      addConstantForce (FluidSimulator.h:35-50; FluidSimulator.cpp:64-150) */
int  xf_sim_set_solid_boundary(xf_sim* s);
int  xf_sim_add_sphere_fluid(xf_sim* s, double x, double y, double z, double r);
int  xf_sim_add_cubic_fluid(xf_sim* s, const double lu[3], const double rd[3]);
int  xf_sim_add_constant_force(xf_sim* s, double fx, double fy, double fz);

/* ---- fluid sources: addFluidSource (FluidSimulator.h:53; FluidSimulator.cpp:152-166)
 *      and CubicFluidSource (fluidsource.cpp:8-33, 75-92).  The source cells are
 *      fixed when the source is added: the cells from (int)((c - w) / dx) to
 *      (int)((c + w) / dx) per axis, i slowest, that are in range and not solid
 *      at that moment.  After every substep's advection each source marks its
 *      cells fluid and appends (int)(8 |v| dt / dx) particles per cell, with
 *      velocity v, at the cell centre plus randomDouble(-0.4999 dx, 0.4999 dx)
 *      per axis (xfluidutil.cpp:67-70) drawn in x, y, z order from a 31-bit
 *      SplitMix64 stream (RAND_MAX = 2^31 - 1) in place of the reference's
 *      unseeded rand().  Particle count and ids grow; ids continue in emission
 *      order as the reference's push_back does. */
int  xf_sim_add_fluid_source(xf_sim* s, const double center[3], double wx, double wy, double wz,
                             const double velocity[3]);
/* Seed of the source stream (default 0); also restarts the stream. */
int  xf_sim_set_source_seed(xf_sim* s, uint64_t seed);

/* ---- init: initParticles / init (FluidSimulator.h:61-66;
 *      FluidSimulator.cpp:168-222).  The reference seeds with unseeded rand();
 *      here the jitter comes from SplitMix64(seed) so runs are reproducible. */
int  xf_sim_seed_particles(xf_sim* s, int mode, uint64_t seed);
int  xf_sim_init(xf_sim* s);

/* ---- stepping: frameStep / timeStep (FluidSimulator.h:72-88;
 *      FluidSimulator.cpp:224-335) ------------------------------------------ */
int  xf_sim_frame_step(xf_sim* s, double dt);
int  xf_sim_time_step(xf_sim* s, double dt, int output_mesh);
int  xf_sim_stage(xf_sim* s, int stage, double dt);
int  xf_sim_last_solve(const xf_sim* s, xf_solve_stats* st);
int  xf_sim_set_timing(xf_sim* s, int on);
int  xf_sim_last_timing(const xf_sim* s, xf_step_timing* t);
int  xf_sim_synchronize(xf_sim* s);
/* Testing the sweep-timeout fallback (no counterpart in the reference): sets
 * the pipelined sweeps' error word, as a timed-out wait would, so that every
 * pipelined wait of the next solve gives up at once; that solve is then redone
 * with per-hyperplane sweeps (xf_solve_stats.fallback = 1). */
int  xf_sim_debug_fail_sweeps(xf_sim* s);

/* ---- state transfer (the reference keeps these as members,
 *      FluidSimulator.h:179-197; here they live in HBM) -------------------- */
int64_t xf_sim_num_particles(const xf_sim* s);
int  xf_sim_set_particles(xf_sim* s, int64_t n, const double* px, const double* py, const double* pz,
                          const double* vx, const double* vy, const double* vz);
int  xf_sim_get_particles(xf_sim* s, double* px, double* py, double* pz,
                          double* vx, double* vy, double* vz);
/* Device-memory variants (no host copy): positions and velocities as n x 3
 * interleaved arrays in particle-id order (the order they were set in; the
 * library keeps its own copies sorted by cell); fields as in xf_sim_set_field. */
int  xf_sim_set_particles_device(xf_sim* s, int64_t n, const double* pos, const double* vel);
int  xf_sim_get_particles_device(xf_sim* s, double* pos, double* vel);
int  xf_sim_set_field_device(xf_sim* s, int which, const void* dev, size_t bytes);
int  xf_sim_get_field_device(xf_sim* s, int which, void* dev, size_t bytes);
size_t xf_sim_field_bytes(const xf_sim* s, int which);
int  xf_sim_set_field(xf_sim* s, int which, const void* host, size_t bytes);
int  xf_sim_get_field(xf_sim* s, int which, void* host, size_t bytes);
int  xf_sim_frame_count(const xf_sim* s);

/* ---- surface / marching cubes: DirectedDistanceField::updateSurfaceInfo
 *      sub-surface cell list and FluidMesher cube classification
 *      (directeddistancefield.cpp:171-244,330-352; fluidmesher.cpp:395-420).
 *      Cells come in the reference's traversal order; each entry is the flat
 *      sub-cell index i + (d*ni)*(j + (d*nj)*k), cube_index the 8-bit
 *      marching-cubes case.  Returns the count (or <0 on error); at most
 *      `cap` entries are written. */
int64_t xf_sim_surface_cells(xf_sim* s, int32_t* cells, uint8_t* cube_index, int64_t cap);

/* ---- surface mesh: FluidMesher::generateMesh (fluidmesher.cpp:364-392,
 *      _generateSurfaceCellMesh :395-545, vertexInterp :4-38) on the current
 *      density field, with its smoothing (_smoothMesh :547-556 ->
 *      TriangleMesh::smooth, trianglemesh.cpp:736-810) and vertex normals
 *      (trianglemesh.cpp:332-360).  Vertices and triangles come in the
 *      reference's order.  smooth = 0 skips the smoothing pass. */
int  xf_sim_generate_mesh(xf_sim* s, int smooth, int64_t* nverts, int64_t* ntris);
/* Copies the last mesh: verts 3*nverts doubles, tris 3*ntris ints, normals
 * 3*nverts doubles; any pointer may be NULL. */
int  xf_sim_get_mesh(xf_sim* s, double* verts, int32_t* tris, double* normals);
/* TriangleMesh::writeMeshToPLY (trianglemesh.cpp:95-252): binary little-endian
 * PLY of the last mesh (float x,y,z; uchar 3 + int indices per face). */
int  xf_sim_write_mesh_ply(xf_sim* s, const char* path);
/* Frame output of FluidSimulator::timeStep (FluidSimulator.cpp:290-305):
 * with a prefix set, each xf_sim_time_step(..., output_mesh=1) (the last
 * substep of xf_sim_frame_step) builds the mesh and writes
 * <prefix><frame>.ply.  NULL disables writing (the mesh is still built). */
int  xf_sim_set_mesh_output(xf_sim* s, const char* prefix);

/* ---- differentiation (no counterpart in the reference, which has no
 *      gradients; for gradient-based callers such as torch.autograd) ------
 *      Both work on the sim's current material and overwrite its velocity
 *      (u,v,w) and pressure fields.  Arrays are the three MAC face arrays;
 *      device_ptrs != 0 means they are device (HBM) pointers, else host. */
/* u' = u - G A^-1 B u in place (VariationalPressureSolver::solve + applyPressure). */
int  xf_sim_project_fields(xf_sim* s, double* u, double* v, double* w, int device_ptrs);
/* Vector-Jacobian product of the projection: g := g + B^T A^-1 (-G^T g), in
 * place (A symmetric: the transposed solve is a PCG solve with A). */
int  xf_sim_project_adjoint(xf_sim* s, double* gu, double* gv, double* gw, int device_ptrs);

/* VJP of the velocity extrapolation (_expolateVelocity, FluidSimulator.cpp:
 * 511-650) on the sim's material, in place on the three face arrays: layer L
 * faces (the mean of their layer L-1 neighbours) hand g/count back to those
 * neighbours, last layer first; faces of layers >= 1 then drop theirs. */
int  xf_sim_extrapolate_adjoint(xf_sim* s, double* gu, double* gv, double* gw, int device_ptrs);

/* Particle transfers, particles as n x 3 arrays (x,y,z interleaved), on the
 * sim's current grid fields; grid gradients gu/gv/gw are ADDED to.
 * VJP of G2P sampling v_p = S(u; p_p) (_getVelocity, FluidSimulator.cpp:
 * 814-894): gvel -> gu/gv/gw and gpos (added).  saved != 0 samples the saved
 * velocities (the FLIP "old" field). */
int  xf_sim_sample_adjoint(xf_sim* s, int64_t n, const double* pos, const double* gvel, double* gpos,
                           double* gu, double* gv, double* gw, int saved, int device_ptrs);
/* VJP of the advection stage's position update: the RK4/RK2 target
 * (FluidSimulator.cpp:896-907) and the collision branch on the sim's material
 * (FluidSimulator.cpp:686-722: target kept, old position kept, or pulled back
 * from the first solid voxel): gnew (gradient on the new positions) -> gpos
 * (written), gu/gv/gw (added). */
int  xf_sim_advect_adjoint(xf_sim* s, int64_t n, const double* pos, double dt, const double* gnew,
                           double* gpos, double* gu, double* gv, double* gw, int device_ptrs);
/* VJP of P2G (FluidSimulator.cpp:909-1085) for the given particles: gu/gv/gw
 * (gradient on the face arrays) -> gvel (written) and gpos (added). */
int  xf_sim_p2g_adjoint(xf_sim* s, int64_t n, const double* pos, const double* vel, const double* gu,
                        const double* gv, const double* gw, double* gvel, double* gpos, int device_ptrs);
/* Allocation-free variant of xf_sim_p2g_adjoint.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as xf_sim_p2g_adjoint with device_ptrs = 1 (all
 * arrays are device pointers).  It calls no hipMalloc/hipFree: the weight and
 * value sums live in a static device array of 2 * XF_P2G_ADJOINT_NOALLOC_MAX_FACES
 * doubles (34.6 MB) that each GPU the library runs on holds until the process
 * ends.  Grids whose largest face array exceeds the maximum are rejected (<0).
 * Every call on one GPU shares that array: calls that overlap on one GPU race. */
#define XF_P2G_ADJOINT_NOALLOC_MAX_FACES 2162688
int  xf_sim_p2g_adjoint_noalloc(xf_sim* s, int64_t n, const double* pos, const double* vel, const double* gu,
                                const double* gv, const double* gw, double* gvel, double* gpos);
/* VJP of the particle-blob density (DirectedDistanceField::updateFeild,
 * directeddistancefield.cpp:122-169, 300-328) for the given particles: gphi
 * (gradient on the (d*ni+1)(d*nj+1)(d*nk+1) density lattice, d =
 * sub_division) -> gpos (added). */
int  xf_sim_density_adjoint(xf_sim* s, int64_t n, const double* pos, const double* gphi, double* gpos,
                            int device_ptrs);

/* ---- operator diagnostics on the last assembled system (parity hooks):
 *      y = A x (FixedSparseMatrix multiply, sparse_matrix.h:266-276) and
 *      z = (L L^T)^-1 r (PCGSolver::apply_preconditioner, pcg_solver.h:305-309).
 *      Host buffers of ni*nj*nk doubles. */
int  xf_sim_apply_matrix(xf_sim* s, const double* x, double* y);
int  xf_sim_apply_precond(xf_sim* s, const double* r, double* z);

/* ---- standalone projection: VariationalPressureSolver setters + solve() +
 *      applyPressure() (variational_pressuresolver.h:35-44).  Solves on the
 *      given material/face velocities and updates u,v,w in place; pressure
 *      (ni*nj*nk) is optional output. ------------------------------------- */
int  xf_project(int ni, int nj, int nk, const uint8_t* material,
                double* u, double* v, double* w, double* pressure,
                const xf_params* p, xf_solve_stats* st);

#ifdef __cplusplus
}
#endif
#endif /* XFLUID_H */
