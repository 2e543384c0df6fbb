// ref_harness.cpp -- C entry points that drive the reference's own classes.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) together
// with the reference's translation units compiled where they lie under the
// reference tree, into oracle/_ref/libxfref.so.  Nothing of the reference is
// copied: this file #includes FluidSimulator.cpp from the reference tree.
//
// Two build accommodations, both local to this translation unit:
//  * `#define private public` so the harness can call the stage methods of
//    FluidSimulator::timeStep (FluidSimulator.cpp:265-335) one at a time and
//    read the members that hold the fields.
//  * FluidSimulator.cpp:179 binds the temporary GridIndex3d(i,j,k) to a
//    non-const reference, which only MSVC accepts.  While that file is
//    included, `GridIndex3d(...)` expressions are routed through a function
//    returning an lvalue copy.  No arithmetic is affected.
#include <vector>
#include <string>
#include <iostream>
#include <sstream>
#include <fstream>
#include <algorithm>
#include <queue>
#include <map>
#include <climits>
#include <limits>
#include <cmath>
#include <cassert>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <sys/time.h>

#include "gridindex.h"
static GridIndex3d xr_gi_slot;
static inline GridIndex3d& xr_gi_lvalue(const GridIndex3d& g) { xr_gi_slot = g; return xr_gi_slot; }

#define private public
#define protected public
#include "FluidSimulator.h"
#define GridIndex3d(...) xr_gi_lvalue(GridIndex3d{__VA_ARGS__})
#include "FluidSimulator.cpp"
#undef GridIndex3d
#undef private
#undef protected

namespace {
struct Quiet {  // the reference prints per-stage timings; keep test output clean
    std::streambuf* old;
    std::ostringstream sink;
    Quiet() { old = std::cout.rdbuf(sink.rdbuf()); }
    ~Quiet() { std::cout.rdbuf(old); }
};

struct Ref {
    FluidSimulator sim;
    Array3d<double> pressure;
    std::vector<double> rhs, micinv;
    int iters = 0;
    double res = 0;
    int status = 0;
    double blob = 0;
    std::vector<Vec3d> mverts;  // last mesh (copied out of a TriangleMesh)
    std::vector<Triangle> mtris;
    Ref(int ni, int nj, int nk) : sim(ni, nj, nk) {}
};

enum { F_U, F_V, F_W, F_US, F_VS, F_WS, F_MAT, F_PHI, F_P, F_RHS, F_INV };

template <class T>
void to_host(Array3d<T>& a, void* dst) {
    std::memcpy(dst, &a(0), sizeof(T) * (size_t)a.sizei() * a.sizej() * a.sizek());
}
}  // namespace

extern "C" {

// Directory prefix of the reference's frame meshes (config.h:4-7, used at
// FluidSimulator.cpp:298-301); the harness points it at a scratch directory.
void xr_set_output(const char* prefix) { Config::outputFild = prefix; }

void* xr_create(int ni, int nj, int nk, double dx) {
    Ref* r = new Ref(ni, nj, nk);
    r->sim.setCellWidth(dx);
    return r;
}
void xr_destroy(void* h) { delete (Ref*)h; }

// Member values of FluidSimulator.h:187-224 and the PCG parameters
// (pcg_solver.h:234-241; tau/sigma there are not used by the factorisation).
void xr_set_params(void* h, double flip, double cfl, double max_dt, double pcg_tol, int pcg_max_iter,
                   double blob_radius, double iso) {
    Ref* r = (Ref*)h;
    r->sim._ratioFLIP = flip;
    r->sim._CFLnumber = cfl;
    r->sim._maxStepTime = max_dt;
    r->sim._threshold = iso;
    r->sim._pressureSolver.solver.set_solver_parameters(pcg_tol, pcg_max_iter);
    r->blob = blob_radius;
}
void xr_set_solid_boundary(void* h) { ((Ref*)h)->sim.setSolidBoundary(); }
void xr_add_cubic_fluid(void* h, const double* lu, const double* rd) {
    ((Ref*)h)->sim.addCubicFluid(Vec3d(lu[0], lu[1], lu[2]), Vec3d(rd[0], rd[1], rd[2]));
}
void xr_add_sphere_fluid(void* h, double x, double y, double z, double rad) { ((Ref*)h)->sim.addSpereFluid(x, y, z, rad); }
void xr_add_constant_force(void* h, double fx, double fy, double fz) { ((Ref*)h)->sim.addConstantForce(fx, fy, fz); }

// FluidSimulator::addFluidSource (FluidSimulator.cpp:152-166)
void xr_add_fluid_source(void* h, const double* c, double wx, double wy, double wz, const double* v) {
    ((Ref*)h)->sim.addFluidSource(Vec3d(c[0], c[1], c[2]), wx, wy, wz, Vec3d(v[0], v[1], v[2]));
}
// _applyFluidSource (FluidSimulator.cpp:1087-1093) with rand() restarted at srand(seed)
void xr_apply_sources(void* h, double dt, unsigned seed) {
    srand(seed);
    ((Ref*)h)->sim._applyFluidSource(dt);
}
// FluidSimulator::initParticles (FluidSimulator.cpp:168-184, 347-370) with
// rand() restarted at srand(seed)
void xr_init_particles(void* h, unsigned seed) {
    srand(seed);
    ((Ref*)h)->sim.initParticles();
}
// the first n values rand() returns after srand(seed), and RAND_MAX
int xr_rand(unsigned seed, int64_t n, uint32_t* out) {
    srand(seed);
    for (int64_t i = 0; i < n; ++i) out[i] = (uint32_t)rand();
    return RAND_MAX;
}

// FluidSimulator::init (FluidSimulator.cpp:186-222) minus initParticles, whose
// rand() jitter is replaced by injected particles (xr_set_particles).
void xr_init(void* h) {
    Ref* r = (Ref*)h;
    FluidSimulator& s = r->sim;
    s.setSolidBoundary();
    s._dis_field.setParticleVector(&s._particles);
    s._dis_field.setSubDivisionLevel(s._subDivisionLevel);
    s._dis_field.setParticleRadius(r->blob > 0 ? r->blob : s._particleRadius * 3.0);
    s._dis_field.setDx(s._dx);
    s._dis_field.setThreshold(s._threshold);
    s._dis_field.initField();
    s._pressureSolver.setDensity(s._density);
    s._pressureSolver.setDisField(&s._dis_field);
    s._pressureSolver.setDx(s._dx);
    s._pressureSolver.setMaterial(&s._material);
    s._pressureSolver.setSize(s._ni, s._nj, s._nk);
    s._pressureSolver.setVelocityField(&s._u, &s._v, &s._w);
    s._usaved = s._u; s._vsaved = s._v; s._wsaved = s._w;
}

void xr_set_particles(void* h, int64_t n, const double* px, const double* py, const double* pz, const double* vx,
                      const double* vy, const double* vz) {
    FluidSimulator& s = ((Ref*)h)->sim;
    s._particles.clear();
    for (int64_t i = 0; i < n; ++i) s._particles.push_back(Particle(Vec3d(px[i], py[i], pz[i]), Vec3d(vx[i], vy[i], vz[i])));
}
int64_t xr_num_particles(void* h) { return (int64_t)((Ref*)h)->sim._particles.size(); }
void xr_get_particles(void* h, double* px, double* py, double* pz, double* vx, double* vy, double* vz) {
    FluidSimulator& s = ((Ref*)h)->sim;
    for (size_t i = 0; i < s._particles.size(); ++i) {
        px[i] = s._particles[i].p.x; py[i] = s._particles[i].p.y; pz[i] = s._particles[i].p.z;
        vx[i] = s._particles[i].v.x; vy[i] = s._particles[i].v.y; vz[i] = s._particles[i].v.z;
    }
}

int xr_get_field(void* h, int which, void* dst) {
    Ref* r = (Ref*)h;
    FluidSimulator& s = r->sim;
    switch (which) {
        case F_U: to_host(s._u, dst); return 0;
        case F_V: to_host(s._v, dst); return 0;
        case F_W: to_host(s._w, dst); return 0;
        case F_US: to_host(s._usaved, dst); return 0;
        case F_VS: to_host(s._vsaved, dst); return 0;
        case F_WS: to_host(s._wsaved, dst); return 0;
        case F_MAT: to_host(s._material, dst); return 0;
        case F_PHI: to_host(s._dis_field._phi, dst); return 0;
        case F_P: to_host(r->pressure, dst); return 0;
        case F_RHS: std::memcpy(dst, r->rhs.data(), r->rhs.size() * 8); return 0;
        case F_INV: std::memcpy(dst, r->micinv.data(), r->micinv.size() * 8); return 0;
    }
    return -1;
}
int xr_set_field(void* h, int which, const void* src) {
    FluidSimulator& s = ((Ref*)h)->sim;
    Array3d<double>* a = nullptr;
    switch (which) {
        case F_U: a = &s._u; break;
        case F_V: a = &s._v; break;
        case F_W: a = &s._w; break;
        case F_US: a = &s._usaved; break;
        case F_VS: a = &s._vsaved; break;
        case F_WS: a = &s._wsaved; break;
        case F_MAT:
            std::memcpy(&s._material(0), src, (size_t)s._ni * s._nj * s._nk);
            return 0;
        default: return -1;
    }
    std::memcpy(&(*a)(0), src, sizeof(double) * (size_t)a->sizei() * a->sizej() * a->sizek());
    return 0;
}

// One stage of timeStep, by the reference's own methods.
int xr_stage(void* h, int stage, double dt) {
    Ref* r = (Ref*)h;
    FluidSimulator& s = r->sim;
    Quiet q;
    switch (stage) {
        case 0: {  // density + surface list (material written there is discarded by the step)
            Array3d<Material> keep = s._material;
            s._dis_field.updateFeild();
            s._dis_field.updateSurfaceInfo(s._material);
            s._material = keep;
            return 0;
        }
        case 1: s._updateMaterial(); return 0;
        case 2: s._saveVelocity(); return 0;
        case 3: s._applyForces(dt); return 0;
        case 4: {  // VariationalPressureSolver::solve (variational_pressuresolver.cpp:50-101) pieces
            s._pressureSolver.setDt(dt);
            SparseMatrix<double> A;
            std::vector<double> b;
            s._pressureSolver._calculate_coefficient(A, b);
            int n = s._ni * s._nj * s._nk;
            std::vector<double> p(n, 0);
            double res = 0;
            int it = 0;
            bool ok = s._pressureSolver.solver.solve(A, b, p, res, it);
            r->rhs = b;
            r->micinv = s._pressureSolver.solver.ic_factor.invdiag;
            r->micinv.resize(n, 0.0);
            r->iters = it; r->res = res; r->status = ok ? 0 : 1;
            r->pressure.resize(s._ni, s._nj, s._nk);
            for (int c = 0; c < n; ++c) r->pressure(c) = p[c];
            return r->status;
        }
        case 5: s._pressureSolver.applyPressure(r->pressure); return 0;
        case 6: s._expolateVelocity(); return 0;
        case 7: {  // G2P + advection: _advectParticles also runs P2G, whose grid result is put back
            Array3d<double> u = s._u, v = s._v, w = s._w;
            s._advectParticles(dt);
            s._u = u; s._v = v; s._w = w;
            return 0;
        }
        case 8: s._updateVelocityByParticles(); return 0;
    }
    return -1;
}
void xr_last_solve(void* h, int* iters, double* res, int* status) {
    Ref* r = (Ref*)h;
    *iters = r->iters; *res = r->res; *status = r->status;
}

int xr_time_step(void* h, double dt) {
    Quiet q;
    ((Ref*)h)->sim.timeStep(dt, false);
    return 0;
}
int xr_frame_step(void* h, double dt) {
    Quiet q;
    ((Ref*)h)->sim.frameStep(dt);
    return 0;
}

int64_t xr_surface_cells(void* h, int32_t* cells, int64_t cap) {
    FluidSimulator& s = ((Ref*)h)->sim;
    const std::vector<GridIndex3d>& sc = s._dis_field.getSubSurfaceCell();
    int d = s._subDivisionLevel;
    int64_t n = (int64_t)sc.size();
    for (int64_t i = 0; i < n && i < cap; ++i)
        cells[i] = (int32_t)(sc[i].i + (int64_t)(d * s._ni) * (sc[i].j + (int64_t)(d * s._nj) * sc[i].k));
    return n;
}

// FluidMesher::generateMesh (fluidmesher.cpp:364-392) on the current density
// field and surface list.  smooth = 0 runs the same per-cell loop without the
// final smoothing pass (for checking vertex creation and triangles alone).
int64_t xr_mesh(void* h, int smooth, int64_t* ntris) {
    Ref* r = (Ref*)h;
    FluidSimulator& s = r->sim;
    Quiet q;
    TriangleMesh mesh;
    if (smooth) {
        s._mesher.generateMesh(s._dis_field, mesh);
    } else {
        FluidMesher& m = s._mesher;
        m._cur_field = &s._dis_field;
        Vec3i dim = s._dis_field.subSize();
        Array3d<int> eu(dim.x + 1, dim.y + 1, dim.z + 1), ev(dim.x + 1, dim.y + 1, dim.z + 1),
            ew(dim.x + 1, dim.y + 1, dim.z + 1);
        eu.fill(-1); ev.fill(-1); ew.fill(-1);
        const std::vector<GridIndex3d>& sc = s._dis_field.getSubSurfaceCell();
        for (size_t i = 0; i < sc.size(); ++i) m._generateSurfaceCellMesh(sc[i], mesh, eu, ev, ew);
        m._cur_field = nullptr;
    }
    r->mverts = mesh.vertices;
    r->mtris = mesh.triangles;
    *ntris = (int64_t)r->mtris.size();
    return (int64_t)r->mverts.size();
}
void xr_get_mesh(void* h, double* verts, int32_t* tris) {
    Ref* r = (Ref*)h;
    for (size_t i = 0; i < r->mverts.size(); ++i) {
        verts[3 * i] = r->mverts[i].x;
        verts[3 * i + 1] = r->mverts[i].y;
        verts[3 * i + 2] = r->mverts[i].z;
    }
    for (size_t t = 0; t < r->mtris.size(); ++t)
        for (int x = 0; x < 3; ++x) tris[3 * t + x] = r->mtris[t].tri[x];
}
// TriangleMesh::writeMeshToPLY (trianglemesh.cpp:95-252) of the last mesh
void xr_write_ply(void* h, const char* path) {
    Ref* r = (Ref*)h;
    TriangleMesh mesh;
    mesh.vertices = r->mverts;
    mesh.triangles = r->mtris;
    mesh.writeMeshToPLY(path);
}

// Standalone solver pieces on an assembled matrix, for pinning the restated
// MIC(0) and sweeps: returns A in the 7-plane layout, b, the MIC(0) pivots and
// one preconditioner application z = M^-1 r (pcg_solver.h:85-220, 300-309).
int xr_pressure_pieces(void* h, double* A7, double* b_out, double* inv_out, const double* r_in, double* z_out) {
    FluidSimulator& s = ((Ref*)h)->sim;
    SparseMatrix<double> A;
    std::vector<double> b;
    s._pressureSolver._calculate_coefficient(A, b);
    const long N = (long)s._ni * s._nj * s._nk, nn = (long)s._ni * s._nj;
    const long off[7] = {-nn, -(long)s._ni, -1, 0, 1, (long)s._ni, nn};
    std::fill(A7, A7 + 7 * N, 0.0);
    for (long c = 0; c < N; ++c)
        for (size_t e = 0; e < A.index[c].size(); ++e) {
            long col = A.index[c][e];
            int slot = -1;
            for (int t = 0; t < 7; ++t) if (c + off[t] == col) slot = t;
            if (slot < 0) return -2;
            A7[slot * N + c] = A.value[c][e];
        }
    std::copy(b.begin(), b.end(), b_out);
    SparseColumnLowerFactor<double> f;
    factor_modified_incomplete_cholesky0(A, f);
    std::copy(f.invdiag.begin(), f.invdiag.end(), inv_out);
    std::vector<double> rr(r_in, r_in + N), zz(N);
    solve_lower(f, rr, zz);
    solve_lower_transpose_in_place(f, zz);
    std::copy(zz.begin(), zz.end(), z_out);
    return 0;
}

// NormalPressureSolver (pressuresolver.cpp:50-347) on the simulator's material
// and velocities: its matrix in 7 slots, its right-hand side, the pressure of
// its solve (default PCGSolver parameters, pcg_solver.h:228-231) and, when the
// solve succeeds, the velocities after its applyPressure (left in the
// simulator's u, v, w).  Returns 0 on success, 1 when the solve fails.
int xr_normal_pieces(void* h, double* A7, double* b_out, double* p_out) {
    FluidSimulator& s = ((Ref*)h)->sim;
    NormalPressureSolver ns;
    ns.setSize(s._ni, s._nj, s._nk);
    ns.setVelocityField(&s._u, &s._v, &s._w);
    ns.setMaterial(&s._material);
    ns.setDx(s._dx);
    SparseMatrix<double> A;
    std::vector<double> b;
    ns._calculate_coefficient(A, b);
    const long N = (long)s._ni * s._nj * s._nk, nn = (long)s._ni * s._nj;
    const long off[7] = {-nn, -(long)s._ni, -1, 0, 1, (long)s._ni, nn};
    std::fill(A7, A7 + 7 * N, 0.0);
    for (long c = 0; c < N; ++c)
        for (size_t e = 0; e < A.index[c].size(); ++e) {
            long col = A.index[c][e];
            int slot = -1;
            for (int t = 0; t < 7; ++t) if (c + off[t] == col) slot = t;
            if (slot < 0) return -2;
            A7[slot * N + c] = A.value[c][e];
        }
    std::copy(b.begin(), b.end(), b_out);
    Quiet q;
    Array3d<double> pressure;
    if (!ns.solve(pressure)) return 1;
    for (int k = 0; k < s._nk; ++k)
        for (int j = 0; j < s._nj; ++j)
            for (int i = 0; i < s._ni; ++i) p_out[i + s._ni * (j + (long)s._nj * k)] = pressure(i, j, k);
    ns.applyPressure(pressure);
    return 0;
}

}  // extern "C"
