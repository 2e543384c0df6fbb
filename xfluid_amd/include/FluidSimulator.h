// FluidSimulator.h -- drop-in C++ facade with the public interface of the
// reference's FluidSimulator (xFluid/xFluid/FluidSimulator.h:16-88), backed by
// the MI355X C ABI (include/xfluid.h).  Same method names, argument meaning
// and return types; errors surface as in the reference (printed, never thrown)
// except that construction without a HIP device aborts loudly.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/xfluid.h"
#include "config.h"

// Vec3 with the reference's operators (array3d.h:11-74).
template <class T>
struct Vec3 {
    Vec3() {}
    Vec3(T x_, T y_, T z_) : x(x_), y(y_), z(z_) {}
    Vec3<T> operator*(T v) const { return Vec3<T>(x * v, y * v, z * v); }
    T operator*(const Vec3<T>& o) const { return x * o.x + y * o.y + z * o.z; }
    Vec3<T> operator+(const Vec3<T>& o) const { return Vec3<T>(x + o.x, y + o.y, z + o.z); }
    Vec3<T> operator-(const Vec3<T>& o) const { return Vec3<T>(x - o.x, y - o.y, z - o.z); }
    T x, y, z;
};
typedef Vec3<double> Vec3d;

class FluidSimulator {
public:
    FluidSimulator() {}
    FluidSimulator(int ni, int nj, int nk) { setSize(ni, nj, nk); }
    ~FluidSimulator() { xf_sim_destroy(_sim); }
    FluidSimulator(const FluidSimulator&) = delete;
    FluidSimulator& operator=(const FluidSimulator&) = delete;

    void setSize(int ni, int nj, int nk) {
        xf_sim_destroy(_sim);
        _sim = nullptr;
        if (xf_sim_create(ni, nj, nk, &_sim) != 0) {
            std::fprintf(stderr, "FluidSimulator: %s\n", xf_last_error());
            std::abort();
        }
        if (_dx > 0) xf_sim_set_cell_width(_sim, _dx);
    }
    void setCellWidth(double dx) { _dx = dx; report(xf_sim_set_cell_width(_sim, dx)); }
    void setSolidBoundary() { report(xf_sim_set_solid_boundary(_sim)); }
    void addSpereFluid(double px, double py, double pz, double r) {
        report(xf_sim_add_sphere_fluid(_sim, px, py, pz, r));
    }
    void addSpereFluid(Vec3d p, double r) { addSpereFluid(p.x, p.y, p.z, r); }
    void addCubicFluid(Vec3d lu, Vec3d rd) {
        double a[3] = {lu.x, lu.y, lu.z}, b[3] = {rd.x, rd.y, rd.z};
        report(xf_sim_add_cubic_fluid(_sim, a, b));
    }
    void addConstantForce(const Vec3d& f) { addConstantForce(f.x, f.y, f.z); }
    void addConstantForce(double fx, double fy, double fz) { report(xf_sim_add_constant_force(_sim, fx, fy, fz)); }
    // FluidSimulator.cpp:152-166 (jitter stream seeded by setSeed)
    void addFluidSource(const Vec3d& c, double wx, double wy, double wz, const Vec3d& v) {
        double cc[3] = {c.x, c.y, c.z}, vv[3] = {v.x, v.y, v.z};
        report(xf_sim_add_fluid_source(_sim, cc, wx, wy, wz, vv));
    }
    void initParticles() { report(xf_sim_seed_particles(_sim, XF_SEED_REFERENCE, _seed)); }
    void init() { report(xf_sim_init(_sim)); }
    // frameStep: CFL substeps; the last one writes <Config::outputFild><frame>.ply
    void frameStep(double dt) {
        report(xf_sim_set_mesh_output(_sim, Config::outputFild.c_str()));
        int rc = xf_sim_frame_step(_sim, dt);
        report(rc);
        std::printf("Frame   :   %d\n", xf_sim_frame_count(_sim) - 1);
    }
    void timeStep(double dt, bool outputMesh) {
        report(xf_sim_set_mesh_output(_sim, Config::outputFild.c_str()));
        int rc = xf_sim_time_step(_sim, dt, outputMesh ? 1 : 0);
        report(rc);
    }

    // -- beyond the reference interface --
    xf_sim* handle() { return _sim; }
    void setParams(const xf_params& p) { report(xf_sim_set_params(_sim, &p)); }
    // seed of initParticles' jitter and of the fluid sources' stream
    void setSeed(uint64_t s) {
        _seed = s;
        report(xf_sim_set_source_seed(_sim, s));
    }

private:
    void report(int rc) {
        if (rc < 0) std::fprintf(stderr, "xfluid error %d: %s\n", rc, xf_last_error());
        else if (rc > 0) std::printf("pressure solve success:  0\n");
    }
    xf_sim* _sim = nullptr;
    double _dx = 0.0;
    uint64_t _seed = 0;
};
