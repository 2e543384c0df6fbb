// source_demo.cpp -- the reference's default scenario, demo_source (main.cpp:5-26),
// on the MI355X facade: a cubic fluid source emitting into an empty 64^3 box.
// Build: make -C xfluid_amd example.  Frame meshes go to Config::outputFild
// (empty here: no files).
#include <cstdio>
#include <cstdlib>
#include "../include/FluidSimulator.h"

int main(int argc, char** argv) {
    int frames = argc > 1 ? std::atoi(argv[1]) : 3;
    Config::outputFild = "";
    const int ni = 64, nj = 64, nk = 64;
    const double dx = 0.125;
    FluidSimulator sim(ni, nj, nk);
    sim.setCellWidth(dx);
    sim.addConstantForce(0.0, -20.0, 0.0);
    sim.addFluidSource(Vec3d(4.0 * dx, 30.0 * dx, 30.0 * dx), 2.0 * dx, 5.0 * dx, 10.0 * dx, Vec3d(10.0, 0.0, 0.0));
    sim.init();
    for (int f = 0; f < frames; ++f) sim.frameStep(1.0 / 30.0);
    std::printf("done: %lld particles\n", (long long)xf_sim_num_particles(sim.handle()));
    return 0;
}
