// dam_break.cpp -- the reference's demo usage (main.cpp:28-49) on the MI355X
// facade: a dam break stepped frame by frame.  Build: make -C xfluid_amd example
#include <cstdio>
#include "../include/FluidSimulator.h"

int main(int argc, char** argv) {
    int frames = argc > 1 ? std::atoi(argv[1]) : 3;
    FluidSimulator sim(64, 32, 32);
    sim.setCellWidth(0.125);
    sim.setSolidBoundary();
    sim.addCubicFluid(Vec3d(0.0, 0.0, 0.0), Vec3d(2.8, 2.4, 4.0));
    sim.addConstantForce(0.0, -9.81, 0.0);
    sim.init();
    for (int f = 0; f < frames; ++f) sim.frameStep(1.0 / 30.0);
    std::printf("done: %lld particles\n", (long long)xf_sim_num_particles(sim.handle()));
    return 0;
}
