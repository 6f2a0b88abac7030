// Uses the reference-shaped C++ adapter (include/nbody_sim.hpp) the way
// SimulationState does: CreateNBodySim, Init(vector), Update(dt), read the
// records.  Prints one line per particle of the first few after `steps`.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "nbody_sim.hpp"

int main(int argc, char** argv) {
    const int kind = argc > 1 ? std::atoi(argv[1]) : 0;  // ENBodySim
    const int n = argc > 2 ? std::atoi(argv[2]) : 512;
    const int steps = argc > 3 ? std::atoi(argv[3]) : 3;
    std::vector<float> pos(3 * n), vel(3 * n);
    std::vector<double> mass(n);
    if (nbody_ic_uniform(11, n, 500.0, 1e20, 1e30, pos.data(), vel.data(), mass.data()) != NBODY_OK) return 2;
    std::vector<Particle> particles(n);
    for (int i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) particles[i].position[a] = pos[3 * i + a];
        particles[i].mass = mass[i];
    }
    try {
        auto sim = CreateNBodySim(nullptr, static_cast<ENBodySim>(kind));  // SimulationState.cpp:286
        sim->Init(particles);
        for (int s = 0; s < steps; ++s) sim->Update(0.02f / 60.0f);
        // the octree the last Barnes-Hut step used, as a renderer would take it
        const auto cells = static_cast<DeviceNBodySim*>(sim.get())->Cells();
        std::fprintf(stderr, "cells %zu root %u\n", cells.size(), cells.empty() ? 0u : cells[0].count);
    } catch (const nbody_error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    std::printf("%s\n", NBodySimGetName(static_cast<ENBodySim>(kind)).c_str());
    for (int i = 0; i < n; ++i)
        std::printf("%.9g %.9g %.9g %.17g %.17g %.17g %.17g\n", particles[i].position[0], particles[i].position[1],
                    particles[i].position[2], particles[i].velocity[0], particles[i].velocity[1],
                    particles[i].velocity[2], particles[i].forces[0]);
    return 0;
}
