// The in-app benchmark's counterpart (SimulationState::RunBenchmark,
// src/States/Simulation/SimulationState.cpp:334-362): for every ENBodySim,
// create the simulator, Init it with the same particle set and time ten
// synchronous Update(1.0f) frames, each writing positions, velocities and
// forces back into the records as the reference's Update does; report
// milliseconds per frame (the reference's UI table, src/UI/UI.cpp:158-169).
//
//   nbody_benchmark [-p particles] [--seeder random|galaxy|star] [--frames K]
//                   [--seed S]
//
// The particle set comes from one of the reference's seeders drawn on the
// device (nbody_init_seeded).  The library has no CPU simulator: both
// brute-force entries run the device direct sum, as CreateNBodySim maps them
// here, and the second column says so; the "Brute Force CPU" row is not a
// CPU timing.  The frame time includes the read-back of the records, which
// the reference's Update also performs.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/nbody_sim.hpp"

static void usage() {
    std::printf("usage: nbody_benchmark [-p particles] [--seeder random|galaxy|star] [--frames K] [--seed S]\n");
}

int main(int argc, char** argv) {
    int particles = 10000, frames = 10;
    std::string seeder = "galaxy";
    unsigned long long seed = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "-p" || a == "--particles") particles = std::atoi(next());
        else if (a == "--seeder") seeder = next();
        else if (a == "--frames") frames = std::atoi(next());
        else if (a == "--seed") seed = std::strtoull(next(), nullptr, 0);
        else if (a == "-h" || a == "--help") {
            usage();
            return 0;
        } else {
            usage();
            return 2;
        }
    }
    const int kind = seeder == "random" ? NBODY_SEEDER_RANDOM
                     : seeder == "star" ? NBODY_SEEDER_STAR_SYSTEM
                                        : NBODY_SEEDER_GALAXY;
    if (particles < 1 || frames < 1 || (seeder != "random" && seeder != "galaxy" && seeder != "star")) {
        usage();
        return 2;
    }
    try {
        // the shared particle set
        std::vector<Particle> base(static_cast<size_t>(particles));
        {
            nbody_config cfg;
            nbody_config_reference(&cfg, NBODY_METHOD_DIRECT);
            nbody_sim* gen = nullptr;
            nbody_status st = nbody_create(&cfg, &gen);
            if (st == NBODY_OK) st = nbody_init_seeded(gen, kind, seed, base.size(), 1.0);
            if (st == NBODY_OK)
                st = nbody_read(gen, base.data(), base.size(),
                                NBODY_FIELD_POSITION | NBODY_FIELD_VELOCITY | NBODY_FIELD_MASS);
            nbody_destroy(gen);
            if (st != NBODY_OK) throw nbody_error("seeding", st);
        }
        std::printf("# %d particles, %s seeder, %d frames of Update(1.0f)\n", particles, seeder.c_str(), frames);
        for (int t = 0; t < static_cast<int>(ENBodySim::NumSims); ++t) {
            const ENBodySim type = static_cast<ENBodySim>(t);
            std::vector<Particle> records = base;
            auto sim = CreateNBodySim(nullptr, type);
            sim->Init(records);
            const auto t0 = std::chrono::steady_clock::now();
            for (int f = 0; f < frames; ++f) sim->Update(1.0f);
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / frames;
            const char* ran = type == ENBodySim::BarnesHut ? "device Barnes-Hut" : "device direct sum";
            std::printf("%-16s %-18s %12.3f ms/frame\n", NBodySimGetName(type).c_str(), ran, ms);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[Error] %s\n", e.what());
        return 1;
    }
    return 0;
}
