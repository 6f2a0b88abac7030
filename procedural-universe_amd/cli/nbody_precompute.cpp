// Headless precompute run, the device counterpart of the reference's `-c`
// command line (src/App/NBody.cpp:52-89 -> App::RunSimulation ->
// SimulationState::RunSimulation, SimulationState.cpp:279-332): build a
// particle set, advance it with INBodySim::Update for a wall-clock time (or a
// fixed number of steps) and write the final records as a .nbody snapshot.
//
//   nbody_precompute [-t simtime] [-s timestep] [-p particles] [-f file]
//                    [--method bh|direct] [--steps K] [--out path] [--seed S]
//
// As in the reference, the step is timestep / 60 and the simulator is
// Barnes-Hut with the reference constants.  Without a file the particles come
// from the reference's StarSystem seeder (SimulationState.cpp:286-290 seeds
// with EParticleSeeder::StarSystem; src/Sim/StarSystemSeeder.cpp:17-55), drawn
// on the device by nbody_init_seeded (--seeder picks another of the
// reference's seeders).  A loaded file (-f) is used as the initial state (the
// reference overwrites it with its seeder, SimulationState.cpp:283-289, a
// defect not reproduced).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/nbody_sim.hpp"

static void usage() {
    std::printf(
        "usage: nbody_precompute [-t simtime_s] [-s timestep] [-p particles] [-f file.nbody]\n"
        "                        [--method bh|direct] [--steps K] [--out path] [--seed S]\n"
        "                        [--seeder star|random|galaxy]\n");
}

int main(int argc, char** argv) {
    int simtime = 10, particles = 4000, steps = -1;
    float timestep = 0.02f;
    std::string file, out, method = "bh", seeder = "star";
    unsigned long long seed = 0x5EED0006ULL;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "-c" || a == "--compute") continue;
        if (a == "-t" || a == "--simtime") simtime = std::atoi(next());
        else if (a == "-s" || a == "--timestep") timestep = static_cast<float>(std::atof(next()));
        else if (a == "-p" || a == "--particles") particles = std::atoi(next());
        else if (a == "-f" || a == "--file") file = next();
        else if (a == "--method") method = next();
        else if (a == "--steps") steps = std::atoi(next());
        else if (a == "--out") out = next();
        else if (a == "--seed") seed = std::strtoull(next(), nullptr, 0);
        else if (a == "--seeder") seeder = next();
        else if (a == "-h" || a == "--help") {
            usage();
            return 0;
        } else {
            usage();
            return 2;
        }
    }
    if ((method != "bh" && method != "direct") || (seeder != "star" && seeder != "random" && seeder != "galaxy")) {
        usage();
        return 2;
    }
    try {
        const ENBodySim kind = method == "bh" ? ENBodySim::BarnesHut : ENBodySim::BruteForceGPU;
        auto sim = CreateNBodySim(kind);
        std::vector<Particle> p;
        if (!file.empty()) {
            size_t n = 0;
            if (nbody_snapshot_count(file.c_str(), &n) != NBODY_OK) throw nbody_error("nbody_snapshot_count", NBODY_ERR_IO);
            p.resize(n);
            nbody_status st = nbody_snapshot_read(file.c_str(), p.data(), p.size(), &n, 1);
            if (st != NBODY_OK) throw nbody_error("nbody_snapshot_read", st);
            std::printf("[Info] Read %zu particles from file\n", n);
        } else {
            // CreateParticleSeeder(particles, EParticleSeeder::StarSystem)->Seed()
            nbody_sim* h = static_cast<DeviceNBodySim*>(sim.get())->handle();
            const int32_t kind_id = seeder == "star" ? NBODY_SEEDER_STAR_SYSTEM
                                    : seeder == "random" ? NBODY_SEEDER_RANDOM : NBODY_SEEDER_GALAXY;
            nbody_status st = nbody_init_seeded(h, kind_id, seed, static_cast<size_t>(particles), 1.0);
            if (st != NBODY_OK) throw nbody_error("nbody_init_seeded", st);
            p.assign(particles, Particle{});
            st = nbody_read(h, p.data(), p.size(), NBODY_FIELD_POSITION | NBODY_FIELD_VELOCITY | NBODY_FIELD_MASS);
            if (st != NBODY_OK) throw nbody_error("nbody_read", st);
        }
        sim->Init(p);  // the same state, now bound to the caller's records as the reference's Init does
        const float dt = timestep * (1.0f / 60.0f);  // src/App/NBody.cpp:87
        const auto t0 = std::chrono::steady_clock::now();
        auto tick = t0;
        long iterations = 0;
        while (true) {
            const auto now = std::chrono::steady_clock::now();
            if (steps >= 0 ? iterations >= steps : std::chrono::duration<double>(now - t0).count() >= simtime) break;
            if (std::chrono::duration<double>(now - tick).count() >= 1.0) {
                tick = now;
                std::printf("[Info] Running... (%ld iterations)\n", iterations);
                std::fflush(stdout);
            }
            sim->Update(dt);
            ++iterations;
        }
        if (out.empty()) {
            ::mkdir("data", 0755);
            const long long ticks =
                std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now().time_since_epoch())
                    .count();
            out = "data/" + std::to_string(ticks) + ".nbody";
        }
        nbody_status st = nbody_snapshot_write(out.c_str(), p.data(), p.size());
        if (st != NBODY_OK) throw nbody_error("nbody_snapshot_write", st);
        std::printf("[Info] %ld iterations of %zu particles (%s) written to %s\n", iterations, p.size(),
                    NBodySimGetName(kind).c_str(), out.c_str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[Error] %s\n", e.what());
        return 1;
    }
    return 0;
}
