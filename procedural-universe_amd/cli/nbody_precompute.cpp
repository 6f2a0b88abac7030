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
// Barnes-Hut with the reference constants.  A loaded file (-f) is used as the
// initial state (the reference overwrites it with its seeder,
// SimulationState.cpp:283-289); without one, particles are uniform in
// [-500, 500)^3 with masses uniform in [1e20, 1e30), the ranges of the
// reference's RandomSeeder (src/Sim/RandomSeeder.cpp:16-19).
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
        "                        [--method bh|direct] [--steps K] [--out path] [--seed S]\n");
}

int main(int argc, char** argv) {
    int simtime = 10, particles = 4000, steps = -1;
    float timestep = 0.02f;
    std::string file, out, method = "bh";
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
        else if (a == "-h" || a == "--help") {
            usage();
            return 0;
        } else {
            usage();
            return 2;
        }
    }
    if (method != "bh" && method != "direct") {
        usage();
        return 2;
    }
    try {
        std::vector<Particle> p;
        if (!file.empty()) {
            size_t n = 0;
            if (nbody_snapshot_count(file.c_str(), &n) != NBODY_OK) throw nbody_error("nbody_snapshot_count", NBODY_ERR_IO);
            p.resize(n);
            nbody_status st = nbody_snapshot_read(file.c_str(), p.data(), p.size(), &n, 1);
            if (st != NBODY_OK) throw nbody_error("nbody_snapshot_read", st);
            std::printf("[Info] Read %zu particles from file\n", n);
        } else {
            std::vector<float> pos(3 * static_cast<size_t>(particles)), vel(3 * static_cast<size_t>(particles));
            std::vector<double> mass(particles);
            nbody_status st = nbody_ic_uniform(seed, particles, 500.0, 1e20, 1e30, pos.data(), vel.data(), mass.data());
            if (st != NBODY_OK) throw nbody_error("nbody_ic_uniform", st);
            p.assign(particles, Particle{});
            for (int i = 0; i < particles; ++i) {
                for (int a = 0; a < 3; ++a) p[i].position[a] = pos[3 * i + a];
                p[i].mass = mass[i];
            }
        }
        auto sim = CreateNBodySim(method == "bh" ? ENBodySim::BarnesHut : ENBodySim::BruteForceGPU);
        sim->Init(p);
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
                    NBodySimGetName(method == "bh" ? ENBodySim::BarnesHut : ENBodySim::BruteForceGPU).c_str(),
                    out.c_str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[Error] %s\n", e.what());
        return 1;
    }
    return 0;
}
