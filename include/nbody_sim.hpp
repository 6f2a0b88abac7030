// nbody_sim.hpp - header-only C++ adapter with the reference's simulator
// interface (src/Sim/INBodySim.hpp:11-28) over the C ABI of nbody.h.
//
//     std::vector<Particle> particles = ...;              // 104-byte records
//     auto sim = CreateNBodySim(ENBodySim::BarnesHut);   // reference constants
//     sim->Init(particles);                              // INBodySim::Init
//     sim->Update(dt);                                   // one step, in place
//
// Semantics kept from the reference:
//  * Init adopts the caller's vector; Update advances one step synchronously
//    and writes Position, Velocity and Forces back into it
//    (SimulationState.cpp:52-60 reads them right after Update);
//  * the brute-force simulators leave Forces zero after Update
//    (BruteForceCPU.cpp:72, Update.csh:35), Barnes-Hut leaves the last forces
//    (BarnesHut.cpp:75,108);
//  * CreateNBodySim uses the reference's physical constants unless a
//    configuration is given (Physics.hpp:9-16, Octree.cpp:5, BarnesHut.cpp:14-19).
// Where the reference ignores errors, the adapter throws nbody_error.
#ifndef NBODY_SIM_HPP_
#define NBODY_SIM_HPP_

#include <algorithm>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "nbody.h"

using Particle = nbody_particle;

// src/Sim/INBodySim.hpp:11-17
enum class ENBodySim { BruteForceCPU, BruteForceGPU, BarnesHut, NumSims };

class nbody_error : public std::runtime_error {
public:
    nbody_error(const char* what, nbody_status status)
        : std::runtime_error(std::string(what) + ": " + nbody_status_string(status) + ": " + nbody_last_error()),
          status_(status) {}
    nbody_status status() const { return status_; }

private:
    nbody_status status_;
};

// src/Sim/INBodySim.hpp:19-25
class INBodySim {
public:
    virtual ~INBodySim() = default;
    virtual void Init(std::vector<Particle>& particles) = 0;
    virtual void Update(float dt) = 0;
    // The reference draws the octree's cubes here (BarnesHut.cpp:98-101); the
    // device simulator draws nothing itself: DeviceNBodySim::Cells() gives the
    // cubes to a renderer.
    virtual void RenderDebug(const float* /*view4x4*/, const float* /*proj4x4*/) {}
};

// The device simulator behind INBodySim: one nbody_sim handle.
class DeviceNBodySim : public INBodySim {
public:
    explicit DeviceNBodySim(const nbody_config& cfg)
        : zero_forces_(cfg.method == NBODY_METHOD_DIRECT), barnes_hut_(cfg.method == NBODY_METHOD_BARNES_HUT) {
        check("nbody_create", nbody_create(&cfg, &sim_));
    }
    ~DeviceNBodySim() override { nbody_destroy(sim_); }
    DeviceNBodySim(const DeviceNBodySim&) = delete;
    DeviceNBodySim& operator=(const DeviceNBodySim&) = delete;

    void Init(std::vector<Particle>& particles) override {
        check("nbody_init", nbody_init(sim_, particles.data(), particles.size()));
        particles_ = &particles;
    }

    void Update(float dt) override {
        if (!particles_) throw std::logic_error("Update before Init");
        check("nbody_step", nbody_step(sim_, dt, 1));
        check("nbody_read", nbody_read(sim_, particles_->data(), particles_->size(),
                                       NBODY_FIELD_POSITION | NBODY_FIELD_VELOCITY | NBODY_FIELD_FORCES));
        if (zero_forces_)
            for (auto& p : *particles_) p.forces[0] = p.forces[1] = p.forces[2] = 0.0;
    }

    nbody_sim* handle() { return sim_; }

    // The reference's BHThetaChanged handler (BarnesHut.cpp:29-31) sets the
    // global Octree::Theta from the UI's slider; here the handle's theta.
    // Register it where the reference registers its handler:
    //     EventStream::Register(EEvent::BHThetaChanged, [&](const EventData& d) {
    //         sim.SetTheta(EventValue<FloatEventData>(d)); });
    void SetTheta(double theta) { check("nbody_set_theta", nbody_set_theta(sim_, theta)); }

    // The octree cells of the last Barnes-Hut step (cube corner and side,
    // centre of mass, mass, level, particle range): what the reference's
    // Octree::RenderDebug draws (Octree.cpp:147-175).  Empty for the direct sum.
    std::vector<nbody_bh_cell> Cells() {
        std::vector<nbody_bh_cell> cells;
        if (!barnes_hut_) return cells;
        size_t count = 0;
        check("nbody_bh_cells", nbody_bh_cells(sim_, nullptr, 0, &count));
        cells.resize(count);
        check("nbody_bh_cells", nbody_bh_cells(sim_, cells.data(), cells.size(), &count));
        cells.resize(std::min(count, cells.size()));
        return cells;
    }

private:
    static void check(const char* what, nbody_status st) {
        if (st != NBODY_OK) throw nbody_error(what, st);
    }
    nbody_sim* sim_ = nullptr;
    std::vector<Particle>* particles_ = nullptr;
    bool zero_forces_;
    bool barnes_hut_;
};

// src/Sim/INBodySim.cpp:7-27.  Both brute-force simulators run the device
// direct sum; `cfg` (optional) replaces the reference constants.
inline std::unique_ptr<INBodySim> CreateNBodySim(ENBodySim type, const nbody_config* cfg = nullptr) {
    const int32_t method = type == ENBodySim::BarnesHut ? NBODY_METHOD_BARNES_HUT : NBODY_METHOD_DIRECT;
    if (type == ENBodySim::NumSims) return nullptr;
    nbody_config c;
    if (cfg) {
        c = *cfg;
        c.method = method;
    } else {
        nbody_config_reference(&c, method);
    }
    return std::unique_ptr<INBodySim>(new DeviceNBodySim(c));
}

// The reference's factory signature (src/Sim/INBodySim.hpp:27): the Direct3D
// device context comes first and is ignored here, so the reference's calls
// (SimulationState.cpp:107,169,286,343) compile unchanged.
template <class DeviceContext>
inline std::unique_ptr<INBodySim> CreateNBodySim(DeviceContext* /*context*/, ENBodySim type) {
    return CreateNBodySim(type);
}
inline std::unique_ptr<INBodySim> CreateNBodySim(std::nullptr_t, ENBodySim type) { return CreateNBodySim(type); }

// src/Sim/INBodySim.cpp:29-39
inline std::string NBodySimGetName(ENBodySim type) {
    switch (type) {
        case ENBodySim::BruteForceCPU: return "Brute Force CPU";
        case ENBodySim::BruteForceGPU: return "Brute Force GPU";
        case ENBodySim::BarnesHut: return "Barnes-Hut";
        default: return "Unknown";
    }
}

#endif  // NBODY_SIM_HPP_
