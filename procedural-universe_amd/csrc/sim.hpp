// The simulation handle behind the C ABI's opaque `nbody_sim`.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/nbody.h"
#include "direct.hpp"

namespace nbody {
struct BHState;
nbody_status validate_config(const nbody_config* c);
}  // namespace nbody

struct nbody_sim {
    nbody_config cfg{};
    hipStream_t stream = nullptr;

    size_t n = 0;         // real particles
    size_t n_pad = 0;     // padded to kTile
    size_t capacity = 0;  // allocated records per buffer
    bool initialised = false;

    // Device state, structure of arrays.
    float4* pos[2] = {nullptr, nullptr};  // (x, y, z, G*m), alternating per step
    int cur = 0;                          // index of the current positions
    float4* vel = nullptr;                // (vx, vy, vz, 0)
    float4* acc = nullptr;                // last accelerations (ax, ay, az, 0)
    float4* part = nullptr;               // direct: partial sums [split][n_pad]
    size_t part_elems = 0;
    double* mass_dev = nullptr;           // FP64 masses (diagnostics)
    double* gm_dev = nullptr;             // FP64 G*m (diagnostics)
    double* diag_buf = nullptr;
    double* ic_scratch = nullptr;         // device initial conditions: recentring sums
    size_t diag_elems = 0;
    std::vector<double> mass_host;        // FP64 masses (FORCES read-back = m * a)
    double max_gm = 0.0;                  // largest |G*m| (the direct sum's mass scale)
    nbody::DirectScale dscale{1.f, 1.f, 0.f, 1.f};  // direct sum's scaled units
    bool acc_valid = false;

    nbody::DirectShape shape{4, 256, 1};
    nbody::BHState* bh = nullptr;

    // Timing of the last nbody_step call.
    bool timing = false;
    std::vector<hipEvent_t> events;        // [0] call begin, [1] call end, then per-step pairs
    std::vector<hipEvent_t> tree_ms_pending;  // Barnes-Hut tree-build brackets
    int timed_steps = 0;
    uint64_t step_interactions = 0;
    nbody_timing last{};

    // nbody_set_graph: an nbody_step sequence is captured once and replayed,
    // one graph per parity of the current positions.
    bool use_graph = false;
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        int32_t steps = -1;
        float dt = 0.f;
    };
    StepGraph graphs[2];
};
