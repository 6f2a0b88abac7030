// Barnes-Hut pipeline on the device: bounds, Morton keys, radix sort,
// binary radix tree, octree cells, FP64 mass/centre-of-mass, traversal.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "../../include/nbody.h"

struct nbody_sim;

namespace nbody {

struct BHState;

nbody_status bh_create(nbody_sim* s);
void bh_destroy(BHState* b);
// (Re)size the tree buffers for n particles.
nbody_status bh_resize(BHState* b, size_t n, nbody_sim* s);
// One force evaluation (kick: a full step) of the handle's current state.
nbody_status bh_step(nbody_sim* s, float dt, bool kick, hipEvent_t ev_force_begin, hipEvent_t ev_force_end);

}  // namespace nbody
