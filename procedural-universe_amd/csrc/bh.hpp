// Barnes-Hut pipeline on the device: bounds, Morton keys, radix sort,
// binary radix tree, octree cells, FP64 mass/centre-of-mass, traversal.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

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
// Recompute what depends on the handle's theta (a fixed root's opening
// thresholds; a dynamic root recomputes them every step).  Enqueued on the
// handle's stream.
nbody_status bh_set_theta(nbody_sim* s);
// Check a fixed root (corner on the root's level-21 grid); out_params may be null.
nbody_status bh_fixed_root(const nbody_config* c, void* out_params);
// Counters of the steps since the last reset.
struct BHCounters {
    unsigned long long interactions;  // lanes x list entries they take part in
    unsigned long long entries;       // list entries (sources evaluated by a whole wave)
    unsigned long long opened;        // cells opened (stack pops)
    unsigned outside;                 // particles clamped into a fixed root
    unsigned overflow;                // nonzero if a traversal stack overflowed
};
nbody_status bh_reset_counters(BHState* b, hipStream_t st);
nbody_status bh_read_counters(BHState* b, hipStream_t st, BHCounters* out);
// Sorted keys and particle order of the last tree build.
nbody_status bh_export_order(nbody_sim* s, uint64_t* keys, uint32_t* order, size_t n);
// Octree cells of the last tree build.
nbody_status bh_export_cells(nbody_sim* s, nbody_bh_cell* out, size_t capacity, size_t* count);

}  // namespace nbody
