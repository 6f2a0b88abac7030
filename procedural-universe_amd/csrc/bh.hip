// Barnes-Hut: placeholder until the device tree lands.
#include "bh.hpp"
#include "common.hpp"
#include "sim.hpp"

namespace nbody {
struct BHState {};
nbody_status bh_create(nbody_sim*) {
    set_error("Barnes-Hut is not built yet");
    return NBODY_ERR_UNSUPPORTED;
}
void bh_destroy(BHState* b) { delete b; }
nbody_status bh_resize(BHState*, size_t, nbody_sim*) { return NBODY_OK; }
nbody_status bh_step(nbody_sim*, float, bool, hipEvent_t, hipEvent_t) { return NBODY_ERR_UNSUPPORTED; }
}  // namespace nbody
