// FP64 energy / momentum diagnostics of a device-resident state.
#pragma once

#include <hip/hip_runtime.h>

namespace nbody {

// Number of partial records each kernel writes (local: 8 doubles per block,
// potential: 1 double per block).
int diag_local_blocks(int n);
int diag_potential_blocks(int n);

// pos/vel: n float4 records; mass: FP64 masses; gm: FP64 G*m.
hipError_t diag_launch(const float4* pos, const float4* vel, const double* mass, const double* gm, int n,
                       int law, double soft, double* local_out, double* pot_out, hipStream_t stream);

}  // namespace nbody
