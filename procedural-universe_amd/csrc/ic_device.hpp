// Device-side initial conditions (counter-based stream shared with ic.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace nbody {

// Doubles of scratch ic_device_plummer needs.
size_t ic_device_scratch_doubles();

// Plummer sphere of n equal masses written to pos[0..n) as (x, y, z, G m) and
// vel[0..n), recentred; mass[i] = M / n and gm[i] = G M / n where those are not
// null; pos_copy (if not null) receives the positions too (the second buffer
// of a ping-pong pair).
hipError_t ic_device_plummer(unsigned long long seed, int n, double a, double total_mass, double G, float4* pos,
                             float4* vel, double* mass, double* gm, float4* pos_copy, double* scratch,
                             hipStream_t stream);

}  // namespace nbody
