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

// The reference's particle seeders (src/Sim/IParticleSeeder.hpp:12-17,
// RandomSeeder.cpp, GalaxySeeder.cpp, StarSystemSeeder.cpp) as distributions
// drawn on the device: kind 0 random cube, 1 galaxy (two spiral arms and a
// disc, randomly oriented), 2 star system.  Writes pos (x, y, z, G m), vel,
// mass and gm for i < n (pos_copy too when not null).
struct SeederTable {
    int loops;             // spiral-arm steps (61: a float angle += 0.1f below 6)
    float angle[64];       // the reference's float-accumulated arm angles and radii
    float radius[64];
    float rot[9];          // row-vector rotation of the galaxy (yaw, pitch, roll)
};
hipError_t ic_device_seeded(int kind, unsigned long long seed, int n, float scale, double G, const SeederTable& T,
                            float4* pos, float4* vel, double* mass, double* gm, float4* pos_copy, hipStream_t stream);
// Host side of the table for a galaxy of this seed.
void ic_seeder_table(unsigned long long seed, SeederTable* T);

}  // namespace nbody
