// Host initial-condition generators (counter-based random stream).
#pragma once

#include <cstddef>
#include <cstdint>

namespace nbody {

void ic_plummer(uint64_t seed, size_t n, double a, double total_mass, double G, float* pos, float* vel,
                double* mass, double x_offset = 0.0);
void ic_disc_bulge(uint64_t seed, size_t n, double Rd, double z0, double bulge_fraction, double ab,
                   double total_mass, double G, float* pos, float* vel, double* mass);
void ic_uniform(uint64_t seed, size_t n, double half, double mass_lo, double mass_hi, float* pos, float* vel,
                double* mass);
void ic_recentre(size_t n, float* pos, float* vel, const double* mass);
void ic_disc_encounter(uint64_t seed, size_t n, double Rd, double z0, double separation, double pericentre,
                       double tilt_deg, double G, float* pos, float* vel, double* mass);

}  // namespace nbody
