// The reference's snapshot file: a raw array of 104-byte Particle records,
// written by the headless precompute run (SimulationState.cpp:317-331) and
// read back, recentred on the centre of mass, by InitParticlesFromFile
// (SimulationState.cpp:229-277).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.hpp"

using namespace nbody;

extern "C" {

nbody_status nbody_snapshot_count(const char* path, size_t* n) {
    clear_error();
    NB_CHECK(path && n, NBODY_ERR_INVALID_ARG, "null argument");
    FILE* f = std::fopen(path, "rb");
    NB_CHECK(f, NBODY_ERR_IO, "cannot open %s", path);
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fclose(f);
    NB_CHECK(bytes >= 0, NBODY_ERR_IO, "cannot size %s", path);
    *n = static_cast<size_t>(bytes) / sizeof(nbody_particle);  // a trailing partial record is not read
    return NBODY_OK;
}

nbody_status nbody_snapshot_read(const char* path, nbody_particle* out, size_t capacity, size_t* n, int32_t recentre) {
    clear_error();
    NB_CHECK(path && n && (out || capacity == 0), NBODY_ERR_INVALID_ARG, "null argument");
    size_t count = 0;
    NB_TRY(nbody_snapshot_count(path, &count));
    NB_CHECK(count <= capacity, NBODY_ERR_CAPACITY, "%s holds %zu particles, capacity %zu", path, count, capacity);
    FILE* f = std::fopen(path, "rb");
    NB_CHECK(f, NBODY_ERR_IO, "cannot open %s", path);
    const size_t got = count ? std::fread(out, sizeof(nbody_particle), count, f) : 0;
    std::fclose(f);
    NB_CHECK(got == count, NBODY_ERR_IO, "short read of %s", path);
    *n = count;
    if (recentre && count) {
        // SimulationState.cpp:252-274: long double total mass, FP64 mass-weighted
        // position, centre rounded to FP32 and subtracted in FP32.
        long double total = 0.0L;
        double cx = 0, cy = 0, cz = 0;
        for (size_t i = 0; i < count; ++i) {
            total += out[i].mass;
            cx += static_cast<double>(out[i].position[0]) * out[i].mass;
            cy += static_cast<double>(out[i].position[1]) * out[i].mass;
            cz += static_cast<double>(out[i].position[2]) * out[i].mass;
        }
        const double t = static_cast<double>(total);
        const float c[3] = {static_cast<float>(cx / t), static_cast<float>(cy / t), static_cast<float>(cz / t)};
        for (size_t i = 0; i < count; ++i)
            for (int a = 0; a < 3; ++a) out[i].position[a] -= c[a];
    }
    return NBODY_OK;
}

nbody_status nbody_snapshot_write(const char* path, const nbody_particle* p, size_t n) {
    clear_error();
    NB_CHECK(path && (p || n == 0), NBODY_ERR_INVALID_ARG, "null argument");
    FILE* f = std::fopen(path, "wb");
    NB_CHECK(f, NBODY_ERR_IO, "cannot create %s", path);
    const size_t put = n ? std::fwrite(p, sizeof(nbody_particle), n, f) : 0;
    const int closed = std::fclose(f);
    NB_CHECK(put == n && closed == 0, NBODY_ERR_IO, "short write of %s", path);
    return NBODY_OK;
}

}  // extern "C"
