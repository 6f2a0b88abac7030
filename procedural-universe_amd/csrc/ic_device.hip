// Initial conditions generated in device memory.
//
// The reference seeds its particles on the host (IParticleSeeder::Seed,
// src/Sim/IParticleSeeder.hpp:12-50) and hands the array to INBodySim::Init.
// A server that admits galaxies by parameters instead generates them where
// they will live: a Plummer sphere from the same counter-based stream as the
// host generator (ic.cpp), particle i drawing only from (seed, i), then
// recentred on its centre of mass.  Positions and velocities are computed in
// FP64 and rounded once, as on the host; the recentring sums are FP64.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "ic_device.hpp"

namespace nbody {

namespace {

constexpr int kIcBlock = 256;
constexpr int kIcMaxBlocks = 1024;

__device__ inline unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct DStream {
    unsigned long long key;
    unsigned long long k;
    __device__ DStream(unsigned long long seed, unsigned long long i)
        : key(mix64(seed * 0x9E3779B97F4A7C15ULL + mix64(i + 0x632BE59BD9B4E019ULL))), k(0) {}
    __device__ double u() {
        const unsigned long long h = mix64(key + 0xD1B54A32D192ED03ULL * (++k));
        return (static_cast<double>(h >> 11) + 0.5) * 0x1.0p-53;
    }
    __device__ void direction(double& x, double& y, double& z) {
        const double c = 2.0 * u() - 1.0;
        const double s = sqrt(fmax(0.0, 1.0 - c * c));
        const double p = 2.0 * 3.14159265358979323846 * u();
        x = s * cos(p);
        y = s * sin(p);
        z = c;
    }
};

// Plummer (Aarseth, Henon & Wielen 1974), as the host generator: radius from
// the inverted mass profile with the tail beyond 20 a rejected, speed by von
// Neumann rejection on q^2 (1 - q^2)^(7/2).  Writes (x, y, z, G m), (v, 0) and
// per-block FP64 sums of positions and velocities for the recentring.
__global__ __launch_bounds__(kIcBlock) void plummer_kernel(unsigned long long seed, int n, double a, double m,
                                                           double G, float4* __restrict__ pos,
                                                           float4* __restrict__ vel, double* __restrict__ mass,
                                                           double* __restrict__ gm, double* __restrict__ partial) {
    __shared__ double red[6][kIcBlock];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const double vscale = sqrt(G * (m * n) / a);
    for (int i = blockIdx.x * kIcBlock + threadIdx.x; i < n; i += gridDim.x * kIcBlock) {
        DStream s(seed, static_cast<unsigned long long>(i));
        double r = 0;
        for (int attempt = 0; attempt < 64; ++attempt) {
            const double U = s.u();
            r = a / sqrt(pow(U, -2.0 / 3.0) - 1.0);
            if (r <= 20.0 * a) break;
            r = 20.0 * a;
        }
        double dx, dy, dz;
        s.direction(dx, dy, dz);
        double q = 0;
        for (int attempt = 0; attempt < 1000; ++attempt) {
            q = s.u();
            const double y = 0.1 * s.u();
            if (y < q * q * pow(1.0 - q * q, 3.5)) break;
        }
        const double ve = sqrt(2.0) * pow(1.0 + r * r / (a * a), -0.25) * vscale;
        double ux, uy, uz;
        s.direction(ux, uy, uz);
        const float px = static_cast<float>(r * dx), py = static_cast<float>(r * dy), pz = static_cast<float>(r * dz);
        const float vx = static_cast<float>(q * ve * ux), vy = static_cast<float>(q * ve * uy),
                    vz = static_cast<float>(q * ve * uz);
        pos[i] = make_float4(px, py, pz, static_cast<float>(G * m));
        vel[i] = make_float4(vx, vy, vz, 0.f);
        if (mass) mass[i] = m;
        if (gm) gm[i] = G * m;
        acc[0] += px, acc[1] += py, acc[2] += pz, acc[3] += vx, acc[4] += vy, acc[5] += vz;
    }
    for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int w = kIcBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int c = 0; c < 6; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Subtract the mean position and velocity (equal masses: the centre of mass).
__global__ __launch_bounds__(kIcBlock) void recentre_kernel(int n, int nblocks, const double* __restrict__ partial,
                                                            float4* __restrict__ pos, float4* __restrict__ vel,
                                                            float4* __restrict__ pos2) {
    __shared__ double mean[6];
    if (threadIdx.x < 6) {
        double s = 0;
        for (int b = 0; b < nblocks; ++b) s += partial[b * 6 + threadIdx.x];
        mean[threadIdx.x] = s / n;
    }
    __syncthreads();
    for (int i = blockIdx.x * kIcBlock + threadIdx.x; i < n; i += gridDim.x * kIcBlock) {
        float4 p = pos[i];
        float4 v = vel[i];
        p.x = static_cast<float>(static_cast<double>(p.x) - mean[0]);
        p.y = static_cast<float>(static_cast<double>(p.y) - mean[1]);
        p.z = static_cast<float>(static_cast<double>(p.z) - mean[2]);
        v.x = static_cast<float>(static_cast<double>(v.x) - mean[3]);
        v.y = static_cast<float>(static_cast<double>(v.y) - mean[4]);
        v.z = static_cast<float>(static_cast<double>(v.z) - mean[5]);
        pos[i] = p;
        vel[i] = v;
        if (pos2) pos2[i] = p;
    }
}

}  // namespace

size_t ic_device_scratch_doubles() { return 6 * kIcMaxBlocks; }

hipError_t ic_device_plummer(unsigned long long seed, int n, double a, double total_mass, double G, float4* pos,
                             float4* vel, double* mass, double* gm, float4* pos_copy, double* scratch,
                             hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    int blocks = (n + kIcBlock - 1) / kIcBlock;
    if (blocks > kIcMaxBlocks) blocks = kIcMaxBlocks;
    hipLaunchKernelGGL(plummer_kernel, dim3(blocks), dim3(kIcBlock), 0, stream, seed, n, a, total_mass / n, G, pos,
                       vel, mass, gm, scratch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(recentre_kernel, dim3(blocks), dim3(kIcBlock), 0, stream, n, blocks, scratch, pos, vel,
                       pos_copy);
    return hipGetLastError();
}

}  // namespace nbody
