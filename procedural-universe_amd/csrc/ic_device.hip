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

// ---- the reference's seeders ---------------------------------------------

struct SeedArgs {
    int kind;
    unsigned long long seed;
    int n;
    float scale;
    double G;
    int arm_total;  // galaxy: particles in the two arms
    int per_loop;   // galaxy: particles per arm step
    SeederTable T;
};

__device__ inline double normal_draw(DStream& s) {  // Box-Muller, standard normal
    const double u1 = s.u(), u2 = s.u();
    return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
}

__device__ inline float3 cross(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

__device__ inline float3 normalized(float3 a) {
    const float l = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    return l > 0.f ? make_float3(a.x / l, a.y / l, a.z / l) : make_float3(0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(kIcBlock) void seeded_kernel(SeedArgs A, float4* __restrict__ pos,
                                                          float4* __restrict__ vel, double* __restrict__ mass,
                                                          double* __restrict__ gm, float4* __restrict__ pos2) {
    for (int i = blockIdx.x * kIcBlock + threadIdx.x; i < A.n; i += gridDim.x * kIcBlock) {
        DStream s(A.seed, static_cast<unsigned long long>(i));
        float3 p;
        double3 v;
        double m;
        if (A.kind == 0) {
            // RandomSeeder.cpp:14-40: cube of half-width 500 / scale, radial
            // velocity 1e16, mass uniform in [1e20, 1e30]
            p = make_float3(static_cast<float>(-500.0 + 1000.0 * s.u()) / A.scale,
                            static_cast<float>(-500.0 + 1000.0 * s.u()) / A.scale,
                            static_cast<float>(-500.0 + 1000.0 * s.u()) / A.scale);
            const float3 d = normalized(p);
            v = make_double3(d.x * 1e16, d.y * 1e16, d.z * 1e16);
            m = 1e20 + (1e30 - 1e20) * s.u();
        } else if (A.kind == 2) {
            // StarSystemSeeder.cpp:20-55: a 1e30 star at the origin; bodies on
            // the z axis at 4-7 AU, velocity (U, U / 20, 0) of 1-5 AU, mass
            // uniform in [1e10, 1e26] (AU = 1.15e12 m, 1 m = 1000 units, the
            // positions divided by 20 AU)
            const double au_m = 1.15e12 * 1000.0;
            if (i == 0) {
                p = make_float3(0.f, 0.f, 0.f);
                v = make_double3(0.0, 0.0, 0.0);
                m = 1e30;
            } else {
                const double r = (4.0 + 3.0 * s.u()) * au_m;
                p = make_float3(0.f, 0.f, static_cast<float>(r / (20.0 * 1.15e12)));
                const double vx = (1.0 + 4.0 * s.u()) * au_m, vy = (1.0 + 4.0 * s.u()) * au_m / 20.0;
                v = make_double3(vx, vy, 0.0);
                m = 1e10 + (1e26 - 1e10) * s.u();
            }
        } else {
            // GalaxySeeder.cpp:42-144
            float3 q;
            if (i < A.arm_total) {
                // two spiral arms (offsets 0 and 3.14), 61 steps of the float
                // angle, per_loop particles per step
                const int per_arm = A.T.loops * A.per_loop;
                const float offset = i < per_arm ? 0.0f : 3.14f;
                const int k = (i % per_arm) / A.per_loop;
                const float angle = A.T.angle[k], r = A.T.radius[k];
                const float3 spiral = make_float3(cosf(angle + offset) * r, sinf(angle + offset) * r, 0.f);
                const float3 spiraln = make_float3(cosf(angle + offset + 0.1f) * (r + 10.f),
                                                   sinf(angle + offset + 0.1f) * (r + 10.f), 0.f);
                float3 nrm = make_float3(spiral.x - spiraln.x, spiral.y - spiraln.y, spiral.z - spiraln.z);
                const float mag = sqrtf(nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z);
                nrm = normalized(nrm);
                const float3 tan = normalized(cross(nrm, make_float3(0.f, 0.f, 1.f)));
                const float lx = static_cast<float>(0.5 + 0.2 * normal_draw(s));  // normal(0.5, 0.2)
                const float ly = static_cast<float>(0.2 + 0.3 * s.u());           // uniform(0.2, 0.5)
                // Lerp(sx, ex, lx) + Lerp(sy, ey, ly), sx/ex = spiral -/+ 140 tangent, sy/ey = spiral -/+ 400 normal
                q = make_float3(spiral.x + tan.x * 140.f * (2.f * lx - 1.f) + spiral.x + nrm.x * 400.f * (2.f * ly - 1.f),
                                spiral.y + tan.y * 140.f * (2.f * lx - 1.f) + spiral.y + nrm.y * 400.f * (2.f * ly - 1.f),
                                0.f);
                q.z = static_cast<float>(16.0 * normal_draw(s));
                const float f = 2e16f * (1000.0f / mag);
                v = make_double3(nrm.x * f, nrm.y * f, nrm.z * f);
                m = 1e20;
            } else {
                // the disc: uniform within 720 of the centre, z normal(0, 16),
                // velocity (y, -x, 0) * uniform(0.8, 1.2) * 1e14, mass uniform in [1e28, 1e30]
                float x = 0.f, y = 0.f, z = 0.f;
                for (int attempt = 0; attempt < 256; ++attempt) {
                    x = static_cast<float>(-2000.0 + 4000.0 * s.u());
                    y = static_cast<float>(-2000.0 + 4000.0 * s.u());
                    z = static_cast<float>(16.0 * normal_draw(s));
                    if (x * x + y * y + z * z <= 720.f * 720.f) break;
                }
                q = make_float3(x, y, z);
                const double f = (0.8 + 0.4 * s.u()) * 1e14;
                v = make_double3(y * f, -x * f, 0.0);
                m = 1e28 + (1e30 - 1e28) * s.u();
            }
            // Transform(q / scale, orientation): row vector times the rotation;
            // the velocities stay unrotated, as in the reference
            const float3 w = make_float3(q.x / A.scale, q.y / A.scale, q.z / A.scale);
            const float* R = A.T.rot;
            p = make_float3(w.x * R[0] + w.y * R[3] + w.z * R[6], w.x * R[1] + w.y * R[4] + w.z * R[7],
                            w.x * R[2] + w.y * R[5] + w.z * R[8]);
        }
        const float4 pp = make_float4(p.x, p.y, p.z, static_cast<float>(A.G * m));
        pos[i] = pp;
        if (pos2) pos2[i] = pp;
        vel[i] = make_float4(static_cast<float>(v.x), static_cast<float>(v.y), static_cast<float>(v.z), 0.f);
        if (mass) mass[i] = m;
        if (gm) gm[i] = A.G * m;
    }
}

}  // namespace

size_t ic_device_scratch_doubles() { return 6 * kIcMaxBlocks; }

void ic_seeder_table(unsigned long long seed, SeederTable* T) {
    // GalaxySeeder.cpp:112-118: for (float angle = 0, r = 2; angle < 6; angle += 0.1f, r += 7.2f)
    float angle = 0.0f, r = 2.0f;
    int k = 0;
    for (; angle < 6.0f && k < 64; angle += 0.1f, r += 7.2f, ++k) {
        T->angle[k] = angle;
        T->radius[k] = r;
    }
    T->loops = k;
    // Orientation from three angles uniform in [0, 2 pi) (GalaxySeeder.cpp:52-54):
    // roll about z, then pitch about x, then yaw about y, for row vectors
    // (Matrix::CreateFromYawPitchRoll).
    const unsigned long long key = seed * 0x9E3779B97F4A7C15ULL + 0x5EEDULL;
    double a[3];
    for (int c = 0; c < 3; ++c) {
        unsigned long long z = key + 0x632BE59BD9B4E019ULL * static_cast<unsigned long long>(c + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        a[c] = 2.0 * 3.14159265358979323846 * ((static_cast<double>(z >> 11) + 0.5) * 0x1.0p-53);
    }
    const double yaw = a[0], pitch = a[1], roll = a[2];
    const double cr = std::cos(roll), sr = std::sin(roll), cp = std::cos(pitch), sp = std::sin(pitch);
    const double cy = std::cos(yaw), sy = std::sin(yaw);
    const double Rz[9] = {cr, sr, 0, -sr, cr, 0, 0, 0, 1};
    const double Rx[9] = {1, 0, 0, 0, cp, sp, 0, -sp, cp};
    const double Ry[9] = {cy, 0, -sy, 0, 1, 0, sy, 0, cy};
    double t[9], M[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            t[3 * i + j] = 0;
            for (int k2 = 0; k2 < 3; ++k2) t[3 * i + j] += Rz[3 * i + k2] * Rx[3 * k2 + j];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            M[3 * i + j] = 0;
            for (int k2 = 0; k2 < 3; ++k2) M[3 * i + j] += t[3 * i + k2] * Ry[3 * k2 + j];
        }
    for (int c = 0; c < 9; ++c) T->rot[c] = static_cast<float>(M[c]);
}

hipError_t ic_device_seeded(int kind, unsigned long long seed, int n, float scale, double G, const SeederTable& T,
                            float4* pos, float4* vel, double* mass, double* gm, float4* pos_copy, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    SeedArgs A{};
    A.kind = kind;
    A.seed = seed;
    A.n = n;
    A.scale = scale;
    A.G = G;
    A.T = T;
    if (kind == 1) {
        // GalaxySeeder.cpp:109-111: numPerLoop = floor(N * 0.4f / floor(6.0f / 0.1f)) per arm step
        const float loops_f = std::floor(6.0f / 0.1f);
        A.per_loop = static_cast<int>(std::floor(static_cast<float>(n) * 0.4f / loops_f));
        A.arm_total = 2 * T.loops * A.per_loop;
    }
    int blocks = (n + kIcBlock - 1) / kIcBlock;
    if (blocks > kIcMaxBlocks) blocks = kIcMaxBlocks;
    hipLaunchKernelGGL(seeded_kernel, dim3(blocks), dim3(kIcBlock), 0, stream, A, pos, vel, mass, gm, pos_copy);
    return hipGetLastError();
}

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
