// FP64 diagnostics of a device-resident state: kinetic energy, pair
// potential energy, linear and angular momentum.
//
// Positions are the FP32 state; masses come from an FP64 copy so that the
// diagnostics do not inherit the FP32 rounding of G*m.  Every sum is formed in
// a fixed order (per-block tree in LDS, then blocks in index order on the
// host), so the result is reproducible run to run.
//
// Pair potentials (the integral of each law's pair force from r to infinity):
//   law R: U(r) = -(G m_i m_j / sqrt(S)) * (pi/2 - atan(r / sqrt(S)));  S = 0: -G m_i m_j / r
//   law P: U(r) = -G m_i m_j / sqrt(r^2 + eps^2)
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/nbody.h"
#include "diag.hpp"

namespace nbody {

namespace {

constexpr int kDiagBlock = 256;
constexpr int kDiagTile = 256;

__device__ double block_sum(double v, double* red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int w = kDiagBlock / 2; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// out[block * 8 + {0: kinetic, 1-3: momentum, 4-6: angular momentum, 7: mass}]
__global__ __launch_bounds__(kDiagBlock) void diag_local_kernel(const float4* __restrict__ pos,
                                                                const float4* __restrict__ vel,
                                                                const double* __restrict__ mass, int n,
                                                                double* __restrict__ out) {
    __shared__ double red[kDiagBlock];
    double q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * kDiagBlock + threadIdx.x; i < n; i += gridDim.x * kDiagBlock) {
        const double m = mass[i];
        const float4 x = pos[i];
        const float4 v = vel[i];
        const double vx = v.x, vy = v.y, vz = v.z;
        const double px = x.x, py = x.y, pz = x.z;
        q[0] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
        q[1] += m * vx;
        q[2] += m * vy;
        q[3] += m * vz;
        q[4] += m * (py * vz - pz * vy);
        q[5] += m * (pz * vx - px * vz);
        q[6] += m * (px * vy - py * vx);
        q[7] += m;
    }
    for (int c = 0; c < 8; ++c) {
        const double s = block_sum(q[c], red);
        if (threadIdx.x == 0) out[blockIdx.x * 8 + c] = s;
    }
}

// out[block] = -1/2 sum_{i in block} m_i sum_{j != i} (G m_j) g(r_ij)
__global__ __launch_bounds__(kDiagBlock) void diag_potential_kernel(const float4* __restrict__ pos,
                                                                    const double* __restrict__ mass,
                                                                    const double* __restrict__ gm, int n,
                                                                    int law, double soft,
                                                                    double* __restrict__ out) {
    __shared__ double red[kDiagBlock];
    __shared__ double tx[kDiagTile], ty[kDiagTile], tz[kDiagTile], tg[kDiagTile];
    const int i = blockIdx.x * kDiagBlock + threadIdx.x;
    const bool active = i < n;
    double xi = 0, yi = 0, zi = 0, mi = 0;
    if (active) {
        const float4 p = pos[i];
        xi = p.x;
        yi = p.y;
        zi = p.z;
        mi = mass[i];
    }
    const double rs = soft > 0 ? sqrt(soft) : 0.0;
    const double half_pi = 1.5707963267948966;
    double phi = 0.0;
    for (int jt = 0; jt < n; jt += kDiagTile) {
        __syncthreads();
        const int j = jt + threadIdx.x;
        if (threadIdx.x < kDiagTile) {
            if (j < n) {
                const float4 p = pos[j];
                tx[threadIdx.x] = p.x;
                ty[threadIdx.x] = p.y;
                tz[threadIdx.x] = p.z;
                tg[threadIdx.x] = gm[j];
            } else {
                tx[threadIdx.x] = ty[threadIdx.x] = tz[threadIdx.x] = 0.0;
                tg[threadIdx.x] = 0.0;
            }
        }
        __syncthreads();
        if (!active) continue;
        const int jn = min(kDiagTile, n - jt);
        for (int l = 0; l < jn; ++l) {
            if (jt + l == i) continue;
            const double dx = tx[l] - xi, dy = ty[l] - yi, dz = tz[l] - zi;
            const double r2 = dx * dx + dy * dy + dz * dz;
            double g;
            if (law == NBODY_LAW_R) {
                if (rs > 0) {
                    g = (half_pi - atan(sqrt(r2) / rs)) / rs;
                } else {
                    g = r2 > 0 ? 1.0 / sqrt(r2) : 0.0;
                }
            } else {
                g = 1.0 / sqrt(r2 + soft);
            }
            phi -= tg[l] * g;
        }
    }
    const double s = block_sum(active ? 0.5 * mi * phi : 0.0, red);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

}  // namespace

int diag_local_blocks(int n) {
    int b = (n + kDiagBlock - 1) / kDiagBlock;
    return b < 1 ? 1 : (b > 1024 ? 1024 : b);
}

int diag_potential_blocks(int n) { return n < 1 ? 1 : (n + kDiagBlock - 1) / kDiagBlock; }

hipError_t diag_launch(const float4* pos, const float4* vel, const double* mass, const double* gm, int n,
                       int law, double soft, double* local_out, double* pot_out, hipStream_t stream) {
    hipLaunchKernelGGL(diag_local_kernel, dim3(diag_local_blocks(n)), dim3(kDiagBlock), 0, stream, pos, vel,
                       mass, n, local_out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(diag_potential_kernel, dim3(diag_potential_blocks(n)), dim3(kDiagBlock), 0, stream,
                       pos, mass, gm, n, law, soft, pot_out);
    return hipGetLastError();
}

}  // namespace nbody
