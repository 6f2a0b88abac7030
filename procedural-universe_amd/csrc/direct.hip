// Direct-sum gravity on gfx950: O(N^2) pair forces with LDS j-tiles.
//
// Replaces the reference's BruteForceCPU::Exec/Update
// (src/Sim/BruteForceCPU.cpp:25-73) and the D3D11 pair of compute shaders
// (shaders/Compute/Gravity.csh:28-48, Update.csh:22-36), written from the
// force law, not translated.
//
// Layout: one float4 (x, y, z, G*m) per particle, padded to a multiple of
// kTile with massless dummies at the origin.  A workgroup owns
// block * i_per_lane targets; each lane keeps i_per_lane targets in registers.
// The workgroup streams its j-range through LDS one tile of kTile particles at
// a time (the next tile is prefetched into registers while the current one is
// consumed), and every lane reads the same LDS record per j, so the read is a
// broadcast.  Per tile the lane sums into plain FP32 registers; the tile sums
// are then added into a per-target compensated (Kahan) accumulator, so the
// rounding of the long sum grows with the number of tiles, not of particles.
//
// The pair arithmetic, the issue-rate choices behind it and the compensated
// summation are described with the shared core in direct_core.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "../../include/nbody.h"
#include "common.hpp"
#include "direct.hpp"
#include "direct_core.hpp"

namespace nbody {

namespace {

using core::kahan_add;

// Kick then drift of target i with the acceleration (ax, ay, az).
__device__ __forceinline__ void kick_drift(const DirectLaunch& L, int i, float ax, float ay, float az) {
    float4 v = L.vel[i];
    float4 x = L.src[i];
    v.x = __builtin_fmaf(ax, L.dt, v.x);
    v.y = __builtin_fmaf(ay, L.dt, v.y);
    v.z = __builtin_fmaf(az, L.dt, v.z);
    x.x = __builtin_fmaf(v.x, L.dts, x.x);
    x.y = __builtin_fmaf(v.y, L.dts, x.y);
    x.z = __builtin_fmaf(v.z, L.dts, x.z);
    L.vel[i] = v;
    L.pos_out[i] = x;
}

// One workgroup: targets [blockIdx.x * BLOCK * K, +BLOCK*K), sources
// [blockIdx.y * jchunk, +jchunk) clipped to n_pad.  KICK fuses the kick-drift
// update (only launched with a single j-chunk).
template <int LAW, int K, int BLOCK, bool KICK>
__global__ __launch_bounds__(BLOCK) void direct_force_kernel(DirectLaunch L, int jchunk) {
    __shared__ core::SrcTile tiles[2];

    const int tid = threadIdx.x;
    const int ibase = blockIdx.x * (BLOCK * K);

    float xi[K], yi[K], zi[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = min(ibase + tid + k * BLOCK, L.n_pad - 1);
        const float4 q = L.src[i];
        xi[k] = q.x * L.sc.inv_len;
        yi[k] = q.y * L.sc.inv_len;
        zi[k] = q.z * L.sc.inv_len;
    }

    const int j0 = blockIdx.y * jchunk;
    const int j1 = min(j0 + jchunk, L.n_pad);
    float sx[K], sy[K], sz[K];
    core::block_forces<LAW, K, BLOCK>(tiles, L.src, j0, j1, L.sc, xi, yi, zi, sx, sy, sz);

#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = ibase + tid + k * BLOCK;
        if constexpr (KICK) {
            if (i < L.n) {
                const float ax = sx[k] * L.sc.out, ay = sy[k] * L.sc.out, az = sz[k] * L.sc.out;
                kick_drift(L, i, ax, ay, az);
                L.acc[i] = make_float4(ax, ay, az, 0.f);
            }
        } else {
            if (i < L.n_pad) L.part[(size_t)blockIdx.y * L.n_pad + i] = make_float4(sx[k], sy[k], sz[k], 0.f);
        }
    }
}

// Adds the j-chunk partial sums of each target in chunk order (compensated),
// scales them back, then either kicks and drifts (kick != 0) or only stores
// the acceleration.
__global__ __launch_bounds__(256) void reduce_kick_kernel(DirectLaunch L, int split, int kick) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.n) return;
    float sx = 0.f, sy = 0.f, sz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
    for (int s = 0; s < split; ++s) {
        const float4 p = L.part[(size_t)s * L.n_pad + i];
        kahan_add(sx, cx, p.x);
        kahan_add(sy, cy, p.y);
        kahan_add(sz, cz, p.z);
    }
    const float ax = sx * L.sc.out, ay = sy * L.sc.out, az = sz * L.sc.out;
    L.acc[i] = make_float4(ax, ay, az, 0.f);
    if (kick) kick_drift(L, i, ax, ay, az);
}

// Acceleration at arbitrary points from all sources: one target per lane,
// every workgroup over the whole source range.
template <int LAW>
__global__ __launch_bounds__(256) void field_kernel(const float4* __restrict__ src, int n_src_pad,
                                                    const float4* __restrict__ tgt, int m, DirectScale sc,
                                                    float4* __restrict__ out) {
    __shared__ core::SrcTile tiles[2];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const float4 q = tgt[min(i, m - 1)];
    float xi[1] = {q.x * sc.inv_len}, yi[1] = {q.y * sc.inv_len}, zi[1] = {q.z * sc.inv_len};
    float sx[1], sy[1], sz[1];
    core::block_forces<LAW, 1, 256>(tiles, src, 0, n_src_pad, sc, xi, yi, zi, sx, sy, sz);
    if (i < m) out[i] = make_float4(sx[0] * sc.out, sy[0] * sc.out, sz[0] * sc.out, 0.f);
}

// (x, y, z) floats -> float4 targets.
__global__ __launch_bounds__(256) void pack_points_kernel(const float* __restrict__ in, int m, float4* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < m) out[i] = make_float4(in[3 * i], in[3 * i + 1], in[3 * i + 2], 0.f);
}

// float4 accelerations -> (ax, ay, az) floats.
__global__ __launch_bounds__(256) void unpack_points_kernel(const float4* __restrict__ in, int m, float* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < m) {
        const float4 a = in[i];
        out[3 * i] = a.x;
        out[3 * i + 1] = a.y;
        out[3 * i + 2] = a.z;
    }
}

using ForceFn = void (*)(DirectLaunch, int);

template <int LAW, int K, int BLOCK>
ForceFn pick_kick(bool kick) {
    return kick ? direct_force_kernel<LAW, K, BLOCK, true> : direct_force_kernel<LAW, K, BLOCK, false>;
}

template <int LAW, int K>
ForceFn pick_block(int block, bool kick) {
    switch (block) {
        case 64: return pick_kick<LAW, K, 64>(kick);
        case 128: return pick_kick<LAW, K, 128>(kick);
        case 256: return pick_kick<LAW, K, 256>(kick);
        default: return nullptr;
    }
}

template <int LAW>
ForceFn pick_k(int k, int block, bool kick) {
    switch (k) {
        case 1: return pick_block<LAW, 1>(block, kick);
        case 2: return pick_block<LAW, 2>(block, kick);
        case 4: return pick_block<LAW, 4>(block, kick);
        case 8: return pick_block<LAW, 8>(block, kick);
        default: return nullptr;
    }
}

ForceFn pick(int law, int k, int block, bool kick) {
    return law == NBODY_LAW_R ? pick_k<NBODY_LAW_R>(k, block, kick) : pick_k<NBODY_LAW_P>(k, block, kick);
}

}  // namespace

bool direct_shape_valid(const DirectShape& s) {
    const bool k_ok = s.i_per_lane == 1 || s.i_per_lane == 2 || s.i_per_lane == 4 || s.i_per_lane == 8;
    const bool b_ok = s.block == 64 || s.block == 128 || s.block == 256;
    return k_ok && b_ok && s.split >= 1;
}

DirectScale direct_scale_for(int law, double softening, double extent, double max_gm) {
    DirectScale sc{};
    // lambda: the power of two nearest the geometric mean of the shortest and
    // longest distances that matter, so that their sixth powers (law R's rsq
    // argument grows as r^6) sit symmetrically inside FP32's range.  The
    // longest is the diagonal of the extent; the shortest is the softening
    // length, or the spacing of FP32 positions across the extent (2^-24 of it)
    // when that is larger.
    const double r_hi = std::sqrt(3.0) * std::max(extent, 0.0);
    const double r_lo = std::max(std::sqrt(std::max(softening, 0.0)), std::ldexp(std::max(extent, 0.0), -24));
    int e_len = 0;
    if (r_hi > 0 && r_lo > 0 && std::isfinite(r_hi) && std::isfinite(r_lo))
        e_len = static_cast<int>(std::lround(0.5 * (std::log2(r_hi) + std::log2(r_lo))));
    else if (r_lo > 0 && std::isfinite(r_lo))
        e_len = static_cast<int>(std::lround(std::log2(r_lo)));
    e_len = std::max(-120, std::min(120, e_len));
    // c: the power of two at or below the largest G m
    int e_mass = 0;
    if (max_gm > 0 && std::isfinite(max_gm)) {
        (void)std::frexp(max_gm, &e_mass);  // max_gm in [2^(e-1), 2^e)
        e_mass -= 1;
    }
    e_mass = std::max(-120, std::min(120, e_mass));
    // keep out = 2^(e_mass - 2 e_len) a normal float
    e_mass = std::max(2 * e_len - 120, std::min(2 * e_len + 120, e_mass));
    sc.inv_len = std::ldexp(1.0f, -e_len);
    sc.mass_k = law == NBODY_LAW_R ? std::ldexp(1.0f, e_mass) : std::ldexp(1.0f, -e_mass);
    sc.soft = static_cast<float>(std::ldexp(softening, -2 * e_len));
    sc.out = std::ldexp(1.0f, e_mass - 2 * e_len);
    return sc;
}

int direct_effective_split(const DirectShape& s, int n_pad) {
    const int tiles = n_pad / kTile;
    int split = s.split < 1 ? 1 : s.split;
    if (split > tiles) split = tiles;
    const int tiles_per_chunk = (tiles + split - 1) / split;
    return (tiles + tiles_per_chunk - 1) / tiles_per_chunk;
}

DirectShape direct_default_shape(int n_pad) {
    // One target per lane keeps the register count, and so the waves per SIMD,
    // highest; below ~1024 workgroups of targets the j-range is split so that
    // the 256 CUs still see about a thousand workgroups.
    DirectShape s{1, 256, 1};
    const int iblocks = (n_pad + s.block - 1) / s.block;
    if (iblocks < 1024) s.split = (1024 + iblocks - 1) / iblocks;
    return s;
}

hipError_t direct_field(const float4* src, int n_src_pad, const float4* tgt, int m, int law, const DirectScale& sc,
                        float4* out, hipStream_t stream) {
    if (m <= 0) return hipSuccess;
    const dim3 grid((m + 255) / 256);
    if (law == NBODY_LAW_R)
        hipLaunchKernelGGL(field_kernel<NBODY_LAW_R>, grid, dim3(256), 0, stream, src, n_src_pad, tgt, m, sc, out);
    else
        hipLaunchKernelGGL(field_kernel<NBODY_LAW_P>, grid, dim3(256), 0, stream, src, n_src_pad, tgt, m, sc, out);
    return hipGetLastError();
}

hipError_t direct_field_staged(const float4* src, int n_src_pad, float4* staging, int cap, int m, int law,
                               const DirectScale& sc, hipStream_t stream) {
    if (m <= 0) return hipSuccess;
    if (m > cap) return hipErrorInvalidValue;
    const dim3 grid((m + 255) / 256);
    float4* half0 = staging;
    float4* half1 = staging + cap;
    hipLaunchKernelGGL(pack_points_kernel, grid, dim3(256), 0, stream, reinterpret_cast<const float*>(half1), m, half0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = direct_field(src, n_src_pad, half0, m, law, sc, half1, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(unpack_points_kernel, grid, dim3(256), 0, stream, half1, m, reinterpret_cast<float*>(half0));
    return hipGetLastError();
}

hipError_t direct_step(const DirectLaunch& L, const DirectShape& s, bool kick, hipStream_t stream,
                       hipEvent_t ev_force_begin, hipEvent_t ev_force_end) {
    const int split = direct_effective_split(s, L.n_pad);
    const int tiles = L.n_pad / kTile;
    const int jchunk = ((tiles + split - 1) / split) * kTile;
    const bool fused = kick && split == 1;
    ForceFn fn = pick(L.law, s.i_per_lane, s.block, fused);
    if (!fn) return hipErrorInvalidValue;
    const int per_block = s.block * s.i_per_lane;
    dim3 grid((L.n_pad + per_block - 1) / per_block, split);
    if (ev_force_begin) {
        hipError_t e = hipEventRecord(ev_force_begin, stream);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fn, grid, dim3(s.block), 0, stream, L, jchunk);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (ev_force_end) {
        e = hipEventRecord(ev_force_end, stream);
        if (e != hipSuccess) return e;
    }
    if (!fused) {
        DirectLaunch R = L;
        hipLaunchKernelGGL(reduce_kick_kernel, dim3((L.n + 255) / 256), dim3(256), 0, stream, R, split,
                           kick ? 1 : 0);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace nbody

extern "C" nbody_status nbody_direct_scale(int32_t law, double softening, double extent, double max_gm,
                                           float* out4) {
    nbody::clear_error();
    NB_CHECK(out4, NBODY_ERR_INVALID_ARG, "out is null");
    NB_CHECK(law == NBODY_LAW_R || law == NBODY_LAW_P, NBODY_ERR_INVALID_ARG, "unknown law %d", law);
    const nbody::DirectScale sc = nbody::direct_scale_for(law, softening, extent, max_gm);
    out4[0] = sc.inv_len;
    out4[1] = sc.mass_k;
    out4[2] = sc.soft;
    out4[3] = sc.out;
    return NBODY_OK;
}
