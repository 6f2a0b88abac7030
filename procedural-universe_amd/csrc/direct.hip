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
// The pair arithmetic is FP32 with the hardware reciprocal square root:
//   law R (reference, Physics.hpp:25-35):  a += G m_j d / (|d| (|d|^2 + S))
//            computed as d * (G m_j) * rsq(r2 * (r2 + S)^2 + tiny)
//   law P (Plummer):                     a += G m_j d / (|d|^2 + eps^2)^(3/2)
// with d = x_j - x_i.  For law R, `tiny` keeps a coincident pair (r2 == 0,
// including j == i) finite: the factor becomes G m_j / sqrt(tiny), which the
// host keeps below FLT_MAX, so d = 0 makes the term exactly 0, as the
// reference's Normalize of a zero vector does; for r2 > 0 it changes the
// factor by tiny / q (below 1e-7 unless q < 1e-30, i.e. pairs far closer than
// FP32 positions of order 1 can resolve).
//
// Issue rates on gfx950 shape the arithmetic: FP32 add/sub/mul/fma whose
// operands are all VGPRs issue at twice the rate of the same instruction with
// an SGPR or literal operand, or of v_min/v_max, and v_rsq_f32 issues at half
// that single rate.  So every per-pair operation is a VGPR-only
// add/sub/mul/fma (12 of them) plus one v_rsq_f32: the softening and `tiny`
// are copied into VGPRs once per workgroup, and the coincident-pair guard is
// an FMA rather than a clamp.  The rsq of four pairs are issued together.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "../../include/nbody.h"
#include "direct.hpp"

namespace nbody {

namespace {

__device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// A uniform value held in a VGPR, so that the arithmetic using it keeps
// all-VGPR operands (an SGPR operand halves the issue rate).
__device__ __forceinline__ float in_vgpr(float x) {
    float v;
    asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(x));
    return v;
}

// Four reciprocal square roots issued back to back.  Interleaved with the
// dual-issued FP32 arithmetic a lone v_rsq_f32 stalls the SIMD for far longer
// than its own issue slots; grouped, the transcendental unit stays busy while
// the surrounding arithmetic of other waves pairs up.  The trailing s_nop
// covers the transcendental-result forwarding wait before the first use.
// The compiler itself separates a v_rsq_f32 result from its first VALU use by
// one wait state; the group ends with NBODY_RSQ_NOP + 1 of them.
#ifndef NBODY_RSQ_NOP
#define NBODY_RSQ_NOP 1
#endif
// Source groups of four unrolled per inner-loop iteration.
#ifndef NBODY_GROUP_UNROLL
#define NBODY_GROUP_UNROLL 4
#endif
#define NBODY_STR2(x) #x
#define NBODY_STR(x) NBODY_STR2(x)

__device__ __forceinline__ void rsq4(float& a, float& b, float& c, float& d) {
    asm volatile(
        "v_rsq_f32 %0, %0\n\t"
        "v_rsq_f32 %1, %1\n\t"
        "v_rsq_f32 %2, %2\n\t"
        "v_rsq_f32 %3, %3\n\t"
        "s_nop " NBODY_STR(NBODY_RSQ_NOP)
        : "+v"(a), "+v"(b), "+v"(c), "+v"(d));
}

// Argument of the reciprocal square root for one pair: r2 (r2 + S)^2 + tiny
// (law R) or r2 + eps^2 (law P), with the difference vector d = x_j - x_i.
template <int LAW>
__device__ __forceinline__ float pair_arg(const float4 pj, const float xi, const float yi, const float zi,
                                          const float soft, const float tiny, float& dx, float& dy, float& dz) {
    dx = pj.x - xi;
    dy = pj.y - yi;
    dz = pj.z - zi;
    if constexpr (LAW == NBODY_LAW_R) {
        const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
        const float t = r2 + soft;
        const float u = r2 * t;
        return __builtin_fmaf(u, t, tiny);
    } else {
        return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, soft)));
    }
}

// acc += d * G m_j * f(rsq): law R uses rsq directly, law P its cube.
template <int LAW>
__device__ __forceinline__ void pair_acc(const float gm, const float inv, const float dx, const float dy,
                                         const float dz, float& ax, float& ay, float& az) {
    float s;
    if constexpr (LAW == NBODY_LAW_R) {
        s = gm * inv;
    } else {
        s = (gm * inv) * (inv * inv);
    }
    ax = __builtin_fmaf(dx, s, ax);
    ay = __builtin_fmaf(dy, s, ay);
    az = __builtin_fmaf(dz, s, az);
}

// One workgroup: targets [blockIdx.x * BLOCK * K, +BLOCK*K), sources
// [blockIdx.y * jchunk, +jchunk) clipped to n_pad.  KICK fuses the kick-drift
// update (only launched with a single j-chunk).
template <int LAW, int K, int BLOCK, bool KICK>
__global__ __launch_bounds__(BLOCK) void direct_force_kernel(DirectLaunch L, int jchunk) {
    static_assert(kTile % BLOCK == 0, "tile must be a multiple of the block");
    constexpr int PER = kTile / BLOCK;
    __shared__ float4 tile[kTile];

    const int tid = threadIdx.x;
    const int ibase = blockIdx.x * (BLOCK * K);
    const float soft = in_vgpr(L.soft);
    const float tiny = in_vgpr(L.tiny);

    float xi[K], yi[K], zi[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = min(ibase + tid + k * BLOCK, L.n_pad - 1);
        const float4 q = L.src[i];
        xi[k] = q.x;
        yi[k] = q.y;
        zi[k] = q.z;
    }

    float sx[K], sy[K], sz[K], cx[K], cy[K], cz[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        sx[k] = sy[k] = sz[k] = 0.f;
        cx[k] = cy[k] = cz[k] = 0.f;
    }

    const int j0 = blockIdx.y * jchunk;
    const int j1 = min(j0 + jchunk, L.n_pad);

    float4 pre[PER];
#pragma unroll
    for (int l = 0; l < PER; ++l) pre[l] = L.src[j0 + tid + l * BLOCK];

    for (int jt = j0; jt < j1; jt += kTile) {
        __syncthreads();
#pragma unroll
        for (int l = 0; l < PER; ++l) tile[tid + l * BLOCK] = pre[l];
        __syncthreads();
        if (jt + kTile < j1) {
#pragma unroll
            for (int l = 0; l < PER; ++l) pre[l] = L.src[jt + kTile + tid + l * BLOCK];
        }

        float ax[K], ay[K], az[K];
#pragma unroll
        for (int k = 0; k < K; ++k) ax[k] = ay[k] = az[k] = 0.f;

        // Groups of four sources: four rsq arguments per target, then one
        // grouped rsq4 per target, then the accumulation.
#pragma unroll NBODY_GROUP_UNROLL
        for (int j = 0; j < kTile; j += 4) {
            float4 pj[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) pj[g] = tile[j + g];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                float dx[4], dy[4], dz[4], q[4];
#pragma unroll
                for (int g = 0; g < 4; ++g) q[g] = pair_arg<LAW>(pj[g], xi[k], yi[k], zi[k], soft, tiny, dx[g], dy[g], dz[g]);
                rsq4(q[0], q[1], q[2], q[3]);
#pragma unroll
                for (int g = 0; g < 4; ++g) pair_acc<LAW>(pj[g].w, q[g], dx[g], dy[g], dz[g], ax[k], ay[k], az[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            kahan_add(sx[k], cx[k], ax[k]);
            kahan_add(sy[k], cy[k], ay[k]);
            kahan_add(sz[k], cz[k], az[k]);
        }
    }

#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = ibase + tid + k * BLOCK;
        if constexpr (KICK) {
            if (i < L.n) {
                float4 v = L.vel[i];
                float4 x = L.src[i];
                v.x = __builtin_fmaf(sx[k], L.dt, v.x);
                v.y = __builtin_fmaf(sy[k], L.dt, v.y);
                v.z = __builtin_fmaf(sz[k], L.dt, v.z);
                x.x = __builtin_fmaf(v.x, L.dts, x.x);
                x.y = __builtin_fmaf(v.y, L.dts, x.y);
                x.z = __builtin_fmaf(v.z, L.dts, x.z);
                L.vel[i] = v;
                L.pos_out[i] = x;
                L.acc[i] = make_float4(sx[k], sy[k], sz[k], 0.f);
            }
        } else {
            if (i < L.n_pad) L.part[(size_t)blockIdx.y * L.n_pad + i] = make_float4(sx[k], sy[k], sz[k], 0.f);
        }
    }
}

// Adds the j-chunk partial sums of each target in chunk order (compensated),
// then either kicks and drifts (kick != 0) or only stores the acceleration.
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
    L.acc[i] = make_float4(sx, sy, sz, 0.f);
    if (kick) {
        float4 v = L.vel[i];
        float4 x = L.src[i];
        v.x = __builtin_fmaf(sx, L.dt, v.x);
        v.y = __builtin_fmaf(sy, L.dt, v.y);
        v.z = __builtin_fmaf(sz, L.dt, v.z);
        x.x = __builtin_fmaf(v.x, L.dts, x.x);
        x.y = __builtin_fmaf(v.y, L.dts, x.y);
        x.z = __builtin_fmaf(v.z, L.dts, x.z);
        L.vel[i] = v;
        L.pos_out[i] = x;
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

float direct_tiny_for(double max_gm) {
    const double lim = 2.0 * (max_gm > 0 ? max_gm : 0.0) / static_cast<double>(FLT_MAX);
    const double t = 4.0 * lim * lim;
    return static_cast<float>(t > static_cast<double>(FLT_MIN) ? t : static_cast<double>(FLT_MIN));
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
