// Device core of the direct sum, shared by the single-system kernel and the
// batch executor so that a system's forces are the same bits either way.
//
// Pair arithmetic (FP32, hardware reciprocal square root), with d = x_j - x_i:
//   law R (reference, src/Sim/Physics.hpp:25-35):  a += G m_j d / (|d| (|d|^2 + S))
//            computed as d * (G m_j) * rsq(r2 * (r2 + S)^2 + tiny)
//   law P (Plummer):                              a += G m_j d / (|d|^2 + eps^2)^(3/2)
// For law R, `tiny` keeps a coincident pair (r2 == 0, including j == i) finite:
// the factor becomes G m_j / sqrt(tiny), which the host keeps below FLT_MAX, so
// d = 0 makes the term exactly 0, as the reference's Normalize of a zero vector
// does; for r2 > 0 it changes the factor by tiny / q (below 1e-7 unless
// q < 1e-30, i.e. pairs far closer than FP32 positions of order 1 resolve).
//
// Issue rates on gfx950 shape the arithmetic: FP32 add/sub/mul/fma whose
// operands are all VGPRs issue at twice the rate of the same instruction with
// an SGPR or literal operand, or of v_min/v_max, and v_rsq_f32 issues at half
// that single rate.  So every per-pair operation is a VGPR-only add/sub/mul/fma
// (12 of them) plus one v_rsq_f32: the softening and `tiny` are copied into
// VGPRs once, the coincident-pair guard is an FMA rather than a clamp, and the
// rsq of four pairs are issued together.
//
// Summation: per LDS tile of kTile sources each target sums in plain FP32
// registers; the tile sums are added into a compensated (Kahan) accumulator,
// so rounding grows with the number of tiles, not of particles.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/nbody.h"
#include "direct.hpp"

namespace nbody {
namespace core {

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

// Sums of the accelerations on the K targets (xi, yi, zi) of this lane from
// the sources src[j0, j1) (j1 - j0 a multiple of kTile), streamed through the
// workgroup's LDS tile; the next tile is prefetched into registers while the
// current one is consumed.  Every thread of the workgroup must call it.
template <int LAW, int K, int BLOCK>
__device__ __forceinline__ void block_forces(float4* tile, const float4* __restrict__ src, const int j0, const int j1,
                                             const float (&xi)[K], const float (&yi)[K], const float (&zi)[K],
                                             const float soft, const float tiny, float (&sx)[K], float (&sy)[K],
                                             float (&sz)[K]) {
    static_assert(kTile % BLOCK == 0, "tile must be a multiple of the block");
    constexpr int PER = kTile / BLOCK;
    const int tid = threadIdx.x;
    float cx[K], cy[K], cz[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        sx[k] = sy[k] = sz[k] = 0.f;
        cx[k] = cy[k] = cz[k] = 0.f;
    }
    float4 pre[PER];
#pragma unroll
    for (int l = 0; l < PER; ++l) pre[l] = src[j0 + tid + l * BLOCK];

    for (int jt = j0; jt < j1; jt += kTile) {
        __syncthreads();
#pragma unroll
        for (int l = 0; l < PER; ++l) tile[tid + l * BLOCK] = pre[l];
        __syncthreads();
        if (jt + kTile < j1) {
#pragma unroll
            for (int l = 0; l < PER; ++l) pre[l] = src[jt + kTile + tid + l * BLOCK];
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

}

}  // namespace core
}  // namespace nbody
