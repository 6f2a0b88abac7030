// Device core of the direct sum, shared by the single-system kernel and the
// batch executor so that a system's forces are the same bits either way.
//
// Scaled units.  Each system carries two powers of two (DirectScale): a length
// lambda and a mass scale c.  Sources and targets are multiplied by 1/lambda
// when they are loaded (exact: a power of two), so d' = d / lambda exactly and
// every rounding of the pair arithmetic is the unscaled one with a shifted
// exponent; lambda is chosen from the system's extent so that the squared
// distances sit in the middle of FP32's range whatever the units.  The sums
// are multiplied by out = c / lambda^2 at the end (exact again).
//
// Law R (reference, src/Sim/Physics.hpp:25-35), per source j staged in LDS as
// (x', y', z', h) and k with h = c / (G m_j), k = S' h (S' = S / lambda^2):
//     r2 = |d'|^2,  t = fma(r2, h, k) = (r2 + S') c / (G m_j),
//     q  = r2 t^2 + tiny,   a' += d' * rsq(q)
// so rsq(q) = (G m_j / c) / (|d'| (r2 + S')): the source's mass is folded into
// the reciprocal square root and the pair costs 12 add/sub/mul/fma and one
// v_rsq_f32 (the unfolded form needs a 13th multiply by G m_j).  A massless
// source (the padding dummies) is staged at (2^100, 2^100, 2^100) with
// h = k = +inf: its r2 overflows to +inf (never 0, so no 0 * inf appears),
// t and q are +inf and rsq = 0 exactly, and d' * 0 = 0.  tiny keeps a
// coincident pair (j == i included) finite: q >= tiny, so d' = 0 makes the
// term exactly 0, as the reference's Normalize of a zero vector does; for
// r2 > 0 it changes q by tiny / q (nothing above q ~ 1e-31).  Masses must be
// >= 0 (the square loses a negative mass's sign).
//
// Law P (Plummer), per source (x', y', z', w) with w = G m_j / c:
//     a' += d' * w * rsq(r2 + eps'^2)^3.
//
// Issue rates on gfx950 shape the arithmetic: FP32 add/sub/mul/fma whose
// operands are all VGPRs issue at twice the rate of the same instruction with
// an SGPR or literal operand, or of v_min/v_max, and v_rsq_f32 issues at half
// that single rate.  So every per-pair operation is a VGPR-only add/sub/mul/fma
// plus one v_rsq_f32: the constants are copied into VGPRs once, the guards are
// FMA operands rather than clamps, and the rsq of four pairs are issued
// together.
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

// One staged source tile: positions in scaled units with the law's per-source
// factor (h for law R, w for law P) in .w, and law R's k = S' h in a second
// array (so that four sources' k come in one 16-byte read).
struct SrcTile {
    float4 p[kTile];
    float k[kTile];
};

// A source's mass factor: h = c / (G m) for law R (+inf when massless), or
// w = G m / c for law P.
template <int LAW>
__device__ __forceinline__ float mass_factor(const float gm, const DirectScale& sc) {
    if constexpr (LAW == NBODY_LAW_R) {
        return gm > 0.f ? __fdiv_rn(sc.mass_k, gm) : __builtin_inff();
    } else {
        return gm * sc.mass_k;
    }
}

// The staged record of a source whose .w already holds its mass factor: law R
// stages a massless source (h = +inf) at (2^100, 2^100, 2^100) with k = +inf.
template <int LAW>
__device__ __forceinline__ void stage_factored(const float4 s, const DirectScale& sc, float4& p, float& k) {
    if constexpr (LAW == NBODY_LAW_R) {
        const bool live = s.w != __builtin_inff();
        const float far = 0x1p100f;
        p.x = live ? s.x * sc.inv_len : far;
        p.y = live ? s.y * sc.inv_len : far;
        p.z = live ? s.z * sc.inv_len : far;
        p.w = s.w;
        k = live ? sc.soft * s.w : __builtin_inff();
    } else {
        p.x = s.x * sc.inv_len;
        p.y = s.y * sc.inv_len;
        p.z = s.z * sc.inv_len;
        p.w = s.w;
        k = 0.f;
    }
}

// The staged record of a source (x, y, z, G m) of this system.
template <int LAW>
__device__ __forceinline__ void stage_source(const float4 s, const DirectScale& sc, float4& p, float& k) {
    stage_factored<LAW>(make_float4(s.x, s.y, s.z, mass_factor<LAW>(s.w, sc)), sc, p, k);
}

// Sums of the accelerations on the K targets (xi, yi, zi) of this lane, in
// scaled units (multiply by sc.out), from the sources src[j0, j1) (j1 - j0 a
// multiple of kTile).  The tiles are staged through the workgroup's two LDS
// buffers in turn: the next tile is loaded from memory into registers while
// the current one is consumed, and written to the other buffer, so one
// barrier per tile separates the writes from the reads.  Every thread of the
// workgroup must call it; targets are already in scaled units.
template <int LAW, int K, int BLOCK>
__device__ __forceinline__ void block_forces(SrcTile* tiles, const float4* __restrict__ src, const int j0, const int j1,
                                             const DirectScale& sc, const float (&xi)[K], const float (&yi)[K],
                                             const float (&zi)[K], float (&sx)[K], float (&sy)[K], float (&sz)[K]) {
    static_assert(kTile % BLOCK == 0, "tile must be a multiple of the block");
    constexpr int PER = kTile / BLOCK;
    const int tid = threadIdx.x;
    const float tiny = in_vgpr(__FLT_MIN__);
    const float soft = in_vgpr(sc.soft);
    float cx[K], cy[K], cz[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        sx[k] = sy[k] = sz[k] = 0.f;
        cx[k] = cy[k] = cz[k] = 0.f;
    }
    float4 pre[PER];
#pragma unroll
    for (int l = 0; l < PER; ++l) pre[l] = src[j0 + tid + l * BLOCK];
#pragma unroll
    for (int l = 0; l < PER; ++l) stage_source<LAW>(pre[l], sc, tiles[0].p[tid + l * BLOCK], tiles[0].k[tid + l * BLOCK]);
    __syncthreads();

    int buf = 0;
    for (int jt = j0; jt < j1; jt += kTile) {
        const bool more = jt + kTile < j1;
        if (more) {
#pragma unroll
            for (int l = 0; l < PER; ++l) pre[l] = src[jt + kTile + tid + l * BLOCK];
        }
        const SrcTile& T = tiles[buf];

        float ax[K], ay[K], az[K];
#pragma unroll
        for (int k = 0; k < K; ++k) ax[k] = ay[k] = az[k] = 0.f;

        // Groups of four sources: four rsq arguments per target, then one
        // grouped rsq4 per target, then the accumulation.
#pragma unroll NBODY_GROUP_UNROLL
        for (int j = 0; j < kTile; j += 4) {
            float4 pj[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) pj[g] = T.p[j + g];
            const float4 kj = *reinterpret_cast<const float4*>(&T.k[j]);
            const float kg[4] = {kj.x, kj.y, kj.z, kj.w};
#pragma unroll
            for (int k = 0; k < K; ++k) {
                float dx[4], dy[4], dz[4], q[4];
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    dx[g] = pj[g].x - xi[k];
                    dy[g] = pj[g].y - yi[k];
                    dz[g] = pj[g].z - zi[k];
                    if constexpr (LAW == NBODY_LAW_R) {
                        const float r2 = __builtin_fmaf(dz[g], dz[g], __builtin_fmaf(dy[g], dy[g], dx[g] * dx[g]));
                        const float t = __builtin_fmaf(r2, pj[g].w, kg[g]);
                        const float u = r2 * t;
                        q[g] = __builtin_fmaf(u, t, tiny);
                    } else {
                        q[g] = __builtin_fmaf(dz[g], dz[g], __builtin_fmaf(dy[g], dy[g], __builtin_fmaf(dx[g], dx[g], soft)));
                    }
                }
                rsq4(q[0], q[1], q[2], q[3]);
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    float f;
                    if constexpr (LAW == NBODY_LAW_R) {
                        f = q[g];
                    } else {
                        f = (pj[g].w * q[g]) * (q[g] * q[g]);
                    }
                    ax[k] = __builtin_fmaf(dx[g], f, ax[k]);
                    ay[k] = __builtin_fmaf(dy[g], f, ay[k]);
                    az[k] = __builtin_fmaf(dz[g], f, az[k]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            kahan_add(sx[k], cx[k], ax[k]);
            kahan_add(sy[k], cy[k], ay[k]);
            kahan_add(sz[k], cz[k], az[k]);
        }
        if (more) {
            // The other buffer was last read before the previous barrier.
#pragma unroll
            for (int l = 0; l < PER; ++l)
                stage_source<LAW>(pre[l], sc, tiles[buf ^ 1].p[tid + l * BLOCK], tiles[buf ^ 1].k[tid + l * BLOCK]);
            __syncthreads();
            buf ^= 1;
        }
    }
}

}  // namespace core
}  // namespace nbody
