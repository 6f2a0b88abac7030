// Direct-sum force pass: launch interface shared by the single-system
// simulator, the batch executor and the autotuner.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace nbody {

// j-particles staged in LDS per tile; particle arrays are padded to a
// multiple of this with massless dummies at the origin.
constexpr int kTile = 256;

// Launch shape of the direct kernel: targets per lane, threads per
// workgroup, and number of j-chunks each target's sum is split into.
struct DirectShape {
    int i_per_lane;  // 1, 2, 4 or 8
    int block;       // 64, 128 or 256
    int split;       // >= 1; > 1 adds a fixed-order reduction pass
};

// Powers of two in which a system's pair arithmetic runs (exact rescalings):
// positions are multiplied by inv_len = 1 / lambda, masses enter as
// c / (G m) (law R) or G m / c (law P), and the sums come back multiplied by
// out = c / lambda^2.  lambda is the power of two nearest the geometric mean
// of the longest distance in the system (its extent's diagonal) and the
// shortest that matters (the softening length, or FP32's position spacing
// across the extent, 2^-24 of it, if larger); c is the power of two at or
// below the largest G m.  Law R's rsq argument r'^2 (r'^2 + S')^2 (c / G m_j)^2
// grows as r^6, so centring the distances' logarithms keeps it inside FP32's
// normal range for every pair that matters, with room for the system to grow
// or shrink some hundred times; a source lighter than about 1e-11 of the
// heaviest loses its pull beyond that range first (it is then dropped).
struct DirectScale {
    float inv_len;  // 1 / lambda
    float mass_k;   // law R: c; law P: 1 / c
    float soft;     // S / lambda^2 (law R) or eps^2 / lambda^2 (law P)
    float out;      // c / lambda^2
};

// extent: the largest coordinate range of the system's positions (0 for a
// single point); max_gm: its largest |G m| (0 when massless).
DirectScale direct_scale_for(int law, double softening, double extent, double max_gm);

struct DirectLaunch {
    const float4* src;  // n_pad records (x, y, z, G*m): positions at step start
    float4* part;       // split > 1 or no kick: partial sums [split][n_pad] (scaled units)
    float4* pos_out;    // kick: positions after the step (other buffer)
    float4* vel;        // kick: velocities, updated in place
    float4* acc;        // accelerations of this evaluation (n_pad)
    int n;              // real particles
    int n_pad;          // padded count, multiple of kTile
    int law;            // nbody_law
    DirectScale sc;     // the system's scaled units
    float dt;           // kick step
    float dts;          // drift factor dt / scale
};

bool direct_shape_valid(const DirectShape& s);


// Effective number of j-chunks for a shape at this size (never an empty chunk).
int direct_effective_split(const DirectShape& s, int n_pad);

// Heuristic shape: enough workgroups to cover all CUs several times.
DirectShape direct_default_shape(int n_pad);

// One force evaluation.  kick = true advances positions/velocities (kick then
// drift) and writes acc; kick = false only writes acc.  With an effective
// split of 1 and kick the whole step is one fused kernel; otherwise the force
// kernel writes partial sums and a second kernel adds them in chunk order.
hipError_t direct_step(const DirectLaunch& L, const DirectShape& s, bool kick, hipStream_t stream,
                       hipEvent_t ev_force_begin, hipEvent_t ev_force_end);

// Accelerations at m points tgt[] (float4, w ignored) from the n_src_pad
// sources src[] (padded with massless dummies) in the sources' scaled units:
// out[i], float4 (ax, ay, az, 0).
hipError_t direct_field(const float4* src, int n_src_pad, const float4* tgt, int m, int law, const DirectScale& sc,
                        float4* out, hipStream_t stream);

// The same in a caller's staging array of 2 * cap float4 records, without any
// allocation: on entry the second half holds the 3m packed floats of the
// points; the first half receives them as float4 targets, the second half
// their accelerations, and on return the first half holds the 3m packed floats
// of the accelerations (m <= cap).  Three stream-ordered kernels, each reading
// one half and writing the other.
hipError_t direct_field_staged(const float4* src, int n_src_pad, float4* staging, int cap, int m, int law,
                               const DirectScale& sc, hipStream_t stream);

}  // namespace nbody
