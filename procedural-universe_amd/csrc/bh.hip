// Barnes-Hut gravity on gfx950.
//
// Replaces the reference's BarnesHut::Update / Exec (src/Sim/BarnesHut.cpp:44-110)
// and its pointer octree (src/Sim/Octree.cpp:16-145), which is rebuilt serially
// by per-particle insertion every step.  Here every stage is a device pass:
//
//  1. root cell      fixed cube (the reference's [-4000, 4000)^3) or the
//                    bounding box grown to a power-of-two cube (two-level
//                    min/max reduction);
//  2. Morton keys    63 bits, 21 levels, x in the lowest bit of each 3-bit
//                    group, which is the reference's child order x + 2y + 4z
//                    (Octree.cpp:25-41).  The cell index along an axis is
//                    floor((x - lo) / cell) computed exactly: x is scaled by a
//                    power of two (exact), floored, and divided by the odd part
//                    of the root size in integers, so keys equal the CPU's bit
//                    for bit and the cells are the reference's half-open cubes;
//  3. sort           least-significant-digit radix sort of (key, index), eight
//                    8-bit passes of histogram / scan / scatter; the scatter
//                    ranks equal digits with wave ballots, orders each tile by
//                    digit in LDS and writes it out in runs, so the order is
//                    stable and equals std::stable_sort of the keys;
//  4. radix tree     binary radix tree over the sorted keys (Karras 2012);
//                    a node whose common prefix reaches a deeper 3-bit group
//                    than its parent's is an octree cell ("kept"), the others
//                    are the binary splits inside one cell; identical keys
//                    form a level-21 bucket;
//  5. mass           FP64 mass and mass-weighted position of every node as a
//                    difference of double-double prefix sums over the sorted
//                    particles (a node is a contiguous range of them); then
//                    every cell writes one 192-byte record of its children
//                    (at most eight, in key order): sources, opening
//                    thresholds and codes, all the traversal reads of it;
//  6. traversal      one lane per particle in Morton order; a wave walks the
//                    octree with one stack in LDS of (cell, lanes) entries,
//                    up to eight cells per step of the walk.
//                    Each lane makes the reference's own decision at each cell
//                    (Octree::CalculateForce, Octree.cpp:107-145): approximate
//                    with the cell's monopole iff side / |p - COM| < theta,
//                    otherwise open it; a cell is pushed with the mask of the
//                    lanes that opened it, so the wave shares the walk while
//                    every particle gets exactly its own set of interactions.
//                    Single particles are exact pair forces (skipping self);
//                    an approximated cell that contains p includes p's own mass,
//                    as in the reference.  Kick and drift are fused.
//
// A compressed octree gives the reference's forces: along a chain of cells
// with one non-empty child the mass and centre of mass do not change and the
// side shrinks, so the reference approximates somewhere on the chain iff it
// approximates at the chain's deepest cell, which is the cell kept here.
//
// Wave priorities (s_setprio, within each kernel): the build's kernels 2, the
// traversal's walk 1 and its list evaluation 0, so that where several
// systems share the device a short build launch or a walk's dependent loads
// issue ahead of long runs of list arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "bh.hpp"
#include "common.hpp"
#include "direct.hpp"
#include "direct_core.hpp"
#include "sim.hpp"

// v_writelane_b32 through the compiler's own intrinsic (this clang has no
// __builtin_amdgcn_writelane): the compiler places the lane select in M0,
// which it reserves and sets itself, and inserts the wait states its hazard
// model requires between the ballot that produces the value and the write.
__device__ int nb_writelane(int value, int lane, int old) __asm("llvm.amdgcn.writelane");

namespace nbody {

namespace {

constexpr int kLevels = 21;
constexpr int kBucketLevel = kLevels;
constexpr int kSortThreads = 256;
// Keys per thread of a sort tile (tile = 256 x items keys): small systems take
// small tiles for more workgroups per pass, large ones large tiles for fewer
// tile counts to scan (measured tree builds: 0.19 / 0.21 / 0.24 ms at 65,536
// particles with 2 / 4 / 8 items, 1.66 / 1.43 / 1.30 ms at 4,194,304).
inline int sort_items(size_t n) { return n <= (size_t(1) << 17) ? 2 : n <= (size_t(1) << 20) ? 4 : 8; }
// The most tiles of any system of at most n particles (the tile grows with
// the size, so a smaller system can have more tiles than a larger one): the
// count buffer of a handle holds that many, whatever size it runs later.
inline size_t sort_max_tiles(size_t n) {
    const auto tiles = [](size_t k, size_t tile) { return (k + tile - 1) / tile; };
    size_t t = tiles(std::min(n, size_t(1) << 17), kSortThreads * 2);
    t = std::max(t, tiles(std::min(n, size_t(1) << 20), kSortThreads * 4));
    return std::max(t, tiles(n, kSortThreads * 8));
}
constexpr int kTravThreads = 64;  // one wavefront per workgroup: the finest tail
constexpr int kTravWaves = kTravThreads / 64;
// Pending (cell, lanes) entries per wave.  Popping one cell at a time (a
// depth-first walk) needs at most 7 * 21 + 1 entries above the cell it starts
// from: seven pending siblings per level below it and the last cell's eight
// children.  A batch of k cells replaces them by at most 8k children, so a
// batch is only as large as keeps that reserve free afterwards, and the stack
// cannot overflow (the checked push that reports overflow is a safeguard).
constexpr int kStackCap = 288;
constexpr int kDfsReserve = 7 * 22;
constexpr int kEmpty = -1;

// Child encoding: >= 0 internal node; <= -2 particle at sorted position (-c - 2).
__host__ __device__ inline int enc_leaf(int k) { return -k - 2; }
__host__ __device__ inline int dec_leaf(int c) { return -c - 2; }

// A level-21 cell of identical keys is marked in its code; particle counts
// stay below 2^30.
constexpr int kBucketFlag = 1 << 30;

// The eight children of an octree cell, everything the traversal needs to
// decide about them: one 192-byte record read with three 64-byte scalar loads
// when the cell is opened.  A child cell holds its centre of mass, the pair
// arithmetic's mass factor of G M (core::mass_factor) and its opening
// threshold (below); a particle its position and the factor of G m; codes are
// kEmpty, enc_leaf(sorted position), or the child cell's node (| kBucketFlag).
struct alignas(64) ChildBlock {
    float4 src[8];
    float thr[8];
    int code[8];
};
static_assert(sizeof(ChildBlock) == 192, "ChildBlock layout");

struct RootParams {
    float lo[3];
    float size;    // edge length of the root cube
    float scale2;  // 2^(21 - p) with size = odd * 2^p
    int odd;       // odd part of the size (1 for a power-of-two root)
    int pad;
    long long base[3];  // floor_div(floor(lo * scale2), odd)
};

__host__ __device__ inline long long floor_div_ll(long long a, long long b) {
    long long q = a / b;
    if ((a % b != 0) && (a < 0)) --q;
    return q;
}

__device__ inline unsigned long long spread3(unsigned long long v) {
    v &= 0x1fffffULL;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

// Exact floor((x - lo) / cell) for a root corner that is a multiple of the cell.
// floor(k / odd) for |k| <= 2^40 and 1 <= odd < 2^24: a dynamic root's odd
// part is 1 (no division); otherwise the quotient rounded from a double
// product (within 2^-12 of the exact one) corrected by the exact remainder,
// in place of a 64-bit integer division.
__device__ inline long long floor_div_odd(long long k, int odd, double inv_odd) {
    if (odd == 1) return k;
    long long q = static_cast<long long>(floor(static_cast<double>(k) * inv_odd));
    const long long r = k - q * odd;
    if (r < 0) --q;
    else if (r >= odd) ++q;
    return q;
}

__device__ inline long long cell_index(float x, const RootParams& R, int a, double inv_odd) {
    float X = x * R.scale2;  // power-of-two scaling: exact
    X = fmaxf(fminf(X, 0x1p40f), -0x1p40f);
    long long k = static_cast<long long>(floorf(X));
    if (X == 0.f && x < 0.f) k = -1;
    return floor_div_odd(k, R.odd, inv_odd) - R.base[a];
}

// ---- opening thresholds ------------------------------------------------------
//
// The reference approximates a cell of side d seen from p iff
// d / |p - COM| < theta (Octree.cpp:122-124), the quotient and the length
// (sqrtf of the squared distance r2) in FP32 and the comparison against the
// double theta, i.e. fl(d / fl(sqrt(r2))) <= theta_lo with theta_lo the
// largest float below theta.  Correctly rounded sqrt and division are
// monotone, so the test is monotone in r2: false at 0 (d / 0 = inf), true at
// +inf (d / inf = 0) when theta_lo >= 0.  The smallest float r2 for which it
// holds is found by bisection over the ordered bit patterns of positive
// floats, and the traversal's test becomes r2 >= thr, exactly the same
// decision.  The side depends only on the level, so one threshold per level.
// The traversal evaluates it as !(r2 < thr), which also holds for a NaN
// distance (a NaN position then reaches the sums instead of being dropped).
// When no finite r2 passes (theta <= 0) the threshold is +inf: only an
// infinite r2 passes, where every pair and every monopole contributes
// exactly 0 to the acceleration, as the reference's opened cell does.
constexpr int kThrLevels = kBucketLevel + 1;

__device__ float open_threshold(float side, float theta_lo) {
    auto approx = [&](unsigned bits) {
        return __fdiv_rn(side, __fsqrt_rn(__uint_as_float(bits))) <= theta_lo;
    };
    const unsigned kInf = 0x7f800000u;
    if (!approx(kInf)) return __uint_as_float(kInf);  // no finite r2 passes
    unsigned lo = 0u, hi = kInf;                               // !approx(lo), approx(hi)
    while (hi - lo > 1u) {
        const unsigned mid = lo + (hi - lo) / 2u;
        if (approx(mid)) hi = mid; else lo = mid;
    }
    return __uint_as_float(hi);
}

// Thresholds of a fixed root: computed once.
__global__ void open_threshold_kernel(const RootParams* __restrict__ R, float theta_lo, float* __restrict__ thr) {
    const int L = threadIdx.x;
    if (L < kThrLevels) thr[L] = open_threshold(ldexpf(R->size, -L), theta_lo);
}

// Every kernel of the tree build raises its wave priority above the
// traversal walk's (1).  When systems are served together, a system's short,
// dependent build launches then issue ahead of other systems' traversals on a
// SIMD they share instead of waiting behind them (a served stream +3%); a
// build that has the device to itself overtakes nothing.
__device__ __forceinline__ void build_priority() { __builtin_amdgcn_s_setprio(2); }

// ---- 1. root -------------------------------------------------------------

__global__ __launch_bounds__(256) void bounds_kernel(const float4* __restrict__ pos, int n,
                                                     float* __restrict__ partial) {
    build_priority();
    __shared__ float red[6][256];
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float4 p = pos[i];
        mn[0] = fminf(mn[0], p.x), mn[1] = fminf(mn[1], p.y), mn[2] = fminf(mn[2], p.z);
        mx[0] = fmaxf(mx[0], p.x), mx[1] = fmaxf(mx[1], p.y), mx[2] = fmaxf(mx[2], p.z);
    }
    for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = mn[a];
        red[3 + a][threadIdx.x] = mx[a];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w)
            for (int a = 0; a < 3; ++a) {
                red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + w]);
                red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Power-of-two cube over the bounding box with a corner on its level-21 grid
// (same arithmetic as the CPU's oracle_dynamic_root), then the opening
// threshold of every level of that root.  One workgroup of 256 threads.
__global__ __launch_bounds__(256) void dynamic_root_kernel(const float* __restrict__ partial, int nblocks,
                                                           RootParams* __restrict__ R, float theta_lo,
                                                           float* __restrict__ thr) {
    build_priority();
    __shared__ float red[6][256];
    __shared__ float size_sh;
    float pm[6] = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX};  // min x,y,z and -max x,y,z
    for (int b = threadIdx.x; b < nblocks; b += 256)
        for (int a = 0; a < 3; ++a) {
            pm[a] = fminf(pm[a], partial[b * 6 + a]);
            pm[3 + a] = fminf(pm[3 + a], -partial[b * 6 + 3 + a]);
        }
    for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = pm[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w)
            for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = fminf(red[c][threadIdx.x], red[c][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float mn[3] = {red[0][0], red[1][0], red[2][0]};
        const float mx[3] = {-red[3][0], -red[4][0], -red[5][0]};
        float ext = 0.f;
        for (int a = 0; a < 3; ++a) ext = fmaxf(ext, mx[a] - mn[a]);
        int e;
        frexpf(fmaxf(ext, 0x1p-100f), &e);
        float size = ldexpf(1.0f, e);
        RootParams out{};
        for (int tries = 0; tries < 4; ++tries) {
            const float cell = ldexpf(size, -kLevels);
            bool fits = true;
            for (int a = 0; a < 3; ++a) {
                const long long k0 = static_cast<long long>(floorf(mn[a] / cell));
                const long long k1 = static_cast<long long>(floorf(mx[a] / cell));
                out.lo[a] = floorf(mn[a] / cell) * cell;
                out.base[a] = k0;
                if (k1 - k0 >= (1LL << kLevels)) fits = false;
            }
            if (fits) break;
            size *= 2.f;
        }
        out.size = size;
        out.scale2 = ldexpf(1.0f, kLevels) / size;
        out.odd = 1;
        *R = out;
        size_sh = size;
    }
    __syncthreads();
    if (threadIdx.x < kThrLevels) thr[threadIdx.x] = open_threshold(ldexpf(size_sh, -threadIdx.x), theta_lo);
}

// ---- 2. keys ---------------------------------------------------------------

__global__ __launch_bounds__(256) void keys_kernel(const float4* __restrict__ pos, int n,
                                                   const RootParams* __restrict__ Rp,
                                                   unsigned long long* __restrict__ keys,
                                                   unsigned* __restrict__ idx, unsigned* __restrict__ outside) {
    build_priority();
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const RootParams R = *Rp;
    const double inv_odd = 1.0 / static_cast<double>(R.odd);
    const float4 p = pos[i];
    const float c[3] = {p.x, p.y, p.z};
    unsigned long long key = 0;
    bool out = false;
    for (int a = 0; a < 3; ++a) {
        long long k = cell_index(c[a], R, a, inv_odd);
        if (k < 0) k = 0, out = true;
        if (k > 0x1fffff) k = 0x1fffff, out = true;
        key |= spread3(static_cast<unsigned long long>(k)) << a;
    }
    keys[i] = key;
    idx[i] = static_cast<unsigned>(i);
    if (out) atomicAdd(outside, 1u);
}

// ---- 3. radix sort -----------------------------------------------------------

// Digits of kRadixBits bits: kRadixPasses passes cover the 63-bit keys.
// Nine-bit digits (seven passes) were slower from 131,072 keys up: a tile's
// digit runs are half as long, so the scatter's writes coalesce less
// (46.5 against 33 us per pass at 4,194,304 keys).
constexpr int kRadixBits = 8;
constexpr int kRadixBins = 1 << kRadixBits;
constexpr unsigned kRadixMask = kRadixBins - 1;
constexpr int kRadixPasses = (63 + kRadixBits - 1) / kRadixBits;
constexpr int kDigitsPerThread = kRadixBins / kSortThreads;  // the digit phases' digits per thread
static_assert(kRadixBins % kSortThreads == 0, "whole digits per thread");

// Digit counts of each tile, digit-major (hist[d * nblocks + tile]).
template <int ITEMS>
__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const unsigned long long* __restrict__ keys, int n,
                                                                  int shift, unsigned* __restrict__ hist,
                                                                  int nblocks) {
    build_priority();
    __shared__ unsigned h[kRadixBins];
    for (int d = threadIdx.x; d < kRadixBins; d += kSortThreads) h[d] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortThreads * ITEMS;
    for (int k = 0; k < ITEMS; ++k) {
        const int i = base + k * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & kRadixMask], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kRadixBins; d += kSortThreads) hist[d * nblocks + blockIdx.x] = h[d];
}

constexpr int kScanThreads2 = 256;
constexpr int kScanPer = 16;

__device__ inline unsigned block_exclusive_scan_u32(unsigned v, unsigned* sh, unsigned* total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads2; off <<= 1) {
        const unsigned a = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    const unsigned incl = sh[t];
    if (total) *total = sh[kScanThreads2 - 1];
    __syncthreads();
    return incl - v;
}

// Workgroup d replaces digit d's row of tile counts by its exclusive scan
// (the tile's offset among the keys with digit d) and writes the row's total
// to totals[d]; the scatter adds the totals of the smaller digits.
__global__ __launch_bounds__(kScanThreads2) void radix_rowscan_kernel(unsigned* __restrict__ hist, int nblocks,
                                                                      unsigned* __restrict__ totals) {
    build_priority();
    __shared__ unsigned sh[kScanThreads2];
    const int d = blockIdx.x, t = threadIdx.x;
    unsigned* row = hist + static_cast<size_t>(d) * nblocks;
    unsigned carry = 0;
    for (int base = 0; base < nblocks; base += kScanThreads2 * kScanPer) {
        const int at = base + t * kScanPer;  // a contiguous run per thread
        unsigned v[kScanPer];
        unsigned sum = 0;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            v[k] = at + k < nblocks ? row[at + k] : 0u;
            sum += v[k];
        }
        unsigned chunk = 0;
        unsigned run = carry + block_exclusive_scan_u32(sum, sh, &chunk);
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            if (at + k < nblocks) row[at + k] = run;
            run += v[k];
        }
        carry += chunk;
    }
    if (t == 0) totals[d] = carry;
}

// Counters: walk[3] (interactions, list entries, cells opened) and
// counters[2] (outside, overflow) on the device, mirrored after every force
// pass into pinned host memory (host[5], same order), so reading them needs
// no copy: a copy is a blit kernel that would queue behind whatever occupies
// the device.
__global__ void zero_counters_kernel(unsigned* __restrict__ counters, unsigned long long* __restrict__ walk,
                                     unsigned long long* __restrict__ host) {
    if (threadIdx.x < 2) counters[threadIdx.x] = 0u;
    if (threadIdx.x < 3) walk[threadIdx.x] = 0ULL;
    if (threadIdx.x < 5) host[threadIdx.x] = 0ULL;
}

__global__ void publish_counters_kernel(const unsigned* __restrict__ counters,
                                        const unsigned long long* __restrict__ walk,
                                        unsigned long long* __restrict__ host) {
    if (threadIdx.x < 3) host[threadIdx.x] = walk[threadIdx.x];
    if (threadIdx.x >= 3 && threadIdx.x < 5) host[threadIdx.x] = counters[threadIdx.x - 3];
}

// One radix pass over a tile of 256 x ITEMS keys, stable.  Each wavefront owns
// a contiguous share of the tile (ITEMS rounds of 64 consecutive keys) and
// ranks every key among the earlier keys of its wavefront with the same digit
// (the lanes with equal digits found by kRadixBits ballots, running per-digit
// counts in LDS; no workgroup barrier).  One pass over the digits then gives each
// wavefront's offset within the tile's digit run and the runs' starts; the
// keys are placed in LDS in digit order and written out in that order, so
// consecutive threads write consecutive addresses within a digit's run of the
// output.
//
// FUSED: `counts` holds the raw per-tile digit counts ([digit * nblocks +
// tile], at most kFusedScanTiles tiles) and each workgroup forms its own
// output places: thread d sums digit d's row before its tile and in total
// (no separate scan launch).  Otherwise `counts` holds each tile's offset
// among the keys of its digit and `totals` each digit's count (the row scan).
// Either way one scan over the digits' totals gives the digits' starts.
// GATHER (the last pass): the sorted positions are written beside the keys.
constexpr int kSortWaves = kSortThreads / 64;

// Above this many tiles the row scan's extra launch costs less than every
// workgroup summing the digits' rows itself (measured: equal at 16 tiles,
// 0.01 ms less at 32, 0.04 at 64, 0.24 at 256).
constexpr int kFusedScanTiles = 16;

template <bool FUSED, bool GATHER, int ITEMS>
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const unsigned long long* __restrict__ keys_in, const unsigned* __restrict__ idx_in,
    unsigned long long* __restrict__ keys_out, unsigned* __restrict__ idx_out, int n, int shift,
    const unsigned* __restrict__ counts, const unsigned* __restrict__ totals, int nblocks,
    const float4* __restrict__ pos_in, float4* __restrict__ pos_out) {
    build_priority();
    constexpr int kSortTile = kSortThreads * ITEMS;
    constexpr int kSortPerWave = 64 * ITEMS;  // a wavefront's share: ITEMS rounds of 64 keys
    __shared__ unsigned cnt[kSortWaves][kRadixBins];  // per-wave digit counts, then per-wave offsets in the run
    __shared__ unsigned run_start[kRadixBins];        // the tile's digit runs
    __shared__ unsigned dst_base[kRadixBins];         // the runs' places in the output
    __shared__ unsigned wave_sum[2][kSortWaves];
    __shared__ unsigned long long skeys[kSortTile];
    __shared__ unsigned sidx[kSortTile];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const unsigned long long below = (1ULL << lane) - 1ULL;
    const int base = blockIdx.x * kSortTile;
    const int wbase = base + wave * kSortPerWave;
#pragma unroll
    for (int d = lane; d < kRadixBins; d += 64) cnt[wave][d] = 0u;  // this wave's counters only
    unsigned long long key[ITEMS];
    unsigned idx[ITEMS], rank[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const int i = wbase + r * 64 + lane;
        const bool valid = i < n;
        key[r] = valid ? keys_in[i] : 0ULL;
        idx[r] = valid ? idx_in[i] : 0u;
        const unsigned digit = static_cast<unsigned>((key[r] >> shift) & kRadixMask);
        unsigned long long peers = __builtin_amdgcn_ballot_w64(valid);
#pragma unroll
        for (int bit = 0; bit < kRadixBits; ++bit) {
            const bool set = (digit >> bit) & 1u;
            const unsigned long long bm = __builtin_amdgcn_ballot_w64(valid && set);
            peers &= set ? bm : ~bm;
        }
        const unsigned within = static_cast<unsigned>(__popcll(peers & below));
        const unsigned prior = valid ? cnt[wave][digit] : 0u;  // read by every lane before the leader's update
        rank[r] = prior + within;
        __builtin_amdgcn_wave_barrier();
        if (valid && within == 0) cnt[wave][digit] = prior + static_cast<unsigned>(__popcll(peers));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    {
        // Thread tid's kDigitsPerThread consecutive digits: offsets of the
        // wavefronts' shares within each digit's run and the run's length,
        // then the runs' starts (an exclusive scan over digits)
        constexpr int D = kDigitsPerThread;
        unsigned run[D], before[D], total[D];
        unsigned rsum = 0, tsum = 0;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const int d = tid * D + j;
            unsigned acc = 0;
#pragma unroll
            for (int w = 0; w < kSortWaves; ++w) {
                const unsigned c = cnt[w][d];
                cnt[w][d] = acc;
                acc += c;
            }
            run[j] = acc;
            // digit d's place in the output: the digit's start (the totals of
            // the smaller digits) plus the tile's offset among the digit's
            // keys, both formed here from the raw counts or given by the row
            // scan
            before[j] = 0;
            total[j] = 0;
            if constexpr (FUSED) {
                const unsigned* row = counts + static_cast<size_t>(d) * nblocks;
                for (int b = 0; b < nblocks; ++b) {
                    const unsigned c = row[b];
                    before[j] += b < static_cast<int>(blockIdx.x) ? c : 0u;
                    total[j] += c;
                }
            } else {
                before[j] = counts[d * nblocks + blockIdx.x];
                total[j] = totals[d];
            }
            rsum += run[j];
            tsum += total[j];
        }
        unsigned x = rsum, y = tsum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned xs = __shfl_up(x, off, 64);
            const unsigned ys = __shfl_up(y, off, 64);
            if (lane >= off) {
                x += xs;
                y += ys;
            }
        }
        if (lane == 63) {
            wave_sum[0][wave] = x;
            wave_sum[1][wave] = y;
        }
        __syncthreads();
        unsigned start = x - rsum, digit_start = y - tsum;
        for (int w = 0; w < wave; ++w) {
            start += wave_sum[0][w];
            digit_start += wave_sum[1][w];
        }
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const int d = tid * D + j;
            run_start[d] = start;
            dst_base[d] = digit_start + before[j];
            start += run[j];
            digit_start += total[j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const int i = wbase + r * 64 + lane;
        if (i < n) {
            const unsigned digit = static_cast<unsigned>((key[r] >> shift) & kRadixMask);
            const unsigned at = run_start[digit] + cnt[wave][digit] + rank[r];
            skeys[at] = key[r];
            sidx[at] = idx[r];
        }
    }
    __syncthreads();
    const int count = min(kSortTile, n - base);
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const int at = r * kSortThreads + tid;
        if (at < count) {
            const unsigned long long k = skeys[at];
            const unsigned digit = static_cast<unsigned>((k >> shift) & kRadixMask);
            const unsigned dst = dst_base[digit] + (static_cast<unsigned>(at) - run_start[digit]);
            keys_out[dst] = k;
            idx_out[dst] = sidx[at];
            if constexpr (GATHER) pos_out[dst] = pos_in[sidx[at]];
        }
    }
}

// ---- 4. radix tree -------------------------------------------------------------

__device__ inline int delta(const unsigned long long* __restrict__ keys, int n, int a, int b) {
    if (b < 0 || b >= n) return -1;
    const unsigned long long ka = keys[a], kb = keys[b];
    if (ka == kb) return 64 + __clz(static_cast<unsigned>(a ^ b));
    return __clzll(ka ^ kb);
}

// Cell level of a binary node with common-prefix length d (keys use 63 bits,
// the top bit is always 0): the number of whole 3-bit groups in common.
__device__ inline int level_of(int d) { return d >= 64 ? kBucketLevel : (d - 1) / 3; }

// A node's level and whether it is an octree cell ("kept": it starts a
// deeper cell than its parent, or it is the root), in one word.
constexpr int kLevelMask = 0xff;
constexpr int kKeptBit = 1 << 8;
__host__ __device__ inline int info_level(int info) { return info & kLevelMask; }
__host__ __device__ inline bool info_kept(int info) { return (info & kKeptBit) != 0; }

struct TreeArgs {
    const unsigned long long* keys;  // sorted
    int n;
    int2* child;  // [n-1] encoded (left, right) children
    int* first;   // [n-1]
    int* last;    // [n-1]
    int* info;    // [n-1] level | kKeptBit
};

__device__ void radix_tree_block(const TreeArgs& T, int block) {
    const int i = block * 256 + threadIdx.x;
    const int n = T.n;
    if (i >= n - 1) return;
    const unsigned long long* keys = T.keys;
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    const int lc = (lo == gamma) ? enc_leaf(gamma) : gamma;
    const int rc = (hi == gamma + 1) ? enc_leaf(gamma + 1) : gamma + 1;
    const int lv = level_of(dnode);
    T.child[i] = make_int2(lc, rc);
    T.first[i] = lo;
    T.last[i] = hi;
    // The node writes its internal children's words: a child's common prefix
    // is that of its range's end points, the same delta its own thread finds.
    if (lc >= 0) {
        const int lvc = level_of(delta(keys, n, lo, gamma));
        T.info[lc] = lvc | (lvc > lv ? kKeptBit : 0);
    }
    if (rc >= 0) {
        const int lvc = level_of(delta(keys, n, gamma + 1, hi));
        T.info[rc] = lvc | (lvc > lv ? kKeptBit : 0);
    }
    if (i == 0) T.info[0] = lv | kKeptBit;
}

// Every octree cell below the bucket level writes its own child block: its
// child cells, then its child particles, each in key order (which is octant
// order), packed into the first slots, the remaining codes kEmpty.  A child
// cell gets (COM, G M), its opening threshold and its node; a particle
// (position, G m), threshold 0 (r2 >= 0 always holds) and its code.  The
// children are the first kept nodes and the particles below the cell in the
// radix tree.  A node that is not kept has its parent's level, so its common
// prefix ends inside the same 3-bit group as the cell's; prefixes lengthen
// strictly from parent to child, so at most two such nodes lie on any path
// below the cell and a descent of three binary levels reaches every child.
// Runs after the cells' centres of mass are final; particles of a bucket are
// reached by its range instead.
struct BlockArgs {
    const int2* child;
    const int* info;
    const float4* node_com;
    const float* thr;  // opening threshold per level
    const float4* pos_sorted;
    int m;
    ChildBlock* blocks;
    DirectScale sc;    // the pair arithmetic's scaled units (mass factors)
};

// A workgroup takes kBlockNodes radix nodes and first gathers the cells among
// them (kept, level below `limit`) into an LDS list in node order, so that
// whole wavefronts work on cells (45% of the nodes on a disc encounter).
// Returns the list's length (uniform).
constexpr int kBlockNodes = 256;
constexpr int kBlockCells = 4;  // cells per lane group in flight together
constexpr int kBlockRounds = kBlockNodes / 256;

__device__ inline int gather_cells(const int* __restrict__ info, int m, int base, int limit, int* list,
                                   int* wave_cnt) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const unsigned long long below = (1ULL << lane) - 1ULL;
    unsigned long long mask[kBlockRounds];
#pragma unroll
    for (int r = 0; r < kBlockRounds; ++r) {
        const int node = base + r * 256 + tid;
        const int w = node < m ? info[node] : 0;
        mask[r] = __builtin_amdgcn_ballot_w64(info_kept(w) && info_level(w) < limit);
        if (lane == 0) wave_cnt[r * 4 + wave] = static_cast<int>(__popcll(mask[r]));
    }
    __syncthreads();
    int total = 0;
#pragma unroll
    for (int k = 0; k < 4 * kBlockRounds; ++k) total += wave_cnt[k];
#pragma unroll
    for (int r = 0; r < kBlockRounds; ++r) {
        int at = 0;
        for (int k = 0; k < r * 4 + wave; ++k) at += wave_cnt[k];
        if ((mask[r] >> lane) & 1ULL) list[at + static_cast<int>(__popcll(mask[r] & below))] = base + r * 256 + tid;
    }
    __syncthreads();
    return total;
}

// Eight lanes per cell.  Lane s of a cell's group follows the path (a, b, c)
// = the bits of s down the binary levels below the cell and takes the child
// at its end; a child met higher up (a kept node or a particle) belongs to
// the first lane of its subtree, the others of that subtree take none.  Each
// level's word and children are loaded together (a radix node always has two
// children, so every load is in range).  Then each lane writes one slot of
// the block: the lanes with child cells the first slots, then those with
// particles, each in lane order (= key order), then the empty ones, so the
// 192 bytes of a block go out in three stores of eight consecutive lanes.
// Every group takes kBlockCells cells at a time (independent load chains in
// flight together).
template <int LAW>
__device__ __forceinline__ void descend(const BlockArgs& A, int i, int s, int& code, int& word) {
    // Without branches: every level's loads are issued (at node 0 where the
    // path has ended; every index is in range), so that the cells' chains
    // overlap, and the results are selected afterwards.
    const int a = s >> 2, b = (s >> 1) & 1, c = s & 1;
    const int2 top = A.child[i];
    const int c1 = a ? top.y : top.x;
    const int n1 = max(c1, 0);
    const int r1 = A.info[n1];
    const int w1 = c1 >= 0 ? r1 : 0;
    const int2 mid = A.child[n1];
    const int c2 = b ? mid.y : mid.x;
    const int n2 = max(c2, 0);
    const int r2 = A.info[n2];
    const int w2 = c2 >= 0 ? r2 : 0;
    const int2 low = A.child[n2];
    const int c3 = c ? low.y : low.x;
    const int r3 = A.info[max(c3, 0)];
    const int w3 = c3 >= 0 ? r3 : 0;
    const bool stop1 = c1 < 0 || info_kept(w1);
    const bool stop2 = c2 < 0 || info_kept(w2);
    // selects, not branches (a branch would take the deeper loads with it)
    const int code2 = stop2 ? (c == 0 ? c2 : kEmpty) : c3;
    const int word2 = stop2 ? (c == 0 ? w2 : 0) : w3;
    code = stop1 ? ((b | c) == 0 ? c1 : kEmpty) : code2;
    word = stop1 ? ((b | c) == 0 ? w1 : 0) : word2;
}

// The slot of lane s's child in its cell's block (see above) and where its
// record is (node 0's for an empty slot, so that every load is in range).
__device__ __forceinline__ void child_slot(const BlockArgs& A, int code, int s, int gshift, int& slot,
                                           const float4*& at) {
    const unsigned below = (1u << s) - 1u;
    const bool found = code != kEmpty;
    const bool cell = found && code >= 0;
    const unsigned gf = static_cast<unsigned>(__builtin_amdgcn_ballot_w64(found) >> gshift) & 0xffu;
    const unsigned gc = static_cast<unsigned>(__builtin_amdgcn_ballot_w64(cell) >> gshift) & 0xffu;
    if (cell) slot = __popc(gc & below);
    else if (found) slot = __popc(gc) + __popc(gf & ~gc & below);
    else slot = __popc(gf) + __popc(~gf & below);
    at = (found && code < 0) ? A.pos_sorted + dec_leaf(code) : A.node_com + (cell ? code : 0);
}

// Writes lane s's slot of cell i's block from the loaded record r and the
// threshold t of the child's level.
template <int LAW>
__device__ __forceinline__ void write_child(const BlockArgs& A, int i, int slot, int code, int word, float4 r,
                                            float t) {
    const bool found = code != kEmpty;
    const bool cell = found && code >= 0;
    // the list's mass factor in place of G m, once per step
    const float4 src =
        found ? make_float4(r.x, r.y, r.z, core::mass_factor<LAW>(r.w, A.sc)) : make_float4(0.f, 0.f, 0.f, 0.f);
    ChildBlock& B = A.blocks[i];
    B.src[slot] = src;
    B.thr[slot] = cell ? t : 0.f;
    B.code[slot] = (cell && info_level(word) >= kBucketLevel) ? (code | kBucketFlag) : code;
}

template <int LAW>
__global__ __launch_bounds__(256) void block_kernel(BlockArgs A) {
    build_priority();
    __shared__ int list[kBlockNodes];
    __shared__ int wave_cnt[4 * kBlockRounds];
    const int total = gather_cells(A.info, A.m, blockIdx.x * kBlockNodes, kBucketLevel, list, wave_cnt);
    const int lane = threadIdx.x & 63, s = threadIdx.x & 7, group = threadIdx.x >> 3;
    const int gshift = lane & ~7;  // the group's bits in a wave's ballot
    for (int k0 = 0; k0 < total; k0 += 32 * kBlockCells) {
        int idx[kBlockCells], code[kBlockCells], word[kBlockCells], slot[kBlockCells];
        bool live[kBlockCells];
        const float4* at[kBlockCells];
#pragma unroll
        for (int u = 0; u < kBlockCells; ++u) {
            const int k = k0 + 32 * u + group;
            live[u] = k < total;
            idx[u] = list[live[u] ? k : 0];
        }
#pragma unroll
        for (int u = 0; u < kBlockCells; ++u) descend<LAW>(A, idx[u], s, code[u], word[u]);
#pragma unroll
        for (int u = 0; u < kBlockCells; ++u) child_slot(A, code[u], s, gshift, slot[u], at[u]);
        // every cell's records and thresholds in flight together
        float4 r[kBlockCells];
        float t[kBlockCells];
#pragma unroll
        for (int u = 0; u < kBlockCells; ++u) {
            r[u] = *at[u];
            t[u] = A.thr[info_level(word[u])];
        }
#pragma unroll
        for (int u = 0; u < kBlockCells; ++u)
            asm volatile("" ::"v"(r[u].x), "v"(r[u].y), "v"(r[u].z), "v"(r[u].w), "v"(t[u]));
#pragma unroll
        for (int u = 0; u < kBlockCells; ++u)
            if (live[u]) write_child<LAW>(A, idx[u], slot[u], code[u], word[u], r[u], t[u]);
    }
}

// ---- 5. FP64 mass and centre of mass ------------------------------------------------
//
// Every radix-tree node covers a contiguous range [first, last] of the sorted
// particles, so its mass and mass-weighted position are differences of
// exclusive prefix sums P[last + 1] - P[first] of (m, m x, m y, m z).  The
// prefix sums are double-double (error-free TwoSum additions, Knuth), so a
// node's FP64 sums are exact to rounding however small the node is next to
// the total, with no atomics and no ordering between nodes.

struct DD {
    double hi, lo;
};

__device__ inline DD dd_add(DD a, DD b) {
    const double s = a.hi + b.hi;
    const double bb = s - a.hi;
    const double e = (a.hi - (s - bb)) + (b.hi - bb) + a.lo + b.lo;
    const double hi = s + e;
    return DD{hi, e - (hi - s)};
}

__device__ inline double dd_diff(DD a, DD b) {  // a - b rounded to double
    const DD nb{-b.hi, -b.lo};
    const DD d = dd_add(a, nb);
    return d.hi + d.lo;
}

constexpr int kMassThreads = 256;
constexpr int kMassPer = 16;    // rows of kMassThreads elements in a chunk, large systems
constexpr int kMassMinRows = 4;  // and in the smallest chunks
constexpr int kMassItems = 4;    // consecutive elements per lane in the prefix pass
static_assert(kMassMinRows % kMassItems == 0 && kMassPer % kMassMinRows == 0, "whole rows per chunk");

constexpr int kMassChunk = kMassThreads * kMassPer;

struct MassArgs {
    const float4* pos_sorted;
    const unsigned* order;  // sorted position -> particle
    const double* mass;     // by particle
    double* mass_sorted;    // by sorted position: written by the chunk sums, read by the prefix pass
    int n;
    DD* prefix;             // [4][n + 1]
    DD* chunk;              // [4][nchunks]
    int nchunks;
    int rows;               // rows of kMassThreads elements per chunk
};

// Element k's mass and mass moments.  The chunk sums (GATHER) read the mass
// through the sorted order and store it in sorted order; the prefix pass then
// reads it contiguously instead of gathering it a second time.
template <bool GATHER>
__device__ inline void leaf_vals(const MassArgs& A, int k, double v[4]) {
    if (k >= A.n) {
        v[0] = v[1] = v[2] = v[3] = 0.0;
        return;
    }
    double m;
    if constexpr (GATHER) {
        m = A.mass[A.order[k]];
        A.mass_sorted[k] = m;
    } else {
        m = A.mass_sorted[k];
    }
    const float4 p = A.pos_sorted[k];
    v[0] = m;
    v[1] = m * static_cast<double>(p.x);
    v[2] = m * static_cast<double>(p.y);
    v[3] = m * static_cast<double>(p.z);
}

__device__ inline DD shfl_up_dd(DD x, int off) {
    return DD{__shfl_up(x.hi, off, 64), __shfl_up(x.lo, off, 64)};
}

// Inclusive scan over the 64 lanes of a wave (fixed order: earlier lanes on
// the left of every addition).
__device__ inline DD wave_scan_dd(DD x, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const DD y = shfl_up_dd(x, off);
        if (lane >= off) x = dd_add(y, x);
    }
    return x;
}

// The exclusive value of a lane from the inclusive scan: the previous lane's.
__device__ inline DD wave_exclusive(DD incl, int lane) {
    const DD e = shfl_up_dd(incl, 1);
    return lane > 0 ? e : DD{0.0, 0.0};
}

constexpr int kMassWaves = kMassThreads / 64;

// Chunk c covers elements [c * C, (c + 1) * C) of [0, n], C = rows *
// kMassThreads (element n is 0: its exclusive prefix is the total), in rows
// of kMassThreads consecutive elements, read with coalesced loads.  A row is
// scanned by wave scans plus the totals of the waves before it.

// Sum of each chunk, per quantity: each thread adds its `rows` elements in
// order (rows of kMassThreads consecutive elements, so every row's loads are
// coalesced), then each wave's lanes are added towards lane 0 (a fixed order)
// and thread q adds the waves' sums in order.  The chunk sums only carry the
// prefix from one chunk into the next; their rounding (double-double, about
// 2^-106 of the total) is far below FP64 resolution of any node's sum.
__device__ void mass_reduce_block(const MassArgs& A, int block) {
    __shared__ DD wsum[4][kMassWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    DD acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    // four rows' loads in flight at a time: fully unrolled, the loads of
    // every row were held in registers (144 VGPRs) and the launch's occupancy,
    // the tree's workgroups included, fell to three waves per SIMD
#pragma unroll 4
    for (int r = 0; r < A.rows; ++r) {
        const int i = (block * A.rows + r) * kMassThreads + threadIdx.x;
        double v[4];
        leaf_vals<true>(A, i, v);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = dd_add(acc[q], DD{v[q], 0.0});
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const DD o{__shfl_down(acc[q].hi, off, 64), __shfl_down(acc[q].lo, off, 64)};
            acc[q] = dd_add(acc[q], o);
        }
    }
    if (lane == 0)
        for (int q = 0; q < 4; ++q) wsum[q][wave] = acc[q];
    __syncthreads();
    if (threadIdx.x < 4) {
        DD t{0.0, 0.0};
        for (int w = 0; w < kMassWaves; ++w) t = dd_add(t, wsum[threadIdx.x][w]);
        A.chunk[threadIdx.x * A.nchunks + block] = t;
    }
}

// The radix tree's nodes and the chunk sums of the mass prefix are
// independent: one launch, the first A.nchunks workgroups reduce chunks and
// the others build nodes (neither half reads what the other writes).
static_assert(kMassThreads == 256, "one workgroup size for both halves");
__global__ __launch_bounds__(256) void tree_mass_kernel(TreeArgs T, MassArgs A) {
    build_priority();
    if (static_cast<int>(blockIdx.x) < A.nchunks) mass_reduce_block(A, blockIdx.x);
    else radix_tree_block(T, blockIdx.x - A.nchunks);
}

// Exclusive scan of the chunk sums (workgroup q scans quantity q).
__global__ __launch_bounds__(kMassThreads) void mass_chunks_kernel(MassArgs A) {
    build_priority();
    __shared__ DD wsum[kMassWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = blockIdx.x;
    DD carry{0.0, 0.0};
    for (int b = 0; b < A.nchunks; b += kMassThreads) {
        const int i = b + threadIdx.x;
        const DD v = i < A.nchunks ? A.chunk[q * A.nchunks + i] : DD{0.0, 0.0};
        const DD incl = wave_scan_dd(v, lane);
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        DD before = carry;
        for (int w = 0; w < wave; ++w) before = dd_add(before, wsum[w]);
        DD row = carry;
        for (int w = 0; w < kMassWaves; ++w) row = dd_add(row, wsum[w]);
        __syncthreads();
        const DD excl = dd_add(before, wave_exclusive(incl, lane));
        if (i < A.nchunks) A.chunk[q * A.nchunks + i] = excl;
        carry = row;
    }
}

// Exclusive prefix P[k] of every element k in [0, n]; coalesced rows.  With
// FUSED (at most kMassFusedChunks chunks) `chunk` holds the chunks' sums and
// the first wavefront adds those of the earlier chunks itself (lane c sums
// chunks c, c + 64, ... in order, then one wave scan), in place of
// mass_chunks_kernel's exclusive scan.
constexpr int kMassFusedChunks = 256;

template <bool FUSED>
__global__ __launch_bounds__(kMassThreads) void mass_prefix_kernel(MassArgs A) {
    build_priority();
    __shared__ DD wsum[4][kMassWaves];
    __shared__ DD wpre[4][kMassWaves];  // the row's sum before each wave, carry included
    __shared__ DD carry_sh[4];          // the sum before the row
    __shared__ DD stage[kMassWaves][64 * kMassItems];  // a wave's prefixes of one quantity
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t stride = static_cast<size_t>(A.n) + 1;
    if constexpr (FUSED) {
        if (wave == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                DD x{0.0, 0.0};
                for (int c = lane; c < static_cast<int>(blockIdx.x); c += 64) x = dd_add(x, A.chunk[q * A.nchunks + c]);
                const DD incl = wave_scan_dd(x, lane);
                if (lane == 63) carry_sh[q] = incl;
            }
        }
    } else {
        if (threadIdx.x < 4) carry_sh[threadIdx.x] = A.chunk[threadIdx.x * A.nchunks + blockIdx.x];
    }
    // Rows of kMassThreads x kItems elements, kItems consecutive ones per
    // lane (a wave reads a contiguous run): the lane adds its elements in
    // order, one wave scan of the lanes' sums per row and quantity, then each
    // element's prefix is the lane's running sum, the wave's prefixes written
    // through LDS as contiguous runs.  Four per lane take a quarter of the
    // wave scans of one per lane: tree build -10% at 131,072 particles, and
    // 1.126 -> 1.095 ms at 4,194,304 since the masses are read in sorted
    // order (while the pass gathered them, one per lane was faster there).
    constexpr int kItems = kMassItems;
#pragma unroll 1
    for (int r = 0; r < A.rows / kItems; ++r) {
        const int i0 = blockIdx.x * A.rows * kMassThreads + (r * kMassThreads + threadIdx.x) * kItems;
        double v[kItems][4];
#pragma unroll
        for (int e = 0; e < kItems; ++e) leaf_vals<false>(A, i0 + e, v[e]);
        DD incl[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            DD sum{v[0][q], 0.0};
#pragma unroll
            for (int e = 1; e < kItems; ++e) sum = dd_add(sum, DD{v[e][q], 0.0});
            incl[q] = wave_scan_dd(sum, lane);
            if (lane == 63) wsum[q][wave] = incl[q];
        }
        __syncthreads();
        // thread q forms quantity q's sums before each wave and the next
        // carry, adding in the same order as every thread would
        if (threadIdx.x < 4) {
            const int q = threadIdx.x;
            DD acc = carry_sh[q];
            for (int w = 0; w < kMassWaves; ++w) {
                wpre[q][w] = acc;
                acc = dd_add(acc, wsum[q][w]);
            }
            carry_sh[q] = acc;
        }
        __syncthreads();
        const int wave0 = blockIdx.x * A.rows * kMassThreads + (r * kMassThreads + wave * 64) * kItems;  // the wave's run
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            DD p = dd_add(wpre[q][wave], wave_exclusive(incl[q], lane));
#pragma unroll
            for (int e = 0; e < kItems; ++e) {
                stage[wave][lane * kItems + e] = p;
                p = dd_add(p, DD{v[e][q], 0.0});
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int k = 0; k < kItems; ++k) {
                const int at = k * 64 + lane;
                if (wave0 + at <= A.n) A.prefix[q * stride + wave0 + at] = stage[wave][at];
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// Traversal record of every kept cell: its FP64 sums (M, Mx, My, Mz) as
// prefix differences, then (COM, G*M) rounded once to FP32.
__global__ __launch_bounds__(256) void cell_com_kernel(const DD* __restrict__ prefix, int n,
                                                       const int* __restrict__ first, const int* __restrict__ last,
                                                       const int* __restrict__ info, int m, double G,
                                                       float4* __restrict__ com) {
    build_priority();
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m || !info_kept(info[i])) return;
    const size_t stride = static_cast<size_t>(n) + 1;
    const int f = first[i], l = last[i] + 1;
    double r[4];
    for (int q = 0; q < 4; ++q) r[q] = dd_diff(prefix[q * stride + l], prefix[q * stride + f]);
    float cx = 0.f, cy = 0.f, cz = 0.f;
    if (r[0] > 0.0) {
        cx = static_cast<float>(r[1] / r[0]);
        cy = static_cast<float>(r[2] / r[0]);
        cz = static_cast<float>(r[3] / r[0]);
    }
    com[i] = make_float4(cx, cy, cz, static_cast<float>(G * r[0]));
}

// The exported description of every radix-tree node that is an octree cell
// (the others are left unwritten; the host keeps those whose word has the
// kept bit).
struct CellExport {
    const unsigned long long* keys;  // sorted
    const int* first;
    const int* last;
    const int* info;
    const float4* com;
    const DD* prefix;  // [4][n + 1]
    const RootParams* root;
    int n;
    nbody_bh_cell* out;
};

__global__ __launch_bounds__(256) void export_cells_kernel(CellExport E) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= E.n - 1 || !info_kept(E.info[i])) return;
    const RootParams R = *E.root;
    const int L = info_level(E.info[i]), f = E.first[i], l = E.last[i];
    const unsigned long long cell = E.keys[f] >> (3 * (kLevels - L));  // the key's first L groups
    nbody_bh_cell c{};
    const double side = ldexp(static_cast<double>(R.size), -L);
    for (int a = 0; a < 3; ++a) {
        unsigned long long idx = 0;
        for (int g = 0; g < L; ++g) idx |= ((cell >> (3 * g + a)) & 1ULL) << g;
        c.lo[a] = static_cast<float>(static_cast<double>(R.lo[a]) + static_cast<double>(idx) * side);
    }
    c.side = static_cast<float>(side);
    const float4 m = E.com[i];
    c.com[0] = m.x;
    c.com[1] = m.y;
    c.com[2] = m.z;
    c.level = static_cast<unsigned>(L);
    c.mass = dd_diff(E.prefix[l + 1], E.prefix[f]);
    c.first = static_cast<unsigned>(f);
    c.count = static_cast<unsigned>(l - f + 1);
    E.out[i] = c;
}

// ---- 6. traversal ---------------------------------------------------------------------

struct TravArgs {
    const float4* node_com;
    const float* thr;  // opening threshold per level
    const int* node_info;
    const unsigned* order;
    const float4* pos_in;   // by particle
    float4* pos_out;        // kick: by particle
    float4* vel;
    float4* acc;
    unsigned long long* walk;  // [0] interactions, [1] list entries, [2] cells opened
    unsigned* overflow;
    int n;
    int kick;
    DirectScale sc;  // the pair arithmetic's scaled units (the direct sum's)
    float dt;
    float dts;
};

using core::kahan_add;

// The walk and the arithmetic are separated.  The walk (uniform over the
// wave) pops a batch of up to eight cells that some lanes opened, reads their
// child blocks, and decides per lane and per child: a child cell is
// approximated (its monopole goes to the list with the lanes that approximate
// it) or opened (pushed with the lanes that open it); a child particle goes to
// the list with every lane that opened the parent but its own.  When the list
// fills, every lane evaluates all entries in a tight loop - four sources at a
// time with grouped rsq, the source's G*m zeroed for lanes outside the entry -
// and adds the result to its compensated sum.  A child's entry is written by
// the lane that holds the child's record in vector registers (lane 8j + s for
// child s of the batch's cell j, one vector load per batch), so an append
// moves only the address and the lane mask.
constexpr int kListCap = 96;

// Margins of the box verdicts: 64 FP32 ulps, above the 16 that the rounding of
// a lane's squared distance and of the box's nearest and farthest squared
// distances can add up to.
constexpr float kBoxLow = 1.0f - 0x1p-18f;
constexpr float kBoxHigh = 1.0f + 0x1p-18f;

// x where the lane's bit of the entry mask is set, else 0 (the bit
// sign-extended to a full-width and-mask); w is the mask word holding the
// lane's bit (lanes 0-31 read the low word, lanes 32-63 the high one).
__device__ __forceinline__ float lane_masked(float x, unsigned w, int lane) {
    return __int_as_float(__float_as_int(x) & __builtin_amdgcn_sbfe(static_cast<int>(w), lane & 31, 1));
}

// v_writelane_b32 through the compiler's own intrinsic is declared above the
// namespace (nb_writelane).


// (lo, hi) = x in lane l only (register writes; the other lanes keep lo and
// hi).  l and x are uniform over the wave.
__device__ __forceinline__ void write_lane64(unsigned& lo, unsigned& hi, unsigned long long x, int l) {
    lo = static_cast<unsigned>(nb_writelane(static_cast<int>(x), l, static_cast<int>(lo)));
    hi = static_cast<unsigned>(nb_writelane(static_cast<int>(x >> 32), l, static_cast<int>(hi)));
}

// A list entry: the source (x, y, z, mass factor) in the direct sum's scaled
// units, as the direct sum stages its tiles (core::stage_factored):
// (x', y', z', h) in one array and k in another, the lanes it applies to in a
// third.  Built when the entry is written, by the lane that holds the source.
template <int LAW>
__device__ __forceinline__ void write_entry(float4* lsrc, float* lk, uint2* lmask, int at, const float4 src,
                                            unsigned long long mask, const DirectScale& sc) {
    float4 p;
    float k;
    core::stage_factored<LAW>(src, sc, p, k);
    lsrc[at] = p;
    lk[at] = k;
    lmask[at] = make_uint2(static_cast<unsigned>(mask), static_cast<unsigned>(mask >> 32));
}

// The list evaluated by every lane, four entries at a time with grouped rsq
// and the last one to three singly (no padding entries: each would cost a
// full evaluation; the direct sum's pair arithmetic, direct_core.hpp), each term zeroed for
// lanes outside its entry; (px, py, pz) is the lane's particle in scaled
// units.  The sums (scaled units) go to the lane's compensated accumulator.
template <int LAW>
__device__ __forceinline__ void flush_list(const float4* lsrc, const float* lk, const uint2* lmask, int cnt, int lane,
                                           float px, float py, float pz, float soft, float& sx, float& sy, float& sz,
                                           float& cx, float& cy, float& cz) {
    const float tiny = core::in_vgpr(__FLT_MIN__);
    float ax = 0.f, ay = 0.f, az = 0.f;
    // the lane's word of every entry's mask
    const unsigned* words = reinterpret_cast<const unsigned*>(lmask) + (lane >> 5);
    const int cnt4 = cnt & ~3;
    for (int e = 0; e < cnt4; e += 4) {
        float4 src[4];
        float dx[4], dy[4], dz[4], q[4];
        const float4 k4 = *reinterpret_cast<const float4*>(lk + e);
        const float kg[4] = {k4.x, k4.y, k4.z, k4.w};
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            src[g] = lsrc[e + g];
            dx[g] = src[g].x - px;
            dy[g] = src[g].y - py;
            dz[g] = src[g].z - pz;
            if constexpr (LAW == NBODY_LAW_R) {
                const float r2 = __builtin_fmaf(dz[g], dz[g], __builtin_fmaf(dy[g], dy[g], dx[g] * dx[g]));
                const float t = __builtin_fmaf(r2, src[g].w, kg[g]);
                const float u = r2 * t;
                q[g] = __builtin_fmaf(u, t, tiny);
            } else {
                q[g] = __builtin_fmaf(dz[g], dz[g], __builtin_fmaf(dy[g], dy[g], __builtin_fmaf(dx[g], dx[g], soft)));
            }
        }
        core::rsq4(q[0], q[1], q[2], q[3]);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            float f;
            if constexpr (LAW == NBODY_LAW_R) {
                f = q[g];
            } else {
                f = (src[g].w * q[g]) * (q[g] * q[g]);
            }
            f = lane_masked(f, words[(e + g) * 2], lane);
            ax = __builtin_fmaf(dx[g], f, ax);
            ay = __builtin_fmaf(dy[g], f, ay);
            az = __builtin_fmaf(dz[g], f, az);
        }
    }
    // the last one to three entries one at a time (the same arithmetic and
    // order; v_rsq_f32 through the compiler's builtin)
    for (int e = cnt4; e < cnt; ++e) {
        const float4 src = lsrc[e];
        const float dx = src.x - px, dy = src.y - py, dz = src.z - pz;
        float q;
        if constexpr (LAW == NBODY_LAW_R) {
            const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            const float t = __builtin_fmaf(r2, src.w, lk[e]);
            const float u = r2 * t;
            q = __builtin_fmaf(u, t, tiny);
        } else {
            q = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, soft)));
        }
        q = __builtin_amdgcn_rsqf(q);
        float f;
        if constexpr (LAW == NBODY_LAW_R) {
            f = q;
        } else {
            f = (src.w * q) * (q * q);
        }
        f = lane_masked(f, words[e * 2], lane);
        ax = __builtin_fmaf(dx, f, ax);
        ay = __builtin_fmaf(dy, f, ay);
        az = __builtin_fmaf(dz, f, az);
    }
    kahan_add(sx, cx, ax);
    kahan_add(sy, cy, ay);
    kahan_add(sz, cz, az);
}

// Workgroups are dealt to the accelerator's XCDs in turn (workgroup b to XCD
// b mod 8), and each XCD has its own L2.  Neighbouring particle groups in
// Morton order walk nearly the same cells, so the groups are renumbered to
// give every XCD runs of 8 consecutive groups of the order (its L2 then holds
// the child blocks those waves share), the runs dealt round the XCDs so that
// dense and sparse regions spread over all of them.  A bijection on [0, nb).
constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_block(int b, int nb) {
    constexpr int C = 2048 / kTravThreads;  // runs of 2048 particles
    const int full = (nb / (kXcds * C)) * (kXcds * C);
    if (b >= full) return b;
    const int x = b % kXcds, k = b / kXcds;
    return ((k / C) * kXcds + x) * C + k % C;
}

// 72 VGPRs (70 for law P): seven waves per SIMD by registers, while the
// workgroup limit keeps 24 one-wave workgroups per CU (six per SIMD)
// resident.  The box's six bounds stay in vector registers: held in scalar
// registers they were spilled and reloaded every batch (2.3% slower).
template <int LAW>
__global__ __launch_bounds__(kTravThreads) void traverse_kernel(const float4* __restrict__ pos_sorted,
                                                                const ChildBlock* __restrict__ blocks,
                                                                const int* __restrict__ node_first,
                                                                const int* __restrict__ node_last, TravArgs A) {
    // The walk runs at raised wave priority and drops it while the list is
    // evaluated: a SIMD then issues the walk's dependent loads and decisions
    // ahead of other wavefronts' list arithmetic, which fills the walk's
    // memory waits (traversal -1% at N = 4M, -3% at 1M; priority 3 gained
    // less, the opposite assignment nothing).
    __builtin_amdgcn_s_setprio(1);
    __shared__ int stk_node[kTravWaves][kStackCap];
    __shared__ uint2 stk_mask[kTravWaves][kStackCap];
    __shared__ float4 lsrc[kTravWaves][kListCap];
    __shared__ __align__(16) float lk[kTravWaves][kListCap];
    __shared__ uint2 lmask[kTravWaves][kListCap];
    __shared__ unsigned long long red[3][kTravWaves];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: list and stack addresses in SGPRs
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int t = blk * kTravThreads + tid;
    const int wave_base = blk * kTravThreads + wave * 64;
    const bool valid = t < A.n;
    const float4 me = pos_sorted[valid ? t : A.n - 1];
    // the particle in the pair arithmetic's scaled units (the decisions use me)
    const float mx = me.x * A.sc.inv_len, my = me.y * A.sc.inv_len, mz = me.z * A.sc.inv_len;
    const float soft = core::in_vgpr(A.sc.soft);
    float sx = 0.f, sy = 0.f, sz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
    // wave-uniform counters and list length
    unsigned long long count = 0, entries = 0, opened = 0;
    int cnt = 0;
    int sp = 0;
    bool overflowed = false;
    // interactions of the entries this lane wrote, folded into a
    // 64-bit per-lane total at every flush
    unsigned lane_count = 0;
    unsigned long long lane_total = 0;
    // during a batch, child cell s of cell j's mask of the lanes at or beyond
    // its opening threshold, in lane 8j + s; lanes of particle children and
    // lanes beyond the batch's children are not used
    unsigned ge_lo = 0u, ge_hi = 0u;

    // Append a source (x, y, z, G m) held in scalar registers (written by lane 0).
    auto append = [&](const float4 src, unsigned long long mask) {
        if (lane == 0)
            write_entry<LAW>(lsrc[wave], lk[wave], lmask[wave], cnt,
                             make_float4(src.x, src.y, src.z, core::mass_factor<LAW>(src.w, A.sc)), mask, A.sc);
        count += static_cast<unsigned long long>(__popcll(mask));
        ++cnt;
    };
    auto flush = [&]() {
        entries += static_cast<unsigned long long>(cnt);
        const int total = __builtin_amdgcn_readfirstlane(cnt);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_s_setprio(0);  // the list at normal priority, the walk raised (kernel start)
        flush_list<LAW>(lsrc[wave], lk[wave], lmask[wave], total, lane, mx, my, mz, soft, sx, sy, sz, cx, cy, cz);
        __builtin_amdgcn_s_setprio(1);
        __builtin_amdgcn_wave_barrier();
        cnt = 0;
        lane_total += lane_count;
        lane_count = 0;
    };
    // the lanes of a particle's own position never interact with it
    auto without_self = [&](unsigned long long m, int q) {
        const int l = q - wave_base;
        return (l >= 0 && l < 64) ? (m & ~(1ULL << l)) : m;
    };

    // The wave's bounding box (lanes past n hold the last particle, inside it).
    // A child cell whose nearest point of the box is beyond its threshold is
    // approximated by every lane, one whose farthest point is within it is
    // opened by every lane: those verdicts are taken for all 64 children of a
    // batch at once, one child per lane, and only the others are tested lane
    // by lane.  The margins cover the rounding of both computations (each
    // within 16 FP32 ulps of the exact squared distance), so every verdict is
    // the decision each lane would make.  Not used when a position is not
    // finite.
    float bx0 = me.x, by0 = me.y, bz0 = me.z, bx1 = me.x, by1 = me.y, bz1 = me.z;
    for (int off = 32; off > 0; off >>= 1) {
        bx0 = fminf(bx0, __shfl_xor(bx0, off));
        by0 = fminf(by0, __shfl_xor(by0, off));
        bz0 = fminf(bz0, __shfl_xor(bz0, off));
        bx1 = fmaxf(bx1, __shfl_xor(bx1, off));
        by1 = fmaxf(by1, __shfl_xor(by1, off));
        bz1 = fmaxf(bz1, __shfl_xor(bz1, off));
    }
    const bool box_ok =
        __builtin_amdgcn_ballot_w64(!(__builtin_isfinite(me.x) && __builtin_isfinite(me.y) &&
                                      __builtin_isfinite(me.z))) == 0ull;

    const unsigned long long all = __builtin_amdgcn_ballot_w64(valid);
    if (all) {
        // the root: approximated by the lanes beyond its threshold (the
        // reference's side / |p - COM| < theta as r2 >= thr, see
        // open_threshold), opened by the others
        const int root_level = info_level(A.node_info[0]);
        const float4 c = A.node_com[0];
        const float dx = me.x - c.x, dy = me.y - c.y, dz = me.z - c.z;
        const float r2 = (dx * dx + dy * dy) + dz * dz;
        const unsigned long long am = __builtin_amdgcn_ballot_w64(!(r2 < A.thr[root_level])) & all;
        if (am) append(c, am);
        const unsigned long long om = all & ~am;
        if (om) {
            if (lane == 0) {
                stk_node[wave][0] = root_level >= kBucketLevel ? kBucketFlag : 0;
                stk_mask[wave][0] = make_uint2(static_cast<unsigned>(om), static_cast<unsigned>(om >> 32));
            }
            sp = 1;
        }
    }
    // Lane 8j + s holds child s of the batch's j-th cell (j = 0 the top of
    // the stack) in the write phase.
    const int grp = lane >> 3;
    while (sp > 0) {
        __builtin_amdgcn_wave_barrier();
        // The batch: up to eight cells from the top of the stack, as many as
        // the stack has room to replace by their children (each cell nets at
        // most seven more entries) while keeping the depth-first reserve, at
        // least one, and no bucket among them.  Group j reads entry sp-1-j
        // (groups beyond the batch repeat its last entry).
        const int kmax = min(min(8, sp), max((kStackCap - kDfsReserve - sp) / 7, 1));
        const int e = sp - 1 - min(grp, kmax - 1);
        const int vcell = stk_node[wave][e];
        const uint2 vm = stk_mask[wave][e];
        const unsigned long long bk = __builtin_amdgcn_ballot_w64((vcell & kBucketFlag) != 0);
        int k = bk ? min(kmax, static_cast<int>(__builtin_ctzll(bk)) >> 3) : kmax;
        if (k == 0) {
            // identical keys: every particle of the range, pair by pair
            --sp;
            ++opened;
            const int node = __builtin_amdgcn_readfirstlane(vcell) & ~kBucketFlag;
            const unsigned lo32 = __builtin_amdgcn_readfirstlane(vm.x);
            const unsigned hi32 = __builtin_amdgcn_readfirstlane(vm.y);
            const unsigned long long mask = (static_cast<unsigned long long>(hi32) << 32) | lo32;
            const int f = node_first[node], l = node_last[node];
            for (int q = f; q <= l; ++q) {
                if (cnt > kListCap - 4) flush();
                const unsigned long long m = without_self(mask, q);
                if (m) append(pos_sorted[q], m);
            }
            continue;
        }
        // Up to eight list entries per cell: a smaller batch when at least
        // four cells fit, else evaluate the list first.
        if (cnt > kListCap - 8 * k) {
            const int fit = (kListCap - cnt) >> 3;
            if (fit >= 4) k = fit;
            else flush();
        }
        opened += static_cast<unsigned long long>(k);
        const int base = sp - k;
        // Lane 8j + s reads child s of cell j (lanes beyond the batch read the
        // top cell's block and take no part), issued before the decisions.
        const int top = __builtin_amdgcn_readfirstlane(vcell);  // all lanes active: group 0
        const bool in_batch = grp < k;
        const ChildBlock* VB = blocks + (in_batch ? vcell : top);
        const float4 vsrc = VB->src[lane & 7];
        const float vthr = VB->thr[lane & 7];
        const int vcode = VB->code[lane & 7];
        // Box verdicts for the child this lane holds (see the box above).
        bool open_all = false, accept_all = false;
        if (box_ok) {
            const float ex = fmaxf(fmaxf(bx0 - vsrc.x, vsrc.x - bx1), 0.f);
            const float ey = fmaxf(fmaxf(by0 - vsrc.y, vsrc.y - by1), 0.f);
            const float ez = fmaxf(fmaxf(bz0 - vsrc.z, vsrc.z - bz1), 0.f);
            const float fx = fmaxf(vsrc.x - bx0, bx1 - vsrc.x);
            const float fy = fmaxf(vsrc.y - by0, by1 - vsrc.y);
            const float fz = fmaxf(vsrc.z - bz0, bz1 - vsrc.z);
            const float near2 = (ex * ex + ey * ey) + ez * ez;
            const float far2 = (fx * fx + fy * fy) + fz * fz;
            accept_all = near2 * kBoxLow >= vthr;
            open_all = far2 * kBoxHigh < vthr;
        }
        // child cells of the batch that need the lane-by-lane test
        const unsigned long long tested =
            __builtin_amdgcn_ballot_w64(vcode >= 0 && !open_all && !accept_all) &
            (k == 8 ? ~0ull : ((1ull << (8 * k)) - 1ull));
        // Decisions, cell by cell, for the tested children only (a cell none
        // of whose children is tested is skipped).  The block's centres of
        // mass and thresholds in scalar registers at once (the empty asm
        // keeps the compiler from sinking each slot's loads into that slot's
        // branch, which would wait on memory once per cache line).  A child
        // particle needs no test: every lane of the mask but its own takes
        // it.  A child cell is approximated by the lanes with !(r2 < thr) (the
        // reference's side / |p - COM| < theta, see open_threshold) and opened
        // by the others; that lane mask goes to lane 8j + s, unless it is
        // every lane (the lanes' masks start as all lanes).
        ge_lo = open_all ? 0u : ~0u;
        ge_hi = open_all ? 0u : ~0u;
        for (int j = 0; j < k; ++j) {
            const unsigned tj = static_cast<unsigned>(tested >> (8 * j)) & 0xffu;
            if (tj == 0u) continue;  // every child cell of cell j has its box verdict
            const int code = __builtin_amdgcn_readlane(vcell, j * 8);
            const ChildBlock* B = blocks + code;
            float4 srcs[8];
            float thrs[8];
#pragma unroll
            for (int slot = 0; slot < 8; ++slot) {
                srcs[slot] = B->src[slot];
                thrs[slot] = B->thr[slot];
            }
#pragma unroll
            for (int slot = 0; slot < 8; ++slot)
                asm volatile("" ::"s"(srcs[slot].x), "s"(srcs[slot].y), "s"(srcs[slot].z), "s"(thrs[slot]));
            const int lbase = j * 8;
            // the tested children are cells of cell j (its block holds them in
            // its first slots): a scalar bit test per slot
#pragma unroll
            for (int slot = 0; slot < 8; ++slot) {
                if (!((tj >> slot) & 1u)) continue;
                const float dx = me.x - srcs[slot].x, dy = me.y - srcs[slot].y, dz = me.z - srcs[slot].z;
                const float r2 = (dx * dx + dy * dy) + dz * dz;
                const unsigned long long ge = __builtin_amdgcn_ballot_w64(!(r2 < thrs[slot]));
                if (ge != ~0ull) write_lane64(ge_lo, ge_hi, ge, lbase + slot);
            }
        }
        // Lane 8j + s forms child s's two masks: approximating lanes am (for a
        // particle every lane of the mask but the particle's own) and opening
        // lanes om (none for a particle).
        const unsigned long long mask = (static_cast<unsigned long long>(vm.y) << 32) | vm.x;
        const unsigned long long batch = k == 8 ? ~0ull : ((1ull << (8 * k)) - 1ull);
        const bool child = in_batch && vcode != kEmpty;
        const unsigned long long children = __builtin_amdgcn_ballot_w64(vcode != kEmpty) & batch;
        const bool leaf = vcode < 0;
        const unsigned own = static_cast<unsigned>(dec_leaf(vcode) - wave_base);
        const unsigned long long self = (leaf && own < 64u) ? (1ull << own) : 0ull;
        const unsigned long long ge = leaf ? ~0ull : ((static_cast<unsigned long long>(ge_hi) << 32) | ge_lo);
        const unsigned long long am = mask & ge & ~self;
        const unsigned long long om = mask & ~ge;
        // List entries, in lane order, written by the children's lanes.
        const bool wl = child && am != 0ull;
        const unsigned long long has_l = __builtin_amdgcn_ballot_w64(am != 0ull) & children;
        if (wl) {
            const int at = static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                static_cast<unsigned>(has_l >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(has_l), cnt)));
            write_entry<LAW>(lsrc[wave], lk[wave], lmask[wave], at, vsrc, am, A.sc);
            lane_count += __popcll(am);
        }
        cnt += __popcll(has_l);
        // Opened cells onto the stack in place of the batch, in reverse lane
        // order: the top cell's first opened child ends on top.
        const bool ws = child && om != 0ull;  // a cell: a particle's ge is all lanes
        const unsigned long long has_s = __builtin_amdgcn_ballot_w64(om != 0ull) & children;
        const int pushed = __popcll(has_s);
        const int last = base + pushed - 1;
        if (last < kStackCap) {
            if (ws) {
                const int at = last - static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                                          static_cast<unsigned>(has_s >> 32),
                                          __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(has_s), 0u)));
                stk_node[wave][at] = vcode;
                stk_mask[wave][at] = make_uint2(static_cast<unsigned>(om), static_cast<unsigned>(om >> 32));
            }
            sp = base + pushed;
        } else {
            // not reached (the depth-first reserve): the entries beyond the
            // capacity would be lost and reported
            if (ws) {
                const int at = last - static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                                          static_cast<unsigned>(has_s >> 32),
                                          __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(has_s), 0u)));
                if (at < kStackCap) {
                    stk_node[wave][at] = vcode;
                    stk_mask[wave][at] = make_uint2(static_cast<unsigned>(om), static_cast<unsigned>(om >> 32));
                }
            }
            sp = kStackCap;
            overflowed = true;
        }
    }
    if (overflowed && lane == 0) atomicOr(A.overflow, 1u);
    if (cnt) flush();
    for (int off = 32; off > 0; off >>= 1) lane_total += __shfl_xor(lane_total, off);
    count += lane_total;

    if (valid) {
        const unsigned p = A.order[t];
        // back from scaled units (a power of two: exact)
        const float ax = sx * A.sc.out, ay = sy * A.sc.out, az = sz * A.sc.out;
        A.acc[p] = make_float4(ax, ay, az, 0.f);
        if (A.kick) {
            float4 v = A.vel[p];
            float4 x = A.pos_in[p];
            v.x = __builtin_fmaf(ax, A.dt, v.x);
            v.y = __builtin_fmaf(ay, A.dt, v.y);
            v.z = __builtin_fmaf(az, A.dt, v.z);
            x.x = __builtin_fmaf(v.x, A.dts, x.x);
            x.y = __builtin_fmaf(v.y, A.dts, x.y);
            x.z = __builtin_fmaf(v.z, A.dts, x.z);
            A.vel[p] = v;
            A.pos_out[p] = x;
        }
    }
    if (lane == 0) {
        red[0][wave] = count;
        red[1][wave] = entries;
        red[2][wave] = opened;
    }
    __syncthreads();
    if (tid < 3) {
        unsigned long long tot = 0;
        for (int w = 0; w < kTravWaves; ++w) tot += red[tid][w];
        atomicAdd(A.walk + tid, tot);
    }
}

}  // namespace

// ---- host side ----------------------------------------------------------------------

struct BHState {
    size_t cap = 0;  // particles the buffers hold
    RootParams* root = nullptr;
    float* bounds_partial = nullptr;
    unsigned long long* keys[2] = {nullptr, nullptr};
    unsigned* idx[2] = {nullptr, nullptr};
    unsigned* hist = nullptr;
    size_t hist_elems = 0;
    float4* pos_sorted = nullptr;
    int2* child = nullptr;  // radix-tree nodes: encoded (left, right) children
    int* first = nullptr;
    int* last = nullptr;
    int* info = nullptr;    // level | kKeptBit
    ChildBlock* blocks = nullptr;
    unsigned* digit_totals = nullptr;  // sort: keys per digit of the current pass
    size_t chunk_elems = 0;
    DD* prefix = nullptr;            // [4][n + 1] double-double prefix sums
    DD* mass_chunk = nullptr;        // [4][nchunks]
    float4* com = nullptr;
    float* thr = nullptr;                    // [kThrLevels] opening thresholds
    unsigned* counters = nullptr;            // [0] outside, [1] overflow
    unsigned long long* walk = nullptr;      // [3] traversal counters
    unsigned long long* host_ctr = nullptr;  // [5] pinned mirror of walk and counters
    int sorted_buf = 0;                      // which keys/idx buffer holds the sorted result
    RootParams fixed{};
    bool fixed_root = false;
    bool built = false;  // a tree has been built since the buffers were sized
};

namespace {

constexpr int kBoundsBlocks = 512;

template <class T>
void bh_free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void bh_release(BHState* b) {
    bh_free(b->bounds_partial);
    bh_free(b->keys[0]);
    bh_free(b->keys[1]);
    bh_free(b->idx[0]);
    bh_free(b->idx[1]);
    bh_free(b->hist);
    bh_free(b->pos_sorted);
    bh_free(b->child);
    bh_free(b->first);
    bh_free(b->last);
    bh_free(b->info);
    bh_free(b->blocks);
    bh_free(b->digit_totals);
    bh_free(b->prefix);
    bh_free(b->mass_chunk);
    bh_free(b->com);
    bh_free(b->thr);
    b->cap = 0;
    b->hist_elems = 0;
    b->built = false;
}

template <class T>
nbody_status bh_alloc(T*& p, size_t count) {
    bh_free(p);
    NB_HIP(hipMalloc(reinterpret_cast<void**>(&p), (count ? count : 1) * sizeof(T)));
    return NBODY_OK;
}

inline int blocks_for(size_t n, int per) { return static_cast<int>((n + per - 1) / per); }

// Rows per mass chunk: the fewest (small chunks give small systems more
// workgroups in the chunk sums and the prefix pass) that keep the system
// within kMassFusedChunks chunks, kMassPer beyond.
inline int mass_rows(int n) {
    for (int rows = kMassMinRows; rows < kMassPer; rows *= 2)
        if (blocks_for(static_cast<size_t>(n) + 1, rows * kMassThreads) <= kMassFusedChunks) return rows;
    return kMassPer;
}

float theta_below(double theta) {
    float t = static_cast<float>(theta);
    if (static_cast<double>(t) >= theta) t = std::nextafter(t, -INFINITY);
    return t;
}

}  // namespace

// The fixed root's exact quantisation parameters; fails when the corner is not
// a multiple of the level-21 cell.
nbody_status bh_fixed_root(const nbody_config* c, void* out_params) {
    RootParams R{};
    int e;
    const float m = std::frexp(c->root_size, &e);
    long long mi = static_cast<long long>(std::ldexp(m, 24));
    int p = e - 24;
    while ((mi & 1) == 0) {
        mi >>= 1;
        ++p;
    }
    NB_CHECK(mi < (1LL << 24) && 21 - p > -100 && 21 - p < 100, NBODY_ERR_INVALID_ARG, "unusable root size");
    R.size = c->root_size;
    R.odd = static_cast<int>(mi);
    R.scale2 = std::ldexp(1.0f, 21 - p);
    for (int a = 0; a < 3; ++a) {
        R.lo[a] = c->root_lo[a];
        const double L = std::floor(static_cast<double>(c->root_lo[a]) * static_cast<double>(R.scale2));
        NB_CHECK(static_cast<double>(c->root_lo[a]) * static_cast<double>(R.scale2) == L && std::fabs(L) < 9e15,
                 NBODY_ERR_INVALID_ARG, "fixed root corner must lie on the level-21 grid of the root");
        const long long Li = static_cast<long long>(L);
        NB_CHECK(Li % R.odd == 0, NBODY_ERR_INVALID_ARG,
                 "fixed root corner must be a multiple of root_size / 2^21");
        R.base[a] = floor_div_ll(Li, R.odd);
    }
    if (out_params) *static_cast<RootParams*>(out_params) = R;
    return NBODY_OK;
}

nbody_status bh_create(nbody_sim* s) {
    BHState* b = new (std::nothrow) BHState();
    NB_CHECK(b, NBODY_ERR_OUT_OF_MEMORY, "host allocation failed");
    if (s->cfg.root_mode == NBODY_ROOT_FIXED) {
        nbody_status st = bh_fixed_root(&s->cfg, &b->fixed);
        if (st != NBODY_OK) {
            delete b;
            return st;
        }
        b->fixed_root = true;
    }
    hipError_t e1 = hipMalloc(reinterpret_cast<void**>(&b->root), sizeof(RootParams));
    hipError_t e2 = hipMalloc(reinterpret_cast<void**>(&b->counters), 2 * sizeof(unsigned));
    hipError_t e3 = hipMalloc(reinterpret_cast<void**>(&b->walk), 3 * sizeof(unsigned long long));
    hipError_t e4 = hipHostMalloc(reinterpret_cast<void**>(&b->host_ctr), 5 * sizeof(unsigned long long),
                                  hipHostMallocDefault);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
        bh_destroy(b);
        set_error("hipMalloc failed for Barnes-Hut control blocks");
        return NBODY_ERR_OUT_OF_MEMORY;
    }
    // On the handle's stream: synchronous copies would wait for whatever else
    // occupies the device.
    if (b->fixed_root) {
        hipError_t e = hipMemcpyAsync(b->root, &b->fixed, sizeof(RootParams), hipMemcpyHostToDevice, s->stream);
        if (e != hipSuccess) {
            bh_destroy(b);
            set_error("hipMemcpyAsync failed: %s", hipGetErrorString(e));
            return NBODY_ERR_DEVICE;
        }
    }
    for (int k = 0; k < 5; ++k) b->host_ctr[k] = 0;
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s->stream, b->counters, b->walk, b->host_ctr);
    s->bh = b;
    return NBODY_OK;
}

void bh_destroy(BHState* b) {
    if (!b) return;
    bh_release(b);
    bh_free(b->root);
    bh_free(b->counters);
    bh_free(b->walk);
    if (b->host_ctr) (void)hipHostFree(b->host_ctr);
    b->host_ctr = nullptr;
    delete b;
}

nbody_status bh_resize(BHState* b, size_t n, nbody_sim* s) {
    if (n <= b->cap && b->cap) return NBODY_OK;
    bh_release(b);
    const size_t m = n > 1 ? n - 1 : 1;
    NB_TRY(bh_alloc(b->bounds_partial, 6 * kBoundsBlocks));
    NB_TRY(bh_alloc(b->keys[0], n));
    NB_TRY(bh_alloc(b->keys[1], n));
    NB_TRY(bh_alloc(b->idx[0], n));
    NB_TRY(bh_alloc(b->idx[1], n));
    const size_t nb = sort_max_tiles(n);
    NB_TRY(bh_alloc(b->hist, kRadixBins * nb));
    b->hist_elems = kRadixBins * nb;
    NB_TRY(bh_alloc(b->pos_sorted, n));
    NB_TRY(bh_alloc(b->child, m));
    NB_TRY(bh_alloc(b->first, m));
    NB_TRY(bh_alloc(b->last, m));
    NB_TRY(bh_alloc(b->info, m));
    NB_TRY(bh_alloc(b->blocks, m));
    NB_TRY(bh_alloc(b->digit_totals, kRadixBins));
    NB_TRY(bh_alloc(b->prefix, 4 * (n + 1)));
    NB_TRY(bh_alloc(b->mass_chunk, 4 * static_cast<size_t>(std::max(kMassFusedChunks, blocks_for(n + 1, kMassChunk)))));
    NB_TRY(bh_alloc(b->com, m));
    NB_TRY(bh_alloc(b->thr, kThrLevels));
    b->cap = n;
    if (b->fixed_root) {
        // The fixed root's thresholds do not change: computed once per allocation.
        hipLaunchKernelGGL(open_threshold_kernel, dim3(1), dim3(64), 0, s->stream, b->root,
                           theta_below(s->cfg.theta), b->thr);
        NB_HIP(hipGetLastError());
    }
    return NBODY_OK;
}

nbody_status bh_set_theta(nbody_sim* s) {
    BHState* b = s->bh;
    // Before the first allocation there are no thresholds yet: bh_resize
    // computes them from the handle's theta.
    if (!b || !b->fixed_root || !b->thr) return NBODY_OK;
    hipLaunchKernelGGL(open_threshold_kernel, dim3(1), dim3(64), 0, s->stream, b->root, theta_below(s->cfg.theta),
                       b->thr);
    NB_HIP(hipGetLastError());
    return NBODY_OK;
}

// The kRadixPasses LSD passes over the keys in tiles of 256 x ITEMS (the last pass
// also gathers the sorted positions); the sorted buffer is b->sorted_buf.
template <int ITEMS>
void radix_sort_passes(BHState* b, int n, const float4* pos, hipStream_t st) {
    const int nb = blocks_for(n, kSortThreads * ITEMS);
    int cur = 0;
    for (int pass = 0; pass < kRadixPasses; ++pass) {
        const int shift = kRadixBits * pass;
        // Above kFusedScanTiles tiles the tile offsets within each digit and
        // the digits' totals come from a separate scan of the counts.
        const bool fused = nb <= kFusedScanTiles;
        hipLaunchKernelGGL(radix_hist_kernel<ITEMS>, dim3(nb), dim3(kSortThreads), 0, st, b->keys[cur], n, shift, b->hist, nb);
        if (!fused)
            hipLaunchKernelGGL(radix_rowscan_kernel, dim3(kRadixBins), dim3(kScanThreads2), 0, st, b->hist, nb,
                               b->digit_totals);
        // the last pass also writes the sorted positions
        auto scatter = pass < kRadixPasses - 1 ? (fused ? radix_scatter_kernel<true, false, ITEMS> : radix_scatter_kernel<false, false, ITEMS>)
                                : (fused ? radix_scatter_kernel<true, true, ITEMS> : radix_scatter_kernel<false, true, ITEMS>);
        hipLaunchKernelGGL(scatter, dim3(nb), dim3(kSortThreads), 0, st, b->keys[cur], b->idx[cur], b->keys[cur ^ 1],
                           b->idx[cur ^ 1], n, shift, b->hist, b->digit_totals, nb, pos, b->pos_sorted);
        cur ^= 1;
    }
    b->sorted_buf = cur;
}

// Keys, sort, tree, masses: everything before the traversal.
static nbody_status bh_build(nbody_sim* s, BHState* b) {
    const int n = static_cast<int>(s->n);
    hipStream_t st = s->stream;
    const float4* pos = s->pos[s->cur];
    if (!b->fixed_root) {
        const int nbb = std::min(kBoundsBlocks, std::max(1, blocks_for(n, 2048)));
        hipLaunchKernelGGL(bounds_kernel, dim3(nbb), dim3(256), 0, st, pos, n, b->bounds_partial);
        hipLaunchKernelGGL(dynamic_root_kernel, dim3(1), dim3(256), 0, st, b->bounds_partial, nbb, b->root,
                           theta_below(s->cfg.theta), b->thr);
    }
    hipLaunchKernelGGL(keys_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, pos, n, b->root, b->keys[0],
                       b->idx[0], b->counters);
    switch (sort_items(static_cast<size_t>(n))) {
    case 2: radix_sort_passes<2>(b, n, pos, st); break;
    case 4: radix_sort_passes<4>(b, n, pos, st); break;
    default: radix_sort_passes<8>(b, n, pos, st); break;
    }
    const unsigned long long* keys = b->keys[b->sorted_buf];
    const unsigned* order = b->idx[b->sorted_buf];
    const int m = n - 1;
    TreeArgs T{keys, n, b->child, b->first, b->last, b->info};
    const int tree_blocks = blocks_for(m, 256);
    const int mrows = mass_rows(n);
    const int mchunks = blocks_for(n + 1, mrows * kMassThreads);
    // the sorted masses go to the sort's other key buffer, idle until the next step's keys
    double* mass_sorted = reinterpret_cast<double*>(b->keys[b->sorted_buf ^ 1]);
    MassArgs M{b->pos_sorted, order, s->mass_dev, mass_sorted, n, b->prefix, b->mass_chunk, mchunks, mrows};
    hipLaunchKernelGGL(tree_mass_kernel, dim3(mchunks + tree_blocks), dim3(256), 0, st, T, M);
    if (mchunks <= kMassFusedChunks) {
        hipLaunchKernelGGL(mass_prefix_kernel<true>, dim3(mchunks), dim3(kMassThreads), 0, st, M);
    } else {
        hipLaunchKernelGGL(mass_chunks_kernel, dim3(4), dim3(kMassThreads), 0, st, M);
        hipLaunchKernelGGL(mass_prefix_kernel<false>, dim3(mchunks), dim3(kMassThreads), 0, st, M);
    }
    hipLaunchKernelGGL(cell_com_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, st, b->prefix, n, b->first, b->last,
                       b->info, m, s->cfg.G, b->com);
    BlockArgs L{b->child, b->info, b->com, b->thr, b->pos_sorted, m, b->blocks, s->dscale};
    b->built = true;
    if (s->cfg.law == NBODY_LAW_R)
        hipLaunchKernelGGL(block_kernel<NBODY_LAW_R>, dim3(blocks_for(m, kBlockNodes)), dim3(256), 0, st, L);
    else
        hipLaunchKernelGGL(block_kernel<NBODY_LAW_P>, dim3(blocks_for(m, kBlockNodes)), dim3(256), 0, st, L);
    NB_HIP(hipGetLastError());
    return NBODY_OK;
}

nbody_status bh_step(nbody_sim* s, float dt, bool kick, hipEvent_t ev_b, hipEvent_t ev_e) {
    BHState* b = s->bh;
    NB_CHECK(b, NBODY_ERR_NOT_INITIALISED, "Barnes-Hut state missing");
    const int n = static_cast<int>(s->n);
    hipStream_t st = s->stream;
    if (n < 2) {
        // A single particle feels no force; the direct pass gives exactly that.
        DirectLaunch L{};
        L.src = s->pos[s->cur];
        L.pos_out = s->pos[s->cur ^ 1];
        L.vel = s->vel;
        L.acc = s->acc;
        L.part = s->part;
        L.n = n;
        L.n_pad = static_cast<int>(s->n_pad);
        L.law = s->cfg.law;
        L.sc = s->dscale;
        L.dt = dt;
        L.dts = static_cast<float>(static_cast<double>(dt) / s->cfg.scale);
        NB_HIP(direct_step(L, s->shape, kick, st, ev_b, ev_e));
        hipLaunchKernelGGL(publish_counters_kernel, dim3(1), dim3(64), 0, st, b->counters, b->walk, b->host_ctr);
        NB_HIP(hipGetLastError());
        return NBODY_OK;
    }
    hipEvent_t tb = nullptr, te = nullptr;
    if (s->timing) {
        NB_HIP(hipEventCreate(&tb));
        s->tree_ms_pending.push_back(tb);
        NB_HIP(hipEventCreate(&te));
        s->tree_ms_pending.push_back(te);
        NB_HIP(hipEventRecord(tb, st));
    }
    NB_TRY(bh_build(s, b));
    if (te) NB_HIP(hipEventRecord(te, st));

    TravArgs A{};
    A.node_com = b->com;
    A.thr = b->thr;
    A.node_info = b->info;
    A.order = b->idx[b->sorted_buf];
    A.pos_in = s->pos[s->cur];
    A.pos_out = s->pos[s->cur ^ 1];
    A.vel = s->vel;
    A.acc = s->acc;
    A.walk = b->walk;
    A.overflow = b->counters + 1;
    A.n = n;
    A.kick = kick ? 1 : 0;
    A.sc = s->dscale;
    A.dt = dt;
    A.dts = static_cast<float>(static_cast<double>(dt) / s->cfg.scale);
    if (ev_b) NB_HIP(hipEventRecord(ev_b, st));
    const dim3 tgrid(blocks_for(n, kTravThreads));
    if (s->cfg.law == NBODY_LAW_R)
        hipLaunchKernelGGL(traverse_kernel<NBODY_LAW_R>, tgrid, dim3(kTravThreads), 0, st, b->pos_sorted, b->blocks,
                           b->first, b->last, A);
    else
        hipLaunchKernelGGL(traverse_kernel<NBODY_LAW_P>, tgrid, dim3(kTravThreads), 0, st, b->pos_sorted, b->blocks,
                           b->first, b->last, A);
    NB_HIP(hipGetLastError());
    if (ev_e) NB_HIP(hipEventRecord(ev_e, st));
    hipLaunchKernelGGL(publish_counters_kernel, dim3(1), dim3(64), 0, st, b->counters, b->walk, b->host_ctr);
    NB_HIP(hipGetLastError());
    return NBODY_OK;
}

nbody_status bh_reset_counters(BHState* b, hipStream_t st) {
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, st, b->counters, b->walk, b->host_ctr);
    NB_HIP(hipGetLastError());
    return NBODY_OK;
}

nbody_status bh_read_counters(BHState* b, hipStream_t st, BHCounters* out) {
    NB_HIP(hipStreamSynchronize(st));  // the last force pass has published its counters
    const volatile unsigned long long* h = b->host_ctr;
    out->interactions = h[0];
    out->entries = h[1];
    out->opened = h[2];
    out->outside = static_cast<unsigned>(h[3]);
    out->overflow = static_cast<unsigned>(h[4]);
    return NBODY_OK;
}

nbody_status bh_export_cells(nbody_sim* s, nbody_bh_cell* out, size_t capacity, size_t* count) {
    BHState* b = s->bh;
    const size_t n = s->n;
    NB_CHECK(b && b->built && b->cap >= n, NBODY_ERR_NOT_INITIALISED, "no Barnes-Hut tree built yet");
    *count = 0;
    if (n < 2) return NBODY_OK;
    const size_t m = n - 1;
    nbody_bh_cell* d = nullptr;
    NB_HIP(hipMalloc(reinterpret_cast<void**>(&d), m * sizeof(nbody_bh_cell)));
    CellExport E{b->keys[b->sorted_buf], b->first, b->last, b->info, b->com, b->prefix, b->root,
                 static_cast<int>(n), d};
    hipLaunchKernelGGL(export_cells_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, s->stream, E);
    std::vector<nbody_bh_cell> cells(m);
    std::vector<int> info(m);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(cells.data(), d, m * sizeof(nbody_bh_cell), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(info.data(), b->info, m * sizeof(int), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    NB_HIP(e);
    std::vector<nbody_bh_cell> list;
    for (size_t i = 0; i < m; ++i)
        if (info_kept(info[i])) list.push_back(cells[i]);
    // depth-first preorder of the octree: by first particle, enclosing cells first
    std::sort(list.begin(), list.end(), [](const nbody_bh_cell& x, const nbody_bh_cell& y) {
        return x.first != y.first ? x.first < y.first : x.count > y.count;
    });
    *count = list.size();
    for (size_t i = 0; i < list.size() && i < capacity; ++i) out[i] = list[i];
    return NBODY_OK;
}

nbody_status bh_export_order(nbody_sim* s, uint64_t* keys, uint32_t* order, size_t n) {
    BHState* b = s->bh;
    NB_CHECK(b && b->cap >= n && n == s->n, NBODY_ERR_NOT_INITIALISED, "no Barnes-Hut tree of %zu particles", n);
    if (n == 0) return NBODY_OK;
    if (keys)
        NB_HIP(hipMemcpyAsync(keys, b->keys[b->sorted_buf], n * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
    if (order)
        NB_HIP(hipMemcpyAsync(order, b->idx[b->sorted_buf], n * sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream));
    NB_HIP(hipStreamSynchronize(s->stream));
    return NBODY_OK;
}

}  // namespace nbody
