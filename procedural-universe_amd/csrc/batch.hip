// Batched direct sum: many independent systems advanced together.
//
// The reference steps one particle set per INBodySim (src/Sim/INBodySim.hpp:19-25);
// a GPU that steps a 1,024-particle galaxy alone keeps four of its 256 CUs busy.
// Here every system lives in one device arena (the same padded SoA records as a
// single simulation) and one persistent kernel per step drains a queue of
// (system, target block) work items, largest systems first, so small and large
// systems fill the machine together.  Each work item runs the single-system
// core (direct_core.hpp) over its own system's sources, so a system's forces are
// the same bits whether it runs alone or in any mix.
//
// The queue lives in device memory (items, their count, per-step counters), so
// a captured hipGraph of k steps stays valid while systems join and leave: only
// the item list is rewritten between launches.  The per-step counters are
// cleared by a kernel inside the graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "common.hpp"
#include "direct.hpp"
#include "direct_core.hpp"
#include "ic_device.hpp"
#include "sim.hpp"

namespace nbody {

namespace {

struct SysDesc {
    int off;         // first record in the arena
    int n;           // real particles
    int n_pad;       // padded to kTile
    int pad_;
    DirectScale sc;  // the system's scaled units
};

struct WorkItem {
    int sys;    // slot in the descriptor table
    int ibase;  // first target (relative to the system)
};

struct BatchLaunch {
    const float4* pos_in;
    float4* pos_out;
    float4* vel;
    float4* acc;
    const SysDesc* sys;
    const WorkItem* items;
    const int* n_items;
    float dt;
    float dts;
};

__global__ void fill_u32_kernel(unsigned* p, int n, unsigned v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// The barriers around the item fetch also separate one item's last reads of
// the LDS tiles from the next item's first writes.
template <int LAW, int K, int BLOCK>
__global__ __launch_bounds__(BLOCK) void batch_step_kernel(BatchLaunch B, unsigned* counter) {
    __shared__ core::SrcTile tiles[2];
    __shared__ int item_sh;
    const int tid = threadIdx.x;
    const int n_items = *B.n_items;
    for (;;) {
        if (tid == 0) item_sh = static_cast<int>(atomicAdd(counter, 1u));
        __syncthreads();
        const int it = item_sh;
        __syncthreads();
        if (it >= n_items) break;
        const WorkItem w = B.items[it];
        const SysDesc d = B.sys[w.sys];
        const float4* src = B.pos_in + d.off;
        float xi[K], yi[K], zi[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = min(w.ibase + tid + k * BLOCK, d.n_pad - 1);
            const float4 q = src[i];
            xi[k] = q.x * d.sc.inv_len;
            yi[k] = q.y * d.sc.inv_len;
            zi[k] = q.z * d.sc.inv_len;
        }
        float sx[K], sy[K], sz[K];
        core::block_forces<LAW, K, BLOCK>(tiles, src, 0, d.n_pad, d.sc, xi, yi, zi, sx, sy, sz);
        for (int k = 0; k < K; ++k) {  // K is 1 or 2: fully unrolled without a pragma
            const int i = w.ibase + tid + k * BLOCK;
            if (i < d.n) {
                const int r = d.off + i;
                const float ax = sx[k] * d.sc.out, ay = sy[k] * d.sc.out, az = sz[k] * d.sc.out;
                float4 v = B.vel[r];
                float4 x = src[i];
                v.x = __builtin_fmaf(ax, B.dt, v.x);
                v.y = __builtin_fmaf(ay, B.dt, v.y);
                v.z = __builtin_fmaf(az, B.dt, v.z);
                x.x = __builtin_fmaf(v.x, B.dts, x.x);
                x.y = __builtin_fmaf(v.y, B.dts, x.y);
                x.z = __builtin_fmaf(v.z, B.dts, x.z);
                B.vel[r] = v;
                B.pos_out[r] = x;
                B.acc[r] = make_float4(ax, ay, az, 0.f);
            }
        }
    }
}

using BatchFn = void (*)(BatchLaunch, unsigned*);

template <int LAW, int K>
BatchFn pick_b(int block) {
    switch (block) {
        case 128: return batch_step_kernel<LAW, K, 128>;
        case 256: return batch_step_kernel<LAW, K, 256>;
        default: return nullptr;
    }
}

template <int LAW>
BatchFn pick_k(int k, int block) {
    switch (k) {
        case 1: return pick_b<LAW, 1>(block);
        case 2: return pick_b<LAW, 2>(block);
        default: return nullptr;
    }
}

BatchFn pick(int law, int k, int block) {
    return law == NBODY_LAW_R ? pick_k<NBODY_LAW_R>(k, block) : pick_k<NBODY_LAW_P>(k, block);
}

bool tuning_valid(const nbody_batch_tuning& t) {
    return (t.i_per_lane == 1 || t.i_per_lane == 2) && (t.block == 128 || t.block == 256) && t.wg_per_cu >= 1 &&
           t.wg_per_cu <= 64;
}

}  // namespace

}  // namespace nbody

using namespace nbody;

struct nbody_batch {
    nbody_config cfg{};
    hipStream_t stream = nullptr;
    int cus = 256;

    size_t capacity = 0;  // records in the arena
    float4* pos[2] = {nullptr, nullptr};
    float4* vel = nullptr;
    float4* acc = nullptr;
    int cur = 0;
    std::map<size_t, size_t> free_list;  // offset -> length (records)

    struct Sys {
        bool live = false;
        size_t off = 0, n = 0, n_pad = 0;
        DirectScale sc{};
        double max_gm = 0.0;
    };
    std::vector<Sys> systems;  // index = id = descriptor slot
    std::vector<int> free_ids;

    SysDesc* d_sys = nullptr;
    size_t sys_cap = 0;
    WorkItem* d_items = nullptr;
    size_t items_cap = 0;
    int* d_n_items = nullptr;
    unsigned* d_counters = nullptr;
    int counters_cap = 0;
    bool items_dirty = true;
    int n_items = 0;
    uint64_t interactions_per_step = 0;

    nbody_batch_tuning tuning{1, 256, 8, {0, 0, 0, 0, 0}, 0.0};

    bool use_graph = true;
    // One captured graph per position-buffer parity (a step of an odd number
    // of steps flips it), each for one (steps, dt, launch shape).
    struct Graph {
        hipGraphExec_t exec = nullptr;
        int steps = -1;
        float dt = 0.f;
        int k = 0, block = 0, wg = 0;
    };
    Graph graph[2];

    double* ic_scratch = nullptr;  // device initial conditions

    // sequential baseline: partial-sum buffer of the plain kernel
    float4* part = nullptr;
    size_t part_elems = 0;

    hipEvent_t ev[2] = {nullptr, nullptr};
    bool timed = false;
    int last_steps = 0;
    uint64_t last_interactions = 0;
};

namespace nbody {
namespace {

// Grow a device buffer; the old one is freed once the steps queued on the
// batch's stream (which may read it) have completed.
nbody_status ensure_dev(void** p, size_t& cap, size_t need, size_t elem, hipStream_t stream) {
    if (need <= cap && *p) return NBODY_OK;
    size_t ncap = std::max<size_t>(need, cap * 2);
    if (ncap == 0) ncap = 1;
    void* q = nullptr;
    NB_HIP(hipMalloc(&q, ncap * elem));
    if (*p) {
        const hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            (void)hipFree(q);
            set_error("hipStreamSynchronize failed: %s", hipGetErrorString(e));
            return NBODY_ERR_DEVICE;
        }
        (void)hipFree(*p);
    }
    *p = q;
    cap = ncap;
    return NBODY_OK;
}

// First fit in the arena's free list; NBODY_ERR_CAPACITY when nothing fits.
nbody_status arena_alloc(nbody_batch* b, size_t len, size_t* off) {
    for (auto it = b->free_list.begin(); it != b->free_list.end(); ++it) {
        if (it->second >= len) {
            *off = it->first;
            const size_t rest = it->second - len;
            const size_t roff = it->first + len;
            b->free_list.erase(it);
            if (rest) b->free_list[roff] = rest;
            return NBODY_OK;
        }
    }
    set_error("batch arena full: %zu records requested", len);
    return NBODY_ERR_CAPACITY;
}

void arena_free(nbody_batch* b, size_t off, size_t len) {
    auto it = b->free_list.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != b->free_list.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        b->free_list.erase(nx);
    }
    if (it != b->free_list.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            b->free_list.erase(it);
        }
    }
}

// Rebuild the work-item list (largest systems first) and upload it.
nbody_status rebuild_items(nbody_batch* b) {
    if (!b->items_dirty) return NBODY_OK;
    const int per = b->tuning.block * b->tuning.i_per_lane;
    std::vector<int> order;
    for (size_t s = 0; s < b->systems.size(); ++s)
        if (b->systems[s].live && b->systems[s].n > 0) order.push_back(static_cast<int>(s));
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return b->systems[a].n_pad > b->systems[c].n_pad; });
    std::vector<WorkItem> items;
    std::vector<SysDesc> descs(b->systems.size());
    uint64_t inter = 0;
    for (size_t s = 0; s < b->systems.size(); ++s) {
        const auto& y = b->systems[s];
        descs[s] = SysDesc{static_cast<int>(y.off), static_cast<int>(y.n), static_cast<int>(y.n_pad), 0, y.sc};
    }
    for (int s : order) {
        const auto& y = b->systems[s];
        inter += static_cast<uint64_t>(y.n) * y.n;
        for (size_t i = 0; i < y.n_pad; i += per) items.push_back(WorkItem{s, static_cast<int>(i)});
    }
    NB_TRY(ensure_dev(reinterpret_cast<void**>(&b->d_items), b->items_cap, std::max<size_t>(items.size(), 1),
                      sizeof(WorkItem), b->stream));
    NB_TRY(ensure_dev(reinterpret_cast<void**>(&b->d_sys), b->sys_cap, std::max<size_t>(descs.size(), 1),
                      sizeof(SysDesc), b->stream));
    // The graph holds the item and descriptor pointers: a reallocation above
    // invalidates it.
    if (!items.empty())
        NB_HIP(hipMemcpyAsync(b->d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                              b->stream));
    if (!descs.empty())
        NB_HIP(hipMemcpyAsync(b->d_sys, descs.data(), descs.size() * sizeof(SysDesc), hipMemcpyHostToDevice,
                              b->stream));
    b->n_items = static_cast<int>(items.size());
    NB_HIP(hipMemcpyAsync(b->d_n_items, &b->n_items, sizeof(int), hipMemcpyHostToDevice, b->stream));
    NB_HIP(hipStreamSynchronize(b->stream));  // host vectors go out of scope
    b->interactions_per_step = inter;
    b->items_dirty = false;
    return NBODY_OK;
}

// A graph exec is only destroyed once its launches have completed (they all
// went to the batch's stream).
void drop_graph(nbody_batch* b) {
    bool any = false;
    for (const auto& g : b->graph) any = any || g.exec;
    if (!any) return;
    (void)hipStreamSynchronize(b->stream);
    for (auto& g : b->graph) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = nbody_batch::Graph{};
    }
}

BatchLaunch launch_for(nbody_batch* b, int parity, float dt) {
    BatchLaunch L{};
    L.pos_in = b->pos[parity];
    L.pos_out = b->pos[parity ^ 1];
    L.vel = b->vel;
    L.acc = b->acc;
    L.sys = b->d_sys;
    L.items = b->d_items;
    L.n_items = b->d_n_items;
    L.dt = dt;
    L.dts = static_cast<float>(static_cast<double>(dt) / b->cfg.scale);
    return L;
}

// Enqueue nsteps persistent launches on the batch stream (captured or not).
nbody_status enqueue_steps(nbody_batch* b, float dt, int nsteps) {
    BatchFn fn = pick(b->cfg.law, b->tuning.i_per_lane, b->tuning.block);
    NB_CHECK(fn, NBODY_ERR_INVALID_ARG, "no batch kernel for this tuning");
    hipLaunchKernelGGL(fill_u32_kernel, dim3((nsteps + 255) / 256), dim3(256), 0, b->stream, b->d_counters, nsteps, 0u);
    const dim3 grid(static_cast<unsigned>(b->tuning.wg_per_cu * b->cus));
    int parity = b->cur;
    for (int s = 0; s < nsteps; ++s) {
        hipLaunchKernelGGL(fn, grid, dim3(b->tuning.block), 0, b->stream, launch_for(b, parity, dt), b->d_counters + s);
        parity ^= 1;
    }
    NB_HIP(hipGetLastError());
    return NBODY_OK;
}

}  // namespace
}  // namespace nbody

extern "C" {

nbody_status nbody_batch_create(const nbody_config* cfg, size_t capacity, nbody_batch** out) {
    clear_error();
    NB_CHECK(out, NBODY_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    NB_TRY(validate_config(cfg));
    NB_CHECK(cfg->method == NBODY_METHOD_DIRECT, NBODY_ERR_UNSUPPORTED, "batches run the direct method");
    NB_CHECK(capacity > 0 && capacity <= (size_t(1) << 30), NBODY_ERR_INVALID_ARG, "capacity must be in [1, 2^30]");
    int count = 0;
    NB_HIP(hipGetDeviceCount(&count));
    NB_CHECK(cfg->device >= 0 && cfg->device < count, NBODY_ERR_INVALID_ARG, "device %d not present", cfg->device);
    NB_HIP(hipSetDevice(cfg->device));
    nbody_batch* b = new (std::nothrow) nbody_batch();
    NB_CHECK(b, NBODY_ERR_OUT_OF_MEMORY, "host allocation failed");
    b->cfg = *cfg;
    auto fail = [&](nbody_status st) {
        nbody_batch_destroy(b);
        return st;
    };
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, cfg->device) == hipSuccess && prop.multiProcessorCount > 0)
        b->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreateWithFlags failed");
        return fail(NBODY_ERR_DEVICE);
    }
    const size_t cap = round_up(capacity, kTile);
    for (float4** p : {&b->pos[0], &b->pos[1], &b->vel, &b->acc}) {
        if (hipMalloc(reinterpret_cast<void**>(p), cap * sizeof(float4)) != hipSuccess) {
            set_error("cannot allocate a batch arena of %zu particles", cap);
            return fail(NBODY_ERR_OUT_OF_MEMORY);
        }
    }
    if (hipMalloc(reinterpret_cast<void**>(&b->d_n_items), sizeof(int)) != hipSuccess) return fail(NBODY_ERR_OUT_OF_MEMORY);
    b->capacity = cap;
    b->free_list[0] = cap;
    if (hipEventCreate(&b->ev[0]) != hipSuccess || hipEventCreate(&b->ev[1]) != hipSuccess) return fail(NBODY_ERR_DEVICE);
    *out = b;
    return NBODY_OK;
}

nbody_status nbody_batch_destroy(nbody_batch* b) {
    clear_error();
    if (!b) return NBODY_OK;
    (void)hipSetDevice(b->cfg.device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (void* p : {static_cast<void*>(b->pos[0]), static_cast<void*>(b->pos[1]), static_cast<void*>(b->vel),
                    static_cast<void*>(b->acc), static_cast<void*>(b->d_sys), static_cast<void*>(b->d_items),
                    static_cast<void*>(b->d_n_items), static_cast<void*>(b->d_counters), static_cast<void*>(b->part),
                    static_cast<void*>(b->ic_scratch)})
        if (p) (void)hipFree(p);
    for (hipEvent_t e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return NBODY_OK;
}

nbody_status nbody_batch_add(nbody_batch* b, const float* pos3, const float* vel3, const double* mass, size_t n,
                             int32_t* id) {
    clear_error();
    NB_CHECK(b && id, NBODY_ERR_INVALID_ARG, "null argument");
    NB_CHECK(n > 0 && pos3 && mass, NBODY_ERR_INVALID_ARG, "a system needs n > 0 particles with positions and masses");
    for (size_t i = 0; i < n; ++i)
        NB_CHECK(mass[i] >= 0 && std::isfinite(mass[i]), NBODY_ERR_INVALID_ARG, "mass[%zu] must be finite and >= 0", i);
    NB_HIP(hipSetDevice(b->cfg.device));
    const size_t n_pad = round_up(n, kTile);
    size_t off = 0;
    NB_TRY(arena_alloc(b, n_pad, &off));
    std::vector<float4> hp(n_pad, make_float4(0.f, 0.f, 0.f, 0.f)), hv(n_pad, make_float4(0.f, 0.f, 0.f, 0.f));
    double max_gm = 0.0;
    float lo[3] = {pos3[0], pos3[1], pos3[2]}, hi[3] = {pos3[0], pos3[1], pos3[2]};
    for (size_t i = 0; i < n; ++i) {
        const double gm = b->cfg.G * mass[i];
        max_gm = std::max(max_gm, std::fabs(gm));
        hp[i] = make_float4(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], static_cast<float>(gm));
        if (vel3) hv[i] = make_float4(vel3[3 * i], vel3[3 * i + 1], vel3[3 * i + 2], 0.f);
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], pos3[3 * i + c]);
            hi[c] = std::max(hi[c], pos3[3 * i + c]);
        }
    }
    const double extent = std::max({static_cast<double>(hi[0]) - lo[0], static_cast<double>(hi[1]) - lo[1],
                                    static_cast<double>(hi[2]) - lo[2]});
    // Enqueued behind any running step: the segment was free, nothing reads it.
    hipError_t e = hipSuccess;
    for (int k = 0; k < 2 && e == hipSuccess; ++k)
        e = hipMemcpyAsync(b->pos[k] + off, hp.data(), n_pad * sizeof(float4), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->vel + off, hv.data(), n_pad * sizeof(float4), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        arena_free(b, off, n_pad);
        set_error("upload failed: %s", hipGetErrorString(e));
        return NBODY_ERR_DEVICE;
    }
    int slot;
    if (!b->free_ids.empty()) {
        slot = b->free_ids.back();
        b->free_ids.pop_back();
    } else {
        slot = static_cast<int>(b->systems.size());
        b->systems.emplace_back();
    }
    auto& y = b->systems[slot];
    y.live = true;
    y.off = off;
    y.n = n;
    y.n_pad = n_pad;
    y.max_gm = max_gm;
    y.sc = direct_scale_for(b->cfg.law, b->cfg.softening, extent, max_gm);
    b->items_dirty = true;
    *id = slot;
    return NBODY_OK;
}

nbody_status nbody_batch_add_plummer(nbody_batch* b, uint64_t seed, size_t n, double a, double total_mass,
                                     int32_t* id) {
    clear_error();
    NB_CHECK(b && id, NBODY_ERR_INVALID_ARG, "null argument");
    NB_CHECK(n > 0 && n <= (size_t(1) << 28) && a > 0 && total_mass > 0, NBODY_ERR_INVALID_ARG,
             "bad Plummer system (n %zu)", n);
    NB_HIP(hipSetDevice(b->cfg.device));
    if (!b->ic_scratch) NB_HIP(hipMalloc(reinterpret_cast<void**>(&b->ic_scratch), ic_device_scratch_doubles() * sizeof(double)));
    const size_t n_pad = round_up(n, kTile);
    size_t off = 0;
    NB_TRY(arena_alloc(b, n_pad, &off));
    const size_t bytes = n_pad * sizeof(float4);
    hipError_t e = hipMemsetAsync(b->pos[b->cur] + off, 0, bytes, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->pos[b->cur ^ 1] + off, 0, bytes, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->vel + off, 0, bytes, b->stream);
    if (e == hipSuccess)
        e = ic_device_plummer(seed, static_cast<int>(n), a, total_mass, b->cfg.G, b->pos[b->cur] + off, b->vel + off,
                              nullptr, nullptr, b->pos[b->cur ^ 1] + off, b->ic_scratch, b->stream);
    if (e != hipSuccess) {
        arena_free(b, off, n_pad);
        set_error("device initial conditions failed: %s", hipGetErrorString(e));
        return NBODY_ERR_DEVICE;
    }
    int slot;
    if (!b->free_ids.empty()) {
        slot = b->free_ids.back();
        b->free_ids.pop_back();
    } else {
        slot = static_cast<int>(b->systems.size());
        b->systems.emplace_back();
    }
    auto& y = b->systems[slot];
    y.live = true;
    y.off = off;
    y.n = n;
    y.n_pad = n_pad;
    y.max_gm = std::fabs(b->cfg.G * total_mass / static_cast<double>(n));
    // Plummer radii are cut at 20 a, so the sphere spans at most 40 a
    y.sc = direct_scale_for(b->cfg.law, b->cfg.softening, 40.0 * a, y.max_gm);
    b->items_dirty = true;
    *id = slot;
    return NBODY_OK;
}

nbody_status nbody_batch_remove(nbody_batch* b, int32_t id) {
    clear_error();
    NB_CHECK(b, NBODY_ERR_INVALID_ARG, "batch is null");
    NB_CHECK(id >= 0 && static_cast<size_t>(id) < b->systems.size() && b->systems[id].live, NBODY_ERR_INVALID_ARG,
             "no system %d", id);
    NB_HIP(hipSetDevice(b->cfg.device));
    NB_HIP(hipStreamSynchronize(b->stream));
    auto& y = b->systems[id];
    arena_free(b, y.off, y.n_pad);
    y = nbody_batch::Sys{};
    b->free_ids.push_back(id);
    b->items_dirty = true;
    return NBODY_OK;
}

nbody_status nbody_batch_step(nbody_batch* b, float dt, int32_t nsteps) {
    clear_error();
    NB_CHECK(b, NBODY_ERR_INVALID_ARG, "batch is null");
    NB_CHECK(nsteps >= 0 && nsteps <= 100000, NBODY_ERR_INVALID_ARG, "nsteps out of range");
    NB_CHECK(std::isfinite(dt), NBODY_ERR_INVALID_ARG, "dt is not finite");
    NB_HIP(hipSetDevice(b->cfg.device));
    b->last_steps = 0;
    b->last_interactions = 0;
    b->timed = false;
    if (nsteps == 0) return NBODY_OK;
    const void* old_items = b->d_items;
    const void* old_sys = b->d_sys;
    NB_TRY(rebuild_items(b));
    if (b->d_items != old_items || b->d_sys != old_sys) drop_graph(b);
    if (b->n_items == 0) return NBODY_OK;
    if (static_cast<int>(nsteps) > b->counters_cap) {
        drop_graph(b);
        NB_HIP(hipStreamSynchronize(b->stream));  // plain launches may still use the counters
        if (b->d_counters) (void)hipFree(b->d_counters);
        b->d_counters = nullptr;
        NB_HIP(hipMalloc(reinterpret_cast<void**>(&b->d_counters), static_cast<size_t>(nsteps) * sizeof(unsigned)));
        b->counters_cap = nsteps;
    }
    NB_HIP(hipEventRecord(b->ev[0], b->stream));
    if (b->use_graph) {
        auto& G = b->graph[b->cur];
        const bool reuse = G.exec && G.steps == nsteps && G.dt == dt && G.k == b->tuning.i_per_lane &&
                           G.block == b->tuning.block && G.wg == b->tuning.wg_per_cu;
        if (!reuse) {
            if (G.exec) {
                // its last launch may still be queued on the stream
                NB_HIP(hipStreamSynchronize(b->stream));
                (void)hipGraphExecDestroy(G.exec);
            }
            G = nbody_batch::Graph{};
            hipGraph_t g = nullptr;
            NB_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            nbody_status st = enqueue_steps(b, dt, nsteps);
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (st != NBODY_OK) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            NB_HIP(e);
            e = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            NB_HIP(e);
            G.steps = nsteps;
            G.dt = dt;
            G.k = b->tuning.i_per_lane;
            G.block = b->tuning.block;
            G.wg = b->tuning.wg_per_cu;
        }
        NB_HIP(hipGraphLaunch(G.exec, b->stream));
    } else {
        NB_TRY(enqueue_steps(b, dt, nsteps));
    }
    NB_HIP(hipEventRecord(b->ev[1], b->stream));
    if (nsteps & 1) b->cur ^= 1;
    b->timed = true;
    b->last_steps = nsteps;
    b->last_interactions = b->interactions_per_step * static_cast<uint64_t>(nsteps);
    return NBODY_OK;
}

nbody_status nbody_batch_step_sequential(nbody_batch* b, float dt, int32_t nsteps, int32_t split_small) {
    clear_error();
    NB_CHECK(b, NBODY_ERR_INVALID_ARG, "batch is null");
    NB_CHECK(nsteps >= 0, NBODY_ERR_INVALID_ARG, "nsteps < 0");
    NB_HIP(hipSetDevice(b->cfg.device));
    NB_TRY(rebuild_items(b));
    b->last_steps = 0;
    b->last_interactions = 0;
    b->timed = false;
    if (nsteps == 0) return NBODY_OK;
    NB_HIP(hipEventRecord(b->ev[0], b->stream));
    // Each galaxy alone, one launch per step of the single-system kernel,
    // waiting for each galaxy before the next: either the plain launch (one
    // workgroup per 256 targets over all sources) or, with split_small, the
    // single-system default that splits the sources of a small system over
    // more workgroups and adds a reduction launch.
    for (const auto& y : b->systems) {
        if (!y.live) continue;
        DirectShape shape = split_small ? direct_default_shape(static_cast<int>(y.n_pad)) : DirectShape{1, 256, 1};
        const size_t split = static_cast<size_t>(direct_effective_split(shape, static_cast<int>(y.n_pad)));
        if (split * y.n_pad > b->part_elems) {
            NB_HIP(hipStreamSynchronize(b->stream));
            if (b->part) (void)hipFree(b->part);
            b->part = nullptr;
            NB_HIP(hipMalloc(reinterpret_cast<void**>(&b->part), split * y.n_pad * sizeof(float4)));
            b->part_elems = split * y.n_pad;
        }
        int parity = b->cur;
        for (int s = 0; s < nsteps; ++s) {
            DirectLaunch L{};
            L.src = b->pos[parity] + y.off;
            L.pos_out = b->pos[parity ^ 1] + y.off;
            L.vel = b->vel + y.off;
            L.acc = b->acc + y.off;
            L.part = b->part;
            L.n = static_cast<int>(y.n);
            L.n_pad = static_cast<int>(y.n_pad);
            L.law = b->cfg.law;
            L.sc = y.sc;
            L.dt = dt;
            L.dts = static_cast<float>(static_cast<double>(dt) / b->cfg.scale);
            NB_HIP(direct_step(L, shape, true, b->stream, nullptr, nullptr));
            parity ^= 1;
        }
        // Steps of one galaxy leave its positions in buffer cur ^ (nsteps & 1);
        // keep every galaxy on the same parity.
        if (nsteps & 1)
            NB_HIP(hipMemcpyAsync(b->pos[b->cur] + y.off, b->pos[b->cur ^ 1] + y.off, y.n_pad * sizeof(float4),
                                  hipMemcpyDeviceToDevice, b->stream));
        NB_HIP(hipStreamSynchronize(b->stream));
    }
    NB_HIP(hipEventRecord(b->ev[1], b->stream));
    b->timed = true;
    b->last_steps = nsteps;
    b->last_interactions = b->interactions_per_step * static_cast<uint64_t>(nsteps);
    return NBODY_OK;
}

nbody_status nbody_batch_synchronize(nbody_batch* b) {
    clear_error();
    NB_CHECK(b, NBODY_ERR_INVALID_ARG, "batch is null");
    NB_HIP(hipSetDevice(b->cfg.device));
    NB_HIP(hipStreamSynchronize(b->stream));
    return NBODY_OK;
}

nbody_status nbody_batch_query(nbody_batch* b, int32_t* done) {
    clear_error();
    NB_CHECK(b && done, NBODY_ERR_INVALID_ARG, "null argument");
    NB_HIP(hipSetDevice(b->cfg.device));
    const hipError_t e = hipStreamQuery(b->stream);
    if (e == hipErrorNotReady) {
        *done = 0;
        return NBODY_OK;
    }
    NB_HIP(e);
    *done = 1;
    return NBODY_OK;
}

nbody_status nbody_batch_read(nbody_batch* b, int32_t id, float* pos, float* vel, float* acc, size_t n) {
    clear_error();
    NB_CHECK(b, NBODY_ERR_INVALID_ARG, "batch is null");
    NB_CHECK(id >= 0 && static_cast<size_t>(id) < b->systems.size() && b->systems[id].live, NBODY_ERR_INVALID_ARG,
             "no system %d", id);
    const auto& y = b->systems[id];
    NB_CHECK(n == y.n, NBODY_ERR_INVALID_ARG, "system %d has %zu particles, not %zu", id, y.n, n);
    NB_HIP(hipSetDevice(b->cfg.device));
    NB_HIP(hipStreamSynchronize(b->stream));
    std::vector<float4> tmp(n);
    auto fetch = [&](const float4* src, float* dst) -> nbody_status {
        if (!dst) return NBODY_OK;
        NB_HIP(hipMemcpy(tmp.data(), src + y.off, n * sizeof(float4), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            dst[3 * i] = tmp[i].x;
            dst[3 * i + 1] = tmp[i].y;
            dst[3 * i + 2] = tmp[i].z;
        }
        return NBODY_OK;
    };
    NB_TRY(fetch(b->pos[b->cur], pos));
    NB_TRY(fetch(b->vel, vel));
    NB_TRY(fetch(b->acc, acc));
    return NBODY_OK;
}

nbody_status nbody_batch_get_info(nbody_batch* b, nbody_batch_info* out) {
    clear_error();
    NB_CHECK(b && out, NBODY_ERR_INVALID_ARG, "null argument");
    std::memset(out, 0, sizeof *out);
    NB_TRY(rebuild_items(b));
    for (const auto& y : b->systems)
        if (y.live) {
            out->systems += 1;
            out->particles += y.n;
        }
    out->interactions_per_step = b->interactions_per_step;
    out->work_items = b->n_items;
    out->capacity = b->capacity;
    size_t free_total = 0, largest = 0;
    for (const auto& f : b->free_list) {
        free_total += f.second;
        largest = std::max(largest, f.second);
    }
    out->free_records = free_total;
    out->largest_free = largest;
    out->tuning = b->tuning;
    return NBODY_OK;
}

nbody_status nbody_batch_set_tuning(nbody_batch* b, const nbody_batch_tuning* t) {
    clear_error();
    NB_CHECK(b && t, NBODY_ERR_INVALID_ARG, "null argument");
    NB_CHECK(tuning_valid(*t), NBODY_ERR_INVALID_ARG, "invalid batch tuning (i_per_lane %d, block %d, wg_per_cu %d)",
             t->i_per_lane, t->block, t->wg_per_cu);
    if (t->i_per_lane != b->tuning.i_per_lane || t->block != b->tuning.block) b->items_dirty = true;
    b->tuning = *t;
    return NBODY_OK;
}

nbody_status nbody_batch_set_graph(nbody_batch* b, int32_t enable) {
    clear_error();
    NB_CHECK(b, NBODY_ERR_INVALID_ARG, "batch is null");
    b->use_graph = enable != 0;
    if (!b->use_graph) drop_graph(b);
    return NBODY_OK;
}

nbody_status nbody_batch_get_timing(nbody_batch* b, nbody_timing* out) {
    clear_error();
    NB_CHECK(b && out, NBODY_ERR_INVALID_ARG, "null argument");
    std::memset(out, 0, sizeof *out);
    NB_HIP(hipSetDevice(b->cfg.device));
    NB_HIP(hipStreamSynchronize(b->stream));
    out->steps = b->last_steps;
    out->interactions = b->last_interactions;
    out->i_per_lane = b->tuning.i_per_lane;
    out->block = b->tuning.block;
    out->split = 1;
    if (b->timed) {
        float ms = 0;
        NB_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
        out->step_ms = ms;
        out->force_ms = ms;
        out->force_launches = b->last_steps;
    }
    return NBODY_OK;
}

// Time one step of each candidate shape on the live batch and keep the
// fastest.  The state is saved before and restored after, so tuning leaves the
// batch as it found it.
nbody_status nbody_batch_tune(nbody_batch* b, float dt, nbody_batch_tuning* out) {
    clear_error();
    NB_CHECK(b, NBODY_ERR_INVALID_ARG, "batch is null");
    NB_HIP(hipSetDevice(b->cfg.device));
    NB_TRY(rebuild_items(b));
    if (b->n_items == 0) {
        if (out) *out = b->tuning;
        return NBODY_OK;
    }
    const size_t cap = b->capacity;
    float4 *sp = nullptr, *sv = nullptr;
    NB_HIP(hipMalloc(reinterpret_cast<void**>(&sp), cap * sizeof(float4)));
    if (hipMalloc(reinterpret_cast<void**>(&sv), cap * sizeof(float4)) != hipSuccess) {
        (void)hipFree(sp);
        set_error("no memory for the tuner's state copy");
        return NBODY_ERR_OUT_OF_MEMORY;
    }
    const int cur0 = b->cur;
    nbody_status st = NBODY_OK;
    auto save = [&]() -> nbody_status {
        NB_HIP(hipMemcpyAsync(sp, b->pos[cur0], cap * sizeof(float4), hipMemcpyDeviceToDevice, b->stream));
        NB_HIP(hipMemcpyAsync(sv, b->vel, cap * sizeof(float4), hipMemcpyDeviceToDevice, b->stream));
        return NBODY_OK;
    };
    auto restore = [&]() -> nbody_status {
        NB_HIP(hipMemcpyAsync(b->pos[cur0], sp, cap * sizeof(float4), hipMemcpyDeviceToDevice, b->stream));
        NB_HIP(hipMemcpyAsync(b->vel, sv, cap * sizeof(float4), hipMemcpyDeviceToDevice, b->stream));
        NB_HIP(hipStreamSynchronize(b->stream));
        b->cur = cur0;
        return NBODY_OK;
    };
    st = save();
    const nbody_batch_tuning keep = b->tuning;
    const bool keep_graph = b->use_graph;
    b->use_graph = false;
    nbody_batch_tuning best = keep;
    double best_ms = 1e300;
    const int ks[] = {1, 2};
    const int blocks[] = {256, 128};
    const int wgs[] = {2, 4, 8, 16};
    for (int k : ks) {
        for (int blk : blocks) {
            for (int wg : wgs) {
                if (st != NBODY_OK) break;
                nbody_batch_tuning t{k, blk, wg, {0, 0, 0, 0, 0}, 0.0};
                b->tuning = t;
                b->items_dirty = true;
                st = rebuild_items(b);
                if (st != NBODY_OK) break;
                // one warm-up step, then one timed step
                for (int rep = 0; rep < 2 && st == NBODY_OK; ++rep) {
                    st = nbody_batch_step(b, dt, 1);
                    if (st == NBODY_OK) {
                        nbody_timing tm;
                        st = nbody_batch_get_timing(b, &tm);
                        if (st == NBODY_OK && rep == 1 && tm.step_ms < best_ms) {
                            best_ms = tm.step_ms;
                            best = t;
                            best.ms_per_step = tm.step_ms;
                        }
                    }
                    if (st == NBODY_OK) st = restore();
                }
            }
        }
    }
    (void)hipFree(sp);
    (void)hipFree(sv);
    b->use_graph = keep_graph;
    b->tuning = (st == NBODY_OK) ? best : keep;
    b->items_dirty = true;
    drop_graph(b);
    if (st != NBODY_OK) return st;
    NB_TRY(rebuild_items(b));
    if (out) *out = b->tuning;
    return NBODY_OK;
}

}  // extern "C"
