// C ABI of the simulation handle: the INBodySim drop-in
// (src/Sim/INBodySim.hpp:19-28, INBodySim.cpp:7-39).
//
// A handle owns its device state (structure of arrays, padded to the direct
// kernel's tile with massless dummies) and one HIP stream.  Positions live in
// two buffers that alternate each step, so the force pass of step k reads the
// positions of step k while it writes step k+1's; nothing is copied between
// steps and nothing returns to the host unless a read asks for it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "bh.hpp"
#include "common.hpp"
#include "diag.hpp"
#include "direct.hpp"
#include "ic.hpp"
#include "ic_device.hpp"
#include "sim.hpp"

static_assert(sizeof(nbody_particle) == 104, "nbody_particle must match the reference's 104-byte Particle");
static_assert(offsetof(nbody_particle, velocity) == 48, "velocity offset");
static_assert(offsetof(nbody_particle, forces) == 72, "forces offset");
static_assert(offsetof(nbody_particle, mass) == 96, "mass offset");
static_assert(sizeof(nbody_bh_cell) == 48 && offsetof(nbody_bh_cell, mass) == 32 &&
                  offsetof(nbody_bh_cell, count) == 44,
              "nbody_bh_cell layout (mirrored by BH_CELL_DTYPE in the Python package)");

namespace nbody {

namespace {
thread_local std::string g_error;
}

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}
const char* last_error() { return g_error.c_str(); }
void clear_error() { g_error.clear(); }

nbody_status validate_config(const nbody_config* c) {
    NB_CHECK(c, NBODY_ERR_INVALID_ARG, "config is null");
    NB_CHECK(c->method == NBODY_METHOD_DIRECT || c->method == NBODY_METHOD_BARNES_HUT, NBODY_ERR_INVALID_ARG,
             "unknown method %d", c->method);
    NB_CHECK(c->law == NBODY_LAW_R || c->law == NBODY_LAW_P, NBODY_ERR_INVALID_ARG, "unknown law %d", c->law);
    NB_CHECK(std::isfinite(c->G) && c->G >= 0, NBODY_ERR_INVALID_ARG, "G must be finite and >= 0");
    NB_CHECK(std::isfinite(c->softening) && c->softening >= 0, NBODY_ERR_INVALID_ARG,
             "softening must be finite and >= 0");
    NB_CHECK(c->law != NBODY_LAW_P || c->softening > 0, NBODY_ERR_INVALID_ARG,
             "law P needs eps^2 > 0 (softening)");
    NB_CHECK(std::isfinite(c->scale) && c->scale > 0, NBODY_ERR_INVALID_ARG, "scale must be finite and > 0");
    NB_CHECK(std::isfinite(c->theta) && c->theta >= 0, NBODY_ERR_INVALID_ARG, "theta must be finite and >= 0");
    NB_CHECK(c->root_mode == NBODY_ROOT_DYNAMIC || c->root_mode == NBODY_ROOT_FIXED, NBODY_ERR_INVALID_ARG,
             "unknown root mode %d", c->root_mode);
    if (c->root_mode == NBODY_ROOT_FIXED) {
        NB_CHECK(std::isfinite(c->root_size) && c->root_size > 0 && std::isfinite(c->root_lo[0]) &&
                     std::isfinite(c->root_lo[1]) && std::isfinite(c->root_lo[2]),
                 NBODY_ERR_INVALID_ARG, "fixed root needs a finite lower corner and size > 0");
    }
    if (c->method == NBODY_METHOD_BARNES_HUT && c->root_mode == NBODY_ROOT_FIXED) NB_TRY(bh_fixed_root(c, nullptr));
    if (c->tune_i_per_lane || c->tune_block || c->tune_split) {
        DirectShape s{c->tune_i_per_lane ? c->tune_i_per_lane : 4, c->tune_block ? c->tune_block : 256,
                      c->tune_split ? c->tune_split : 1};
        NB_CHECK(direct_shape_valid(s), NBODY_ERR_INVALID_ARG,
                 "invalid direct launch shape (i_per_lane %d, block %d, split %d)", s.i_per_lane, s.block,
                 s.split);
    }
    return NBODY_OK;
}

namespace {

void free_dev(void* p) {
    if (p) (void)hipFree(p);
}

template <class T>
nbody_status dev_alloc(T** p, size_t count) {
    free_dev(*p);
    *p = nullptr;
    if (count == 0) return NBODY_OK;
    NB_HIP(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
    return NBODY_OK;
}

// A graph exec is only destroyed once its launches have completed: every
// launch went to the handle's stream, so waiting for the stream suffices.
void drop_graphs(nbody_sim* s) {
    bool any = false;
    for (const auto& g : s->graphs) any = any || g.exec;
    if (!any) return;
    (void)hipStreamSynchronize(s->stream);
    for (auto& g : s->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = nbody_sim::StepGraph{};
    }
}

nbody_status release_state(nbody_sim* s) {
    free_dev(s->pos[0]);
    free_dev(s->pos[1]);
    free_dev(s->vel);
    free_dev(s->acc);
    free_dev(s->part);
    free_dev(s->mass_dev);
    free_dev(s->gm_dev);
    free_dev(s->diag_buf);
    free_dev(s->ic_scratch);
    s->pos[0] = s->pos[1] = s->vel = s->acc = s->part = nullptr;
    s->mass_dev = s->gm_dev = s->diag_buf = s->ic_scratch = nullptr;
    s->capacity = 0;
    s->part_elems = 0;
    s->diag_elems = 0;
    return NBODY_OK;
}

DirectShape shape_for(const nbody_sim* s) {
    DirectShape d = direct_default_shape(static_cast<int>(s->n_pad));
    if (s->cfg.tune_i_per_lane) d.i_per_lane = s->cfg.tune_i_per_lane;
    if (s->cfg.tune_block) d.block = s->cfg.tune_block;
    if (s->cfg.tune_split) d.split = s->cfg.tune_split;
    return d;
}

// Size the device buffers for n particles (keeping larger ones) and set the
// launch shape; the contents are left to the caller.
nbody_status prepare_state(nbody_sim* s, size_t n) {
    NB_HIP(hipSetDevice(s->cfg.device));
    NB_HIP(hipStreamSynchronize(s->stream));
    drop_graphs(s);  // they hold the old sizes, pointers and launch arguments
    const size_t n_pad = round_up(std::max<size_t>(n, 1), kTile);
    if (n_pad > s->capacity) {
        release_state(s);
        NB_TRY(dev_alloc(&s->pos[0], n_pad));
        NB_TRY(dev_alloc(&s->pos[1], n_pad));
        NB_TRY(dev_alloc(&s->vel, n_pad));
        NB_TRY(dev_alloc(&s->acc, n_pad));
        NB_TRY(dev_alloc(&s->mass_dev, n_pad));
        NB_TRY(dev_alloc(&s->gm_dev, n_pad));
        s->capacity = n_pad;
    }
    s->n = n;
    s->n_pad = n_pad;
    s->cur = 0;
    s->shape = shape_for(s);
    const size_t split = static_cast<size_t>(direct_effective_split(s->shape, static_cast<int>(n_pad)));
    // Partial sums serve the direct kernel only (a Barnes-Hut handle uses it
    // for a single particle alone).
    const bool direct = s->cfg.method == NBODY_METHOD_DIRECT || n < 2;
    const size_t need_part = direct ? std::max<size_t>(split, 1) * n_pad : 0;
    if (need_part > s->part_elems) {
        NB_TRY(dev_alloc(&s->part, need_part));
        s->part_elems = need_part;
    }
    return NBODY_OK;
}

// The largest coordinate range of n positions (x, y, z in .x .y .z).
double extent_of(const float4* p, size_t n) {
    if (n == 0) return 0.0;
    float lo[3] = {p[0].x, p[0].y, p[0].z}, hi[3] = {p[0].x, p[0].y, p[0].z};
    for (size_t i = 1; i < n; ++i) {
        const float c[3] = {p[i].x, p[i].y, p[i].z};
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], c[a]);
            hi[a] = std::max(hi[a], c[a]);
        }
    }
    double e = 0.0;
    for (int a = 0; a < 3; ++a) e = std::max(e, static_cast<double>(hi[a]) - static_cast<double>(lo[a]));
    return e;
}

// The direct sum's scaled units for the handle's state (extent and max_gm
// of the initial state; a system that later grows or shrinks by several
// orders of magnitude keeps them, which the pair arithmetic's range allows).
void set_scale(nbody_sim* s, double extent) {
    s->dscale = direct_scale_for(s->cfg.law, s->cfg.softening, extent, s->max_gm);
}

// Masses must be finite and >= 0 (the direct sum folds G m into a square).
nbody_status check_masses(const double* mass, size_t n) {
    for (size_t i = 0; i < n; ++i)
        NB_CHECK(std::isfinite(mass[i]) && mass[i] >= 0, NBODY_ERR_INVALID_ARG, "mass[%zu] = %g must be finite and >= 0",
                 i, mass[i]);
    return NBODY_OK;
}

// Staging of nbody_field_noalloc: the points (first half) and their
// accelerations (second half), 2 x NBODY_FIELD_NOALLOC_MAX_POINTS records,
// 2 MiB held by every device the library runs on until the process ends.
// In host code `field_scratch_array` names a host-side stand-in, not device
// memory: only the address hipGetSymbolAddress returns (per device) is used.
__device__ __align__(256) float4 field_scratch_array[2 * NBODY_FIELD_NOALLOC_MAX_POINTS];

constexpr int kMaxDevices = 64;

// The device address of this device's field_scratch_array, looked up once per
// device with that device current (each device holds its own copy).
nbody_status field_scratch(int device, float4** out) {
    static std::mutex mu;
    static float4* addr[kMaxDevices] = {};
    NB_CHECK(device >= 0 && device < kMaxDevices, NBODY_ERR_UNSUPPORTED, "device %d above %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void* p = nullptr;
        NB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(field_scratch_array)));
        NB_CHECK(p, NBODY_ERR_DEVICE, "hipGetSymbolAddress returned no address");
        addr[device] = static_cast<float4*>(p);
    }
    *out = addr[device];
    return NBODY_OK;
}

// Upload a packed state (n real records) into freshly sized device buffers.
nbody_status upload_state(nbody_sim* s, const std::vector<float4>& pos, const std::vector<float4>& vel,
                          const std::vector<double>& mass, size_t n) {
    NB_TRY(prepare_state(s, n));
    const size_t n_pad = s->n_pad;
    std::vector<float4> hp(n_pad, make_float4(0.f, 0.f, 0.f, 0.f)), hv(n_pad, make_float4(0.f, 0.f, 0.f, 0.f));
    std::vector<double> hm(n_pad, 0.0), hg(n_pad, 0.0);
    std::copy(pos.begin(), pos.begin() + n, hp.begin());
    std::copy(vel.begin(), vel.begin() + n, hv.begin());
    s->max_gm = 0.0;
    for (size_t i = 0; i < n; ++i) {
        hm[i] = mass[i];
        hg[i] = s->cfg.G * mass[i];
        s->max_gm = std::max(s->max_gm, std::fabs(hg[i]));
    }
    set_scale(s, extent_of(pos.data(), n));
    NB_HIP(hipMemcpyAsync(s->pos[0], hp.data(), n_pad * sizeof(float4), hipMemcpyHostToDevice, s->stream));
    NB_HIP(hipMemcpyAsync(s->pos[1], hp.data(), n_pad * sizeof(float4), hipMemcpyHostToDevice, s->stream));
    NB_HIP(hipMemcpyAsync(s->vel, hv.data(), n_pad * sizeof(float4), hipMemcpyHostToDevice, s->stream));
    NB_HIP(hipMemsetAsync(s->acc, 0, n_pad * sizeof(float4), s->stream));
    NB_HIP(hipMemcpyAsync(s->mass_dev, hm.data(), n_pad * sizeof(double), hipMemcpyHostToDevice, s->stream));
    NB_HIP(hipMemcpyAsync(s->gm_dev, hg.data(), n_pad * sizeof(double), hipMemcpyHostToDevice, s->stream));
    NB_HIP(hipStreamSynchronize(s->stream));
    s->mass_host.assign(mass.begin(), mass.begin() + n);
    s->initialised = true;
    s->acc_valid = false;
    if (s->bh) NB_TRY(bh_resize(s->bh, n, s));
    return NBODY_OK;
}

nbody_status ensure_events(nbody_sim* s, size_t pairs) {
    while (s->events.size() < 2 * pairs) {
        hipEvent_t e;
        NB_HIP(hipEventCreate(&e));
        s->events.push_back(e);
    }
    return NBODY_OK;
}

DirectLaunch launch_for(nbody_sim* s, float dt) {
    DirectLaunch L{};
    L.src = s->pos[s->cur];
    L.pos_out = s->pos[s->cur ^ 1];
    L.vel = s->vel;
    L.acc = s->acc;
    L.part = s->part;
    L.n = static_cast<int>(s->n);
    L.n_pad = static_cast<int>(s->n_pad);
    L.law = s->cfg.law;
    L.sc = s->dscale;
    L.dt = dt;
    L.dts = static_cast<float>(static_cast<double>(dt) / s->cfg.scale);
    return L;
}

// One force evaluation (kick: a full step).  ev_b/ev_e bracket the force kernel(s).
nbody_status force_pass(nbody_sim* s, float dt, bool kick, hipEvent_t ev_b, hipEvent_t ev_e) {
    if (s->cfg.method == NBODY_METHOD_BARNES_HUT) {
        NB_TRY(bh_step(s, dt, kick, ev_b, ev_e));
    } else {
        DirectLaunch L = launch_for(s, dt);
        NB_HIP(direct_step(L, s->shape, kick, s->stream, ev_b, ev_e));
        s->step_interactions += static_cast<uint64_t>(s->n) * static_cast<uint64_t>(s->n);
    }
    if (kick) s->cur ^= 1;
    s->acc_valid = true;
    return NBODY_OK;
}

}  // namespace

}  // namespace nbody

using namespace nbody;

extern "C" {

void nbody_config_default(nbody_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->method = NBODY_METHOD_DIRECT;
    c->law = NBODY_LAW_P;
    c->G = 1.0;
    c->softening = 1e-4;
    c->scale = 1.0;
    c->theta = 0.5;
    c->root_mode = NBODY_ROOT_DYNAMIC;
    c->root_size = 0.f;
    c->device = 0;
}

void nbody_config_reference(nbody_config* c, int32_t method) {
    if (!c) return;
    nbody_config_default(c);
    c->method = method;
    c->law = NBODY_LAW_R;
    c->G = 6.674e-11;        // Physics.hpp:9
    c->softening = 1e1;      // Physics.hpp:10
    c->scale = 20 * 1.15e12; // Physics.hpp:13,16 (StarSystemScale = 20 AU)
    c->theta = 2.0;          // Octree.cpp:5
    c->root_mode = NBODY_ROOT_FIXED;  // BarnesHut.cpp:14-19
    c->root_lo[0] = c->root_lo[1] = c->root_lo[2] = -4000.f;
    c->root_size = 8000.f;
}

nbody_status nbody_create(const nbody_config* cfg, nbody_sim** out) {
    clear_error();
    NB_CHECK(out, NBODY_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    NB_TRY(validate_config(cfg));
    int count = 0;
    NB_HIP(hipGetDeviceCount(&count));
    NB_CHECK(cfg->device >= 0 && cfg->device < count, NBODY_ERR_INVALID_ARG, "device %d not present (%d visible)",
             cfg->device, count);
    NB_HIP(hipSetDevice(cfg->device));
    nbody_sim* s = new (std::nothrow) nbody_sim();
    NB_CHECK(s, NBODY_ERR_OUT_OF_MEMORY, "host allocation failed");
    s->cfg = *cfg;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete s;
        set_error("hipStreamCreateWithFlags failed: %s", hipGetErrorString(e));
        return NBODY_ERR_DEVICE;
    }
    if (cfg->method == NBODY_METHOD_BARNES_HUT) {
        nbody_status st = bh_create(s);
        if (st != NBODY_OK) {
            (void)hipStreamDestroy(s->stream);
            delete s;
            return st;
        }
    }
    *out = s;
    return NBODY_OK;
}

nbody_status nbody_destroy(nbody_sim* s) {
    clear_error();
    if (!s) return NBODY_OK;
    (void)hipSetDevice(s->cfg.device);
    (void)hipStreamSynchronize(s->stream);
    drop_graphs(s);
    if (s->bh) bh_destroy(s->bh);
    release_state(s);
    for (hipEvent_t e : s->events) (void)hipEventDestroy(e);
    for (hipEvent_t e : s->tree_ms_pending) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(s->stream);
    delete s;
    return NBODY_OK;
}

nbody_status nbody_init(nbody_sim* s, const nbody_particle* p, size_t n) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    NB_CHECK(p || n == 0, NBODY_ERR_INVALID_ARG, "particles is null");
    NB_CHECK(n <= (size_t(1) << 30), NBODY_ERR_CAPACITY, "n = %zu exceeds 2^30", n);
    std::vector<float4> pos(n), vel(n);
    std::vector<double> mass(n);
    for (size_t i = 0; i < n; ++i) mass[i] = p[i].mass;
    NB_TRY(check_masses(mass.data(), n));
    for (size_t i = 0; i < n; ++i) {
        const double gm = s->cfg.G * p[i].mass;
        pos[i] = make_float4(p[i].position[0], p[i].position[1], p[i].position[2], static_cast<float>(gm));
        vel[i] = make_float4(static_cast<float>(p[i].velocity[0]), static_cast<float>(p[i].velocity[1]),
                             static_cast<float>(p[i].velocity[2]), 0.f);
        mass[i] = p[i].mass;
    }
    return upload_state(s, pos, vel, mass, n);
}

nbody_status nbody_init_soa(nbody_sim* s, const float* pos3, const float* vel3, const double* mass, size_t n) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    NB_CHECK((pos3 && mass) || n == 0, NBODY_ERR_INVALID_ARG, "pos or mass is null");
    NB_CHECK(n <= (size_t(1) << 30), NBODY_ERR_CAPACITY, "n = %zu exceeds 2^30", n);
    NB_TRY(check_masses(mass, n));
    std::vector<float4> pos(n), vel(n);
    std::vector<double> m(n);
    for (size_t i = 0; i < n; ++i) {
        pos[i] = make_float4(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], static_cast<float>(s->cfg.G * mass[i]));
        vel[i] = vel3 ? make_float4(vel3[3 * i], vel3[3 * i + 1], vel3[3 * i + 2], 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
        m[i] = mass[i];
    }
    return upload_state(s, pos, vel, m, n);
}

nbody_status nbody_init_plummer(nbody_sim* s, uint64_t seed, size_t n, double a, double total_mass) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    NB_CHECK(n >= 1 && n <= (size_t(1) << 30), NBODY_ERR_INVALID_ARG, "n = %zu out of range", n);
    NB_CHECK(a > 0 && total_mass > 0, NBODY_ERR_INVALID_ARG, "bad Plummer parameters");
    NB_TRY(prepare_state(s, n));
    if (!s->ic_scratch) NB_TRY(dev_alloc(&s->ic_scratch, ic_device_scratch_doubles()));
    const size_t bytes = s->n_pad * sizeof(float4);
    NB_HIP(hipMemsetAsync(s->pos[0], 0, bytes, s->stream));
    NB_HIP(hipMemsetAsync(s->pos[1], 0, bytes, s->stream));
    NB_HIP(hipMemsetAsync(s->vel, 0, bytes, s->stream));
    NB_HIP(hipMemsetAsync(s->acc, 0, bytes, s->stream));
    NB_HIP(hipMemsetAsync(s->mass_dev, 0, s->n_pad * sizeof(double), s->stream));
    NB_HIP(hipMemsetAsync(s->gm_dev, 0, s->n_pad * sizeof(double), s->stream));
    // Everything is generated on the handle's stream; nothing waits for it
    // here (steps and reads queue behind it).
    NB_HIP(ic_device_plummer(seed, static_cast<int>(n), a, total_mass, s->cfg.G, s->pos[0], s->vel, s->mass_dev,
                             s->gm_dev, s->pos[1], s->ic_scratch, s->stream));
    const double m = total_mass / static_cast<double>(n);
    s->mass_host.assign(n, m);
    s->max_gm = std::fabs(s->cfg.G * m);
    set_scale(s, 40.0 * a);  // radii are cut at 20 a
    s->initialised = true;
    s->acc_valid = false;
    if (s->bh) NB_TRY(bh_resize(s->bh, n, s));
    return NBODY_OK;
}

// nbody_step as one graph launch: the counter reset and the nsteps force
// passes are captured on the first call with this (nsteps, dt, parity) and
// replayed afterwards.  The capture records kernels only (the counters are
// cleared by a kernel, never a memset node).
nbody_status step_graph(nbody_sim* s, float dt, int32_t nsteps) {
    auto& G = s->graphs[s->cur];
    const int parity = s->cur;
    if (!(G.exec && G.steps == nsteps && G.dt == dt)) {
        if (G.exec) {
            // its last launch may still be queued on the stream
            NB_HIP(hipStreamSynchronize(s->stream));
            (void)hipGraphExecDestroy(G.exec);
        }
        G = nbody_sim::StepGraph{};
        hipGraph_t g = nullptr;
        NB_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        nbody_status st = NBODY_OK;
        if (s->bh) st = bh_reset_counters(s->bh, s->stream);
        for (int k = 0; k < nsteps && st == NBODY_OK; ++k) st = force_pass(s, dt, true, nullptr, nullptr);
        const hipError_t e = hipStreamEndCapture(s->stream, &g);
        s->cur = parity;  // nothing ran yet: the launch below advances the state
        if (st != NBODY_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        NB_HIP(e);
        const hipError_t ei = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        NB_HIP(ei);
        G.steps = nsteps;
        G.dt = dt;
    }
    NB_HIP(hipGraphLaunch(G.exec, s->stream));
    s->cur = parity ^ (nsteps & 1);
    s->acc_valid = true;
    s->step_interactions = s->bh ? 0 : static_cast<uint64_t>(nsteps) * s->n * s->n;
    s->last.steps = nsteps;
    s->last.interactions = s->step_interactions;
    s->last.i_per_lane = s->shape.i_per_lane;
    s->last.block = s->shape.block;
    s->last.split = direct_effective_split(s->shape, static_cast<int>(s->n_pad));
    s->timed_steps = 0;
    return NBODY_OK;
}

nbody_status nbody_init_seeded(nbody_sim* s, int32_t seeder, uint64_t seed, size_t n, double scale) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    NB_CHECK(seeder >= NBODY_SEEDER_RANDOM && seeder <= NBODY_SEEDER_STAR_SYSTEM, NBODY_ERR_INVALID_ARG,
             "unknown seeder %d", seeder);
    NB_CHECK(n >= 1 && n <= (size_t(1) << 30), NBODY_ERR_INVALID_ARG, "n = %zu out of range", n);
    NB_CHECK(std::isfinite(scale) && scale > 0, NBODY_ERR_INVALID_ARG, "scale must be positive");
    NB_TRY(prepare_state(s, n));
    const size_t bytes = s->n_pad * sizeof(float4);
    NB_HIP(hipMemsetAsync(s->pos[0], 0, bytes, s->stream));
    NB_HIP(hipMemsetAsync(s->pos[1], 0, bytes, s->stream));
    NB_HIP(hipMemsetAsync(s->vel, 0, bytes, s->stream));
    NB_HIP(hipMemsetAsync(s->acc, 0, bytes, s->stream));
    NB_HIP(hipMemsetAsync(s->mass_dev, 0, s->n_pad * sizeof(double), s->stream));
    NB_HIP(hipMemsetAsync(s->gm_dev, 0, s->n_pad * sizeof(double), s->stream));
    SeederTable T{};
    ic_seeder_table(seed, &T);
    NB_HIP(ic_device_seeded(seeder, seed, static_cast<int>(n), static_cast<float>(scale), s->cfg.G, T, s->pos[0],
                            s->vel, s->mass_dev, s->gm_dev, s->pos[1], s->stream));
    s->mass_host.resize(n);
    std::vector<float4> pos_host(n);
    NB_HIP(hipMemcpyAsync(s->mass_host.data(), s->mass_dev, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    NB_HIP(hipMemcpyAsync(pos_host.data(), s->pos[0], n * sizeof(float4), hipMemcpyDeviceToHost, s->stream));
    NB_HIP(hipStreamSynchronize(s->stream));
    s->max_gm = 0.0;
    for (double m : s->mass_host) s->max_gm = std::max(s->max_gm, std::fabs(s->cfg.G * m));
    set_scale(s, extent_of(pos_host.data(), n));
    s->initialised = true;
    s->acc_valid = false;
    if (s->bh) NB_TRY(bh_resize(s->bh, n, s));
    return NBODY_OK;
}

nbody_status nbody_step(nbody_sim* s, float dt, int32_t nsteps) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    NB_CHECK(s->initialised, NBODY_ERR_NOT_INITIALISED, "nbody_step before nbody_init");
    NB_CHECK(nsteps >= 0, NBODY_ERR_INVALID_ARG, "nsteps < 0");
    NB_CHECK(std::isfinite(dt), NBODY_ERR_INVALID_ARG, "dt is not finite");
    NB_HIP(hipSetDevice(s->cfg.device));
    s->last = nbody_timing{};
    s->step_interactions = 0;
    if (!s->tree_ms_pending.empty()) {
        NB_HIP(hipStreamSynchronize(s->stream));
        for (hipEvent_t e : s->tree_ms_pending) (void)hipEventDestroy(e);
        s->tree_ms_pending.clear();
    }
    if (s->use_graph && !s->timing && nsteps > 0 && s->n > 0) return step_graph(s, dt, nsteps);
    if (s->bh) NB_TRY(bh_reset_counters(s->bh, s->stream));
    if (s->n == 0) return NBODY_OK;
    if (s->timing) NB_TRY(ensure_events(s, static_cast<size_t>(nsteps) + 1));
    if (s->timing) NB_HIP(hipEventRecord(s->events[0], s->stream));
    for (int k = 0; k < nsteps; ++k) {
        hipEvent_t b = s->timing ? s->events[2 + 2 * k] : nullptr;
        hipEvent_t e = s->timing ? s->events[3 + 2 * k] : nullptr;
        NB_TRY(force_pass(s, dt, true, b, e));
    }
    if (s->timing) NB_HIP(hipEventRecord(s->events[1], s->stream));
    s->last.steps = nsteps;
    s->last.force_launches = s->timing ? nsteps : 0;
    s->last.interactions = s->step_interactions;
    s->last.i_per_lane = s->shape.i_per_lane;
    s->last.block = s->shape.block;
    s->last.split = direct_effective_split(s->shape, static_cast<int>(s->n_pad));
    s->timed_steps = s->timing ? nsteps : 0;
    return NBODY_OK;
}

nbody_status nbody_synchronize(nbody_sim* s) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    NB_HIP(hipSetDevice(s->cfg.device));
    NB_HIP(hipStreamSynchronize(s->stream));
    return NBODY_OK;
}

nbody_status nbody_query(nbody_sim* s, int32_t* done) {
    clear_error();
    NB_CHECK(s && done, NBODY_ERR_INVALID_ARG, "null argument");
    NB_HIP(hipSetDevice(s->cfg.device));
    const hipError_t e = hipStreamQuery(s->stream);
    if (e == hipErrorNotReady) {
        *done = 0;
        return NBODY_OK;
    }
    NB_HIP(e);
    *done = 1;
    return NBODY_OK;
}

nbody_status nbody_read_soa(nbody_sim* s, float* pos, float* vel, float* acc, size_t n) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    NB_CHECK(s->initialised, NBODY_ERR_NOT_INITIALISED, "read before nbody_init");
    NB_CHECK(n == s->n, NBODY_ERR_INVALID_ARG, "n = %zu does not match the state's %zu particles", n, s->n);
    NB_HIP(hipSetDevice(s->cfg.device));
    std::vector<float4> tmp(n);
    auto fetch = [&](const float4* src, float* dst) -> nbody_status {
        if (!dst || n == 0) return NBODY_OK;
        NB_HIP(hipMemcpyAsync(tmp.data(), src, n * sizeof(float4), hipMemcpyDeviceToHost, s->stream));
        NB_HIP(hipStreamSynchronize(s->stream));
        for (size_t i = 0; i < n; ++i) {
            dst[3 * i] = tmp[i].x;
            dst[3 * i + 1] = tmp[i].y;
            dst[3 * i + 2] = tmp[i].z;
        }
        return NBODY_OK;
    };
    NB_HIP(hipStreamSynchronize(s->stream));
    NB_TRY(fetch(s->pos[s->cur], pos));
    NB_TRY(fetch(s->vel, vel));
    NB_TRY(fetch(s->acc, acc));
    return NBODY_OK;
}

nbody_status nbody_read(nbody_sim* s, nbody_particle* p, size_t n, uint32_t fields) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    NB_CHECK(p || n == 0, NBODY_ERR_INVALID_ARG, "particles is null");
    NB_CHECK((fields & ~15u) == 0, NBODY_ERR_INVALID_ARG, "unknown field bits 0x%x", fields);
    std::vector<float> pos(3 * n), vel(3 * n), acc(3 * n);
    NB_TRY(nbody_read_soa(s, (fields & NBODY_FIELD_POSITION) ? pos.data() : nullptr,
                          (fields & NBODY_FIELD_VELOCITY) ? vel.data() : nullptr,
                          (fields & NBODY_FIELD_FORCES) ? acc.data() : nullptr, n));
    for (size_t i = 0; i < n; ++i) {
        if (fields & NBODY_FIELD_POSITION)
            for (int c = 0; c < 3; ++c) p[i].position[c] = pos[3 * i + c];
        if (fields & NBODY_FIELD_VELOCITY)
            for (int c = 0; c < 3; ++c) p[i].velocity[c] = vel[3 * i + c];
        if (fields & NBODY_FIELD_FORCES)
            for (int c = 0; c < 3; ++c) p[i].forces[c] = s->mass_host[i] * static_cast<double>(acc[3 * i + c]);
        if (fields & NBODY_FIELD_MASS) p[i].mass = s->mass_host[i];
    }
    return NBODY_OK;
}

nbody_status nbody_compute_accelerations(nbody_sim* s, float* acc, size_t n) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    NB_CHECK(s->initialised, NBODY_ERR_NOT_INITIALISED, "nbody_compute_accelerations before nbody_init");
    NB_CHECK(acc || n == 0, NBODY_ERR_INVALID_ARG, "acc is null");
    NB_CHECK(n == s->n, NBODY_ERR_INVALID_ARG, "n = %zu does not match the state's %zu particles", n, s->n);
    if (n == 0) return NBODY_OK;
    NB_HIP(hipSetDevice(s->cfg.device));
    s->step_interactions = 0;
    if (s->bh) NB_TRY(bh_reset_counters(s->bh, s->stream));  // the stats then describe this evaluation
    NB_TRY(force_pass(s, 0.f, false, nullptr, nullptr));
    return nbody_read_soa(s, nullptr, nullptr, acc, n);
}

nbody_status nbody_field(nbody_sim* s, const float* points, size_t m, float* acc) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    NB_CHECK(s->initialised, NBODY_ERR_NOT_INITIALISED, "nbody_field before nbody_init");
    NB_CHECK((points && acc) || m == 0, NBODY_ERR_INVALID_ARG, "null argument");
    NB_CHECK(m <= (size_t(1) << 30), NBODY_ERR_CAPACITY, "too many points");
    if (m == 0) return NBODY_OK;
    NB_HIP(hipSetDevice(s->cfg.device));
    std::vector<float4> h(m);
    for (size_t i = 0; i < m; ++i) h[i] = make_float4(points[3 * i], points[3 * i + 1], points[3 * i + 2], 0.f);
    float4* d = nullptr;
    NB_HIP(hipMalloc(reinterpret_cast<void**>(&d), 2 * m * sizeof(float4)));
    nbody_status st = NBODY_OK;
    hipError_t e = hipMemcpyAsync(d, h.data(), m * sizeof(float4), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = direct_field(s->pos[s->cur], static_cast<int>(s->n_pad), d, static_cast<int>(m), s->cfg.law, s->dscale,
                         d + m, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d + m, m * sizeof(float4), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        set_error("nbody_field: %s", hipGetErrorString(e));
        st = NBODY_ERR_DEVICE;
    }
    (void)hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    if (st != NBODY_OK) return st;
    for (size_t i = 0; i < m; ++i) {
        acc[3 * i] = h[i].x;
        acc[3 * i + 1] = h[i].y;
        acc[3 * i + 2] = h[i].z;
    }
    return NBODY_OK;
}

nbody_status nbody_field_noalloc(nbody_sim* s, const float* points, size_t m, float* acc) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    NB_CHECK(s->initialised, NBODY_ERR_NOT_INITIALISED, "nbody_field_noalloc before nbody_init");
    NB_CHECK((points && acc) || m == 0, NBODY_ERR_INVALID_ARG, "null argument");
    NB_CHECK(m <= NBODY_FIELD_NOALLOC_MAX_POINTS, NBODY_ERR_CAPACITY,
             "nbody_field_noalloc takes at most %d points (%zu given); use nbody_field", NBODY_FIELD_NOALLOC_MAX_POINTS,
             m);
    if (m == 0) return NBODY_OK;
    NB_HIP(hipSetDevice(s->cfg.device));
    float4* d = nullptr;
    NB_TRY(field_scratch(s->cfg.device, &d));
    const int cap = NBODY_FIELD_NOALLOC_MAX_POINTS;
    // the caller's packed points into the second half; direct_field_staged
    // packs, evaluates and unpacks them inside the array
    NB_HIP(hipMemcpyAsync(reinterpret_cast<float*>(d + cap), points, 3 * m * sizeof(float), hipMemcpyHostToDevice,
                          s->stream));
    NB_HIP(direct_field_staged(s->pos[s->cur], static_cast<int>(s->n_pad), d, cap, static_cast<int>(m), s->cfg.law,
                               s->dscale, s->stream));
    NB_HIP(hipMemcpyAsync(acc, reinterpret_cast<const float*>(d), 3 * m * sizeof(float), hipMemcpyDeviceToHost,
                          s->stream));
    NB_HIP(hipStreamSynchronize(s->stream));
    return NBODY_OK;
}

nbody_status nbody_size(nbody_sim* s, size_t* n) {
    clear_error();
    NB_CHECK(s && n, NBODY_ERR_INVALID_ARG, "null argument");
    *n = s->n;
    return NBODY_OK;
}

nbody_status nbody_diagnostics(nbody_sim* s, nbody_diag* out) {
    clear_error();
    NB_CHECK(s && out, NBODY_ERR_INVALID_ARG, "null argument");
    NB_CHECK(s->initialised, NBODY_ERR_NOT_INITIALISED, "nbody_diagnostics before nbody_init");
    std::memset(out, 0, sizeof *out);
    if (s->n == 0) return NBODY_OK;
    NB_HIP(hipSetDevice(s->cfg.device));
    const int n = static_cast<int>(s->n);
    const size_t lb = static_cast<size_t>(diag_local_blocks(n)), pb = static_cast<size_t>(diag_potential_blocks(n));
    const size_t need = 8 * lb + pb;
    if (need > s->diag_elems) {
        NB_TRY(dev_alloc(&s->diag_buf, need));
        s->diag_elems = need;
    }
    NB_HIP(diag_launch(s->pos[s->cur], s->vel, s->mass_dev, s->gm_dev, n, s->cfg.law, s->cfg.softening,
                       s->diag_buf, s->diag_buf + 8 * lb, s->stream));
    std::vector<double> h(need);
    NB_HIP(hipMemcpyAsync(h.data(), s->diag_buf, need * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    NB_HIP(hipStreamSynchronize(s->stream));
    double q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t b = 0; b < lb; ++b)
        for (int c = 0; c < 8; ++c) q[c] += h[8 * b + c];
    double pot = 0;
    for (size_t b = 0; b < pb; ++b) pot += h[8 * lb + b];
    out->kinetic = q[0];
    out->momentum[0] = q[1];
    out->momentum[1] = q[2];
    out->momentum[2] = q[3];
    out->angular_momentum[0] = q[4];
    out->angular_momentum[1] = q[5];
    out->angular_momentum[2] = q[6];
    out->mass = q[7];
    // Positions advance by v*dt/scale (x' = v/scale, v' = a), so the conserved
    // energy is sum m v^2 / (2 scale) + U; momenta stay in velocity units.
    out->kinetic /= s->cfg.scale;
    out->potential = pot;
    out->total = out->kinetic + out->potential;
    return NBODY_OK;
}

nbody_status nbody_set_graph(nbody_sim* s, int32_t enable) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    NB_HIP(hipSetDevice(s->cfg.device));
    NB_HIP(hipStreamSynchronize(s->stream));
    s->use_graph = enable != 0;
    if (!s->use_graph) drop_graphs(s);
    return NBODY_OK;
}

nbody_status nbody_set_theta(nbody_sim* s, double theta) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    NB_CHECK(std::isfinite(theta) && theta >= 0, NBODY_ERR_INVALID_ARG, "theta must be finite and >= 0");
    NB_HIP(hipSetDevice(s->cfg.device));
    NB_HIP(hipStreamSynchronize(s->stream));
    // A captured Barnes-Hut step holds theta (the dynamic root's kernel
    // argument), so the graphs are captured again on the next step.
    drop_graphs(s);
    s->cfg.theta = theta;
    if (s->bh) NB_TRY(bh_set_theta(s));
    return NBODY_OK;
}

nbody_status nbody_set_timing(nbody_sim* s, int32_t enable) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    s->timing = enable != 0;
    return NBODY_OK;
}

nbody_status nbody_get_timing(nbody_sim* s, nbody_timing* out) {
    clear_error();
    NB_CHECK(s && out, NBODY_ERR_INVALID_ARG, "null argument");
    NB_HIP(hipSetDevice(s->cfg.device));
    NB_HIP(hipStreamSynchronize(s->stream));
    *out = s->last;
    if (s->bh && s->n >= 2) {
        BHCounters c{};
        NB_TRY(bh_read_counters(s->bh, s->stream, &c));
        out->interactions = c.interactions;
    }
    if (s->timed_steps > 0) {
        float ms = 0;
        NB_HIP(hipEventElapsedTime(&ms, s->events[0], s->events[1]));
        out->step_ms = ms;
        double f = 0;
        for (int k = 0; k < s->timed_steps; ++k) {
            NB_HIP(hipEventElapsedTime(&ms, s->events[2 + 2 * k], s->events[3 + 2 * k]));
            f += ms;
        }
        out->force_ms = f;
        out->tree_ms = 0;
        for (size_t k = 0; k < s->tree_ms_pending.size(); k += 2) {
            NB_HIP(hipEventElapsedTime(&ms, s->tree_ms_pending[k], s->tree_ms_pending[k + 1]));
            out->tree_ms += ms;
        }
    }
    return NBODY_OK;
}

nbody_status nbody_bh_stats_get(nbody_sim* s, nbody_bh_stats* out) {
    clear_error();
    NB_CHECK(s && out, NBODY_ERR_INVALID_ARG, "null argument");
    NB_CHECK(s->bh, NBODY_ERR_UNSUPPORTED, "not a Barnes-Hut simulation");
    std::memset(out, 0, sizeof *out);
    NB_HIP(hipSetDevice(s->cfg.device));
    BHCounters c{};
    NB_TRY(bh_read_counters(s->bh, s->stream, &c));
    out->interactions = c.interactions;
    out->outside_root = c.outside;
    out->stack_overflow = c.overflow;
    out->list_entries = c.entries;
    out->cells_opened = c.opened;
    return NBODY_OK;
}

nbody_status nbody_bh_tree_order(nbody_sim* s, uint64_t* keys, uint32_t* order, size_t n) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "sim is null");
    NB_CHECK(s->bh, NBODY_ERR_UNSUPPORTED, "not a Barnes-Hut simulation");
    NB_CHECK(s->initialised && n == s->n, NBODY_ERR_INVALID_ARG, "n = %zu does not match the state's %zu", n, s->n);
    NB_HIP(hipSetDevice(s->cfg.device));
    return bh_export_order(s, keys, order, n);
}

nbody_status nbody_bh_cells(nbody_sim* s, nbody_bh_cell* out, size_t capacity, size_t* count) {
    clear_error();
    NB_CHECK(s && count && (out || capacity == 0), NBODY_ERR_INVALID_ARG, "bad argument");
    NB_CHECK(s->bh, NBODY_ERR_UNSUPPORTED, "not a Barnes-Hut simulation");
    NB_CHECK(s->initialised, NBODY_ERR_NOT_INITIALISED, "no tree before nbody_init");
    NB_HIP(hipSetDevice(s->cfg.device));
    return bh_export_cells(s, out, capacity, count);
}

const char* nbody_method_name(int32_t method) {
    switch (method) {
        case NBODY_METHOD_DIRECT: return "Brute Force GPU";
        case NBODY_METHOD_BARNES_HUT: return "Barnes-Hut";
        default: return "Unknown";
    }
}

const char* nbody_status_string(nbody_status st) {
    switch (st) {
        case NBODY_OK: return "ok";
        case NBODY_ERR_INVALID_ARG: return "invalid argument";
        case NBODY_ERR_OUT_OF_MEMORY: return "out of memory";
        case NBODY_ERR_DEVICE: return "device error";
        case NBODY_ERR_NOT_INITIALISED: return "not initialised";
        case NBODY_ERR_UNSUPPORTED: return "unsupported";
        case NBODY_ERR_CAPACITY: return "capacity exceeded";
        case NBODY_ERR_IO: return "i/o error";
        default: return "unknown status";
    }
}

const char* nbody_last_error(void) { return last_error(); }

int32_t nbody_abi_version(void) { return NBODY_ABI_VERSION; }

int32_t nbody_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

nbody_status nbody_ic_plummer(uint64_t seed, size_t n, double a, double total_mass, double G, float* pos,
                              float* vel, double* mass) {
    clear_error();
    NB_CHECK((pos && vel && mass) || n == 0, NBODY_ERR_INVALID_ARG, "null output");
    NB_CHECK(a > 0 && total_mass > 0 && G >= 0, NBODY_ERR_INVALID_ARG, "bad Plummer parameters");
    ic_plummer(seed, n, a, total_mass, G, pos, vel, mass);
    ic_recentre(n, pos, vel, mass);
    return NBODY_OK;
}

nbody_status nbody_ic_disc_bulge(uint64_t seed, size_t n, double Rd, double z0, double bulge_fraction, double ab,
                                 double total_mass, double G, float* pos, float* vel, double* mass) {
    clear_error();
    NB_CHECK((pos && vel && mass) || n == 0, NBODY_ERR_INVALID_ARG, "null output");
    NB_CHECK(Rd > 0 && z0 > 0 && bulge_fraction >= 0 && bulge_fraction <= 1 && ab > 0 && total_mass > 0 && G >= 0,
             NBODY_ERR_INVALID_ARG, "bad disc/bulge parameters");
    ic_disc_bulge(seed, n, Rd, z0, bulge_fraction, ab, total_mass, G, pos, vel, mass);
    return NBODY_OK;
}

nbody_status nbody_ic_disc_encounter(uint64_t seed, size_t n, double Rd, double z0, double separation,
                                     double pericentre, double tilt_deg, double G, float* pos, float* vel,
                                     double* mass) {
    clear_error();
    NB_CHECK((pos && vel && mass) || n == 0, NBODY_ERR_INVALID_ARG, "null output");
    NB_CHECK(n >= 2 && Rd > 0 && z0 > 0 && separation > 0 && pericentre > 0 && pericentre <= separation && G > 0,
             NBODY_ERR_INVALID_ARG, "bad encounter parameters");
    ic_disc_encounter(seed, n, Rd, z0, separation, pericentre, tilt_deg, G, pos, vel, mass);
    return NBODY_OK;
}

nbody_status nbody_ic_uniform(uint64_t seed, size_t n, double half, double mass_lo, double mass_hi, float* pos,
                              float* vel, double* mass) {
    clear_error();
    NB_CHECK((pos && vel && mass) || n == 0, NBODY_ERR_INVALID_ARG, "null output");
    NB_CHECK(half > 0 && mass_lo >= 0 && mass_hi >= mass_lo, NBODY_ERR_INVALID_ARG, "bad uniform parameters");
    ic_uniform(seed, n, half, mass_lo, mass_hi, pos, vel, mass);
    return NBODY_OK;
}

}  // extern "C"
