// Request loop for galaxy simulations: continuous batching on one GPU.
//
// The reference runs one simulation at a time from its UI or its headless
// precompute loop (SimulationState::RunSimulation, SimulationState.cpp:279-332).
// Here requests (a galaxy's size, seed and step count) arrive at any time and
// are served together:
//
//  * admission  - queued requests are admitted in arrival order when their
//                 device-memory footprint fits the budget (first fit: a small
//                 request may pass a large one that does not fit yet); a
//                 request's state is generated on the device and stays there;
//  * direct     - requests below the size threshold join the batch arena and
//                 advance in its persistent hipGraph launch, one batch step
//                 at a time;
//  * Barnes-Hut - larger requests get a simulation handle of their own (from
//                 a pool of idle handles), on its own stream, and a few steps
//                 are enqueued at a time, so their steps overlap the batch's
//                 and each other's on the device;
//  * retirement - a request that has taken its steps leaves the batch (or
//                 returns its handle), its memory is released and its
//                 completion time recorded.
//
// Nothing waits on the device: a tick credits the work that has completed
// (stream queries), retires, admits, and enqueues more, so the batch steps at
// its own pace and every Barnes-Hut request at its own.  Every request's
// arithmetic is independent of the others (batch items and handles never mix
// systems), so a result is the same alone and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "sim.hpp"

namespace nbody {
namespace {

// Device bytes per particle: a simulation handle (positions x2, velocity,
// acceleration, FP64 mass and G m: 80 B) plus the Barnes-Hut tree (keys and
// indices x2: 24 B, sorted positions: 16 B, radix-tree node arrays: 20 B,
// child blocks: 192 B, double-double prefix sums: 64 B, cell records: 16 B,
// digit counts: under 1 B): 413 B, counted as 432 (an upper bound).
constexpr size_t kBhBytesPerParticle = 80 + 352;
constexpr size_t kDirectBytesPerParticle = 64;

struct Request {
    uint64_t id = 0;
    size_t n = 0;
    int32_t steps = 0;
    int32_t done = 0;
    uint64_t seed = 0;
    double a = 1.0, mass = 1.0;
    double t_arrival = 0, t_admit = 0, t_done = 0;
    int32_t method = NBODY_METHOD_DIRECT;
    int32_t batch_id = -1;
    int32_t inflight = 0;         // Barnes-Hut steps enqueued and not yet credited
    bool in_batch_step = false;   // member of the batch step in flight
    nbody_sim* sim = nullptr;
    size_t sim_cap = 0;
    size_t footprint = 0;
    uint64_t interactions = 0;
};

}  // namespace
}  // namespace nbody

using namespace nbody;

struct nbody_server {
    nbody_server_config cfg{};
    std::chrono::steady_clock::time_point t0;
    nbody_batch* batch = nullptr;
    std::deque<Request> pending;
    std::vector<Request> active;
    std::vector<nbody_request_result> done;
    struct Idle {
        nbody_sim* sim;
        size_t cap;
    };
    std::vector<Idle> idle;
    std::map<uint64_t, std::vector<float>> kept;  // id -> pos[3n] then vel[3n]
    uint64_t next_id = 1;
    bool batch_inflight = false;
    int32_t inject_left = 0;  // injected allocation failures still to come
    size_t mem_used = 0, mem_peak = 0;
    nbody_server_stats stats{};

    double clock() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace nbody {
namespace {

size_t footprint(const nbody_server* s, size_t n, int method) {
    const size_t np = round_up(n, kTile);
    return method == NBODY_METHOD_DIRECT ? np * kDirectBytesPerParticle : np * kBhBytesPerParticle;
}

void release_idle(nbody_server* s, size_t need) {
    // Free idle Barnes-Hut handles (largest first) until `need` bytes fit.
    std::sort(s->idle.begin(), s->idle.end(), [](const nbody_server::Idle& a, const nbody_server::Idle& b) {
        return a.cap > b.cap;
    });
    while (!s->idle.empty() && s->mem_used + need > s->cfg.memory_budget) {
        nbody_destroy(s->idle.front().sim);
        s->mem_used -= footprint(s, s->idle.front().cap, NBODY_METHOD_BARNES_HUT);
        s->idle.erase(s->idle.begin());
    }
}

// A device allocation for a request, or the injected failure that stands in
// for one (nbody_server_config.inject_alloc_failures).
template <class F>
nbody_status allocate(nbody_server* s, F&& call) {
    if (s->inject_left > 0) {
        --s->inject_left;
        set_error("injected allocation failure");
        return NBODY_ERR_OUT_OF_MEMORY;
    }
    return call();
}

// The device is fuller than the footprint model says: give back the idle
// handles' memory and leave the request queued for a later tick.
nbody_status defer(nbody_server* s) {
    release_idle(s, s->cfg.memory_budget + 1);
    s->stats.alloc_deferrals += 1;
    clear_error();
    return NBODY_OK;
}

nbody_status admit(nbody_server* s, Request& r, bool* ok) {
    *ok = false;
    if (r.method == NBODY_METHOD_DIRECT) {
        const size_t f = footprint(s, r.n, r.method);
        if (s->mem_used + f > s->cfg.memory_budget) release_idle(s, f);
        if (s->mem_used + f > s->cfg.memory_budget) return NBODY_OK;
        int32_t id = -1;
        nbody_status st =
            allocate(s, [&] { return nbody_batch_add_plummer(s->batch, r.seed, r.n, r.a, r.mass, &id); });
        if (st == NBODY_ERR_CAPACITY) return NBODY_OK;  // arena fragmented or full: wait
        if (st == NBODY_ERR_OUT_OF_MEMORY) return defer(s);
        NB_TRY(st);
        r.batch_id = id;
        r.footprint = f;
        s->mem_used += f;
    } else {
        if (static_cast<int32_t>(std::count_if(s->active.begin(), s->active.end(), [](const Request& q) {
                return q.method == NBODY_METHOD_BARNES_HUT;
            })) >= s->cfg.max_active_bh)
            return NBODY_OK;
        // Reuse the smallest idle handle that holds n particles.
        int pick = -1;
        for (size_t k = 0; k < s->idle.size(); ++k)
            if (s->idle[k].cap >= r.n && (pick < 0 || s->idle[k].cap < s->idle[pick].cap)) pick = static_cast<int>(k);
        if (pick >= 0) {
            r.sim = s->idle[pick].sim;
            r.sim_cap = s->idle[pick].cap;
            s->idle.erase(s->idle.begin() + pick);
        } else {
            const size_t f = footprint(s, r.n, r.method);
            if (s->mem_used + f > s->cfg.memory_budget) release_idle(s, f);
            if (s->mem_used + f > s->cfg.memory_budget) return NBODY_OK;
            nbody_config c = s->cfg.sim;
            c.method = NBODY_METHOD_BARNES_HUT;
            const nbody_status sc = allocate(s, [&] { return nbody_create(&c, &r.sim); });
            if (sc == NBODY_ERR_OUT_OF_MEMORY) return defer(s);
            NB_TRY(sc);
            // one graph launch per step instead of about thirty kernels
            const nbody_status sg = nbody_set_graph(r.sim, s->cfg.bh_no_graph ? 0 : 1);
            if (sg != NBODY_OK) {
                const std::string msg = nbody_last_error();  // nbody_destroy clears it
                nbody_destroy(r.sim);
                r.sim = nullptr;
                set_error("%s", msg.c_str());
                return sg;
            }
            r.sim_cap = r.n;
            s->mem_used += f;
        }
        r.footprint = footprint(s, r.sim_cap, r.method);
        // a new handle allocates its state here
        nbody_status st = allocate(s, [&] { return nbody_init_plummer(r.sim, r.seed, r.n, r.a, r.mass); });
        if (st != NBODY_OK) {
            const std::string msg = nbody_last_error();  // nbody_destroy clears it
            nbody_destroy(r.sim);
            s->mem_used -= r.footprint;
            r.sim = nullptr;
            if (st == NBODY_ERR_OUT_OF_MEMORY) return defer(s);
            set_error("%s", msg.c_str());
            return st;
        }
    }
    s->mem_peak = std::max(s->mem_peak, s->mem_used);
    r.t_admit = s->clock();
    *ok = true;
    return NBODY_OK;
}

nbody_status retire(nbody_server* s, Request& r) {
    if (s->cfg.keep_results) {
        std::vector<float> buf(6 * r.n);
        if (r.method == NBODY_METHOD_DIRECT)
            NB_TRY(nbody_batch_read(s->batch, r.batch_id, buf.data(), buf.data() + 3 * r.n, nullptr, r.n));
        else
            NB_TRY(nbody_read_soa(r.sim, buf.data(), buf.data() + 3 * r.n, nullptr, r.n));
        s->kept[r.id] = std::move(buf);
    }
    if (r.method == NBODY_METHOD_DIRECT) {
        NB_TRY(nbody_batch_remove(s->batch, r.batch_id));
        s->mem_used -= r.footprint;
    } else {
        s->idle.push_back(nbody_server::Idle{r.sim, r.sim_cap});  // memory stays counted while idle
        r.sim = nullptr;
    }
    r.t_done = s->clock();
    nbody_request_result out{};
    out.id = r.id;
    out.n = static_cast<uint32_t>(r.n);
    out.steps = r.steps;
    out.method = r.method;
    out.arrival = r.t_arrival;
    out.admitted = r.t_admit;
    out.completed = r.t_done;
    out.interactions = r.interactions;
    s->done.push_back(out);
    s->stats.completed += 1;
    s->stats.interactions += r.interactions;
    return NBODY_OK;
}

}  // namespace
}  // namespace nbody

extern "C" {

void nbody_server_config_default(nbody_server_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    nbody_config_default(&c->sim);
    c->sim.law = NBODY_LAW_R;
    c->sim.softening = 0.02 * 0.02;
    c->sim.theta = 0.6;
    c->memory_budget = 64ull << 30;
    c->direct_capacity = 64ull << 20;
    c->bh_threshold = 32768;
    c->max_active_bh = 16;
    c->dt = 1e-3f;
    c->keep_results = 0;
    c->bh_steps_ahead = 1;
}

nbody_status nbody_server_create(const nbody_server_config* cfg, nbody_server** out) {
    clear_error();
    NB_CHECK(cfg && out, NBODY_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    nbody_config c = cfg->sim;
    c.method = NBODY_METHOD_DIRECT;
    NB_TRY(validate_config(&c));
    NB_CHECK(cfg->memory_budget > 0 && cfg->direct_capacity > 0 && cfg->max_active_bh >= 0 && std::isfinite(cfg->dt),
             NBODY_ERR_INVALID_ARG, "bad server configuration");
    nbody_server* s = new (std::nothrow) nbody_server();
    NB_CHECK(s, NBODY_ERR_OUT_OF_MEMORY, "host allocation failed");
    s->cfg = *cfg;
    s->inject_left = std::max(0, cfg->inject_alloc_failures);
    s->t0 = std::chrono::steady_clock::now();
    nbody_status st = nbody_batch_create(&c, cfg->direct_capacity, &s->batch);
    if (st != NBODY_OK) {
        delete s;
        return st;
    }
    *out = s;
    return NBODY_OK;
}

nbody_status nbody_server_destroy(nbody_server* s) {
    clear_error();
    if (!s) return NBODY_OK;
    for (auto& r : s->active)
        if (r.sim) nbody_destroy(r.sim);
    for (auto& i : s->idle) nbody_destroy(i.sim);
    nbody_batch_destroy(s->batch);
    delete s;
    return NBODY_OK;
}

double nbody_server_clock(nbody_server* s) { return s ? s->clock() : 0.0; }

nbody_status nbody_server_submit_plummer(nbody_server* s, uint64_t seed, size_t n, int32_t steps, double a,
                                         double total_mass, double arrival, uint64_t* id) {
    clear_error();
    NB_CHECK(s && id, NBODY_ERR_INVALID_ARG, "null argument");
    NB_CHECK(n >= 1 && n <= (size_t(1) << 28) && steps >= 0 && a > 0 && total_mass > 0, NBODY_ERR_INVALID_ARG,
             "bad request (n %zu, steps %d)", n, steps);
    Request r;
    r.id = s->next_id++;
    r.n = n;
    r.steps = steps;
    r.seed = seed;
    r.a = a;
    r.mass = total_mass;
    r.t_arrival = std::isfinite(arrival) ? arrival : s->clock();
    r.method = n >= s->cfg.bh_threshold ? NBODY_METHOD_BARNES_HUT : NBODY_METHOD_DIRECT;
    s->pending.push_back(r);
    s->stats.submitted += 1;
    *id = r.id;
    return NBODY_OK;
}

nbody_status nbody_server_tick(nbody_server* s, int32_t* completed) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "server is null");
    if (completed) *completed = 0;
    const uint64_t done_before = s->stats.completed;
    const double t_start = s->clock();
    bool progressed = false;
    // 1. credit completed work
    if (s->batch_inflight) {
        int32_t idle = 0;
        NB_TRY(nbody_batch_query(s->batch, &idle));
        if (idle) {
            for (auto& r : s->active)
                if (r.in_batch_step) {
                    r.in_batch_step = false;
                    r.done += 1;
                    r.interactions += static_cast<uint64_t>(r.n) * r.n;
                }
            s->batch_inflight = false;
            progressed = true;
        }
    }
    for (auto& r : s->active) {
        if (r.method != NBODY_METHOD_BARNES_HUT || r.inflight == 0) continue;
        int32_t idle = 0;
        NB_TRY(nbody_query(r.sim, &idle));
        if (!idle) continue;
        nbody_timing t;
        NB_TRY(nbody_get_timing(r.sim, &t));
        r.interactions += t.interactions;
        r.done += r.inflight;
        r.inflight = 0;
        progressed = true;
    }
    // 2. retirement (a direct request only while the batch is idle: leaving
    // the arena changes its items)
    for (size_t k = 0; k < s->active.size();) {
        const Request& r = s->active[k];
        const bool finished = r.done >= r.steps && r.inflight == 0 && !r.in_batch_step &&
                              (r.method != NBODY_METHOD_DIRECT || !s->batch_inflight);
        if (finished) {
            NB_TRY(retire(s, s->active[k]));
            s->active.erase(s->active.begin() + k);
            progressed = true;
        } else {
            ++k;
        }
    }
    // 3. admission, first fit in arrival order (a direct request added while
    // a batch step runs joins from the next step on)
    for (auto it = s->pending.begin(); it != s->pending.end();) {
        bool ok = false;
        NB_TRY(admit(s, *it, &ok));
        if (ok) {
            s->active.push_back(*it);
            s->stats.admitted += 1;
            it = s->pending.erase(it);
            progressed = true;
        } else {
            ++it;
        }
    }
    // Requests of zero steps complete on admission.
    for (size_t k = 0; k < s->active.size();) {
        const Request& r = s->active[k];
        if (r.steps == 0 && (r.method != NBODY_METHOD_DIRECT || !s->batch_inflight)) {
            NB_TRY(retire(s, s->active[k]));
            s->active.erase(s->active.begin() + k);
        } else {
            ++k;
        }
    }
    // 4. enqueue: the next batch step, and the next steps of idle handles
    if (!s->batch_inflight) {
        bool any = false;
        for (auto& r : s->active)
            if (r.method == NBODY_METHOD_DIRECT && r.done < r.steps) {
                r.in_batch_step = true;
                any = true;
            }
        if (any) {
            NB_TRY(nbody_batch_step(s->batch, s->cfg.dt, 1));
            s->batch_inflight = true;
            progressed = true;
        }
    }
    const int32_t ahead = s->cfg.bh_steps_ahead > 0 ? s->cfg.bh_steps_ahead : 1;
    for (auto& r : s->active)
        if (r.method == NBODY_METHOD_BARNES_HUT && r.inflight == 0 && r.done < r.steps) {
            const int32_t k = std::min(ahead, r.steps - r.done);
            NB_TRY(nbody_step(r.sim, s->cfg.dt, k));
            r.inflight = k;
            progressed = true;
        }
    if (progressed) {
        s->stats.ticks += 1;
        s->stats.busy_seconds += s->clock() - t_start;
    } else {
        std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    if (completed) *completed = static_cast<int32_t>(s->stats.completed - done_before);
    return NBODY_OK;
}

nbody_status nbody_server_poll(nbody_server* s, nbody_request_result* out, int32_t max, int32_t* count) {
    clear_error();
    NB_CHECK(s && count && (out || max == 0) && max >= 0, NBODY_ERR_INVALID_ARG, "bad argument");
    const int32_t k = std::min<int32_t>(max, static_cast<int32_t>(s->done.size()));
    for (int32_t i = 0; i < k; ++i) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + k);
    *count = k;
    return NBODY_OK;
}

nbody_status nbody_server_fetch(nbody_server* s, uint64_t id, float* pos, float* vel, size_t n) {
    clear_error();
    NB_CHECK(s, NBODY_ERR_INVALID_ARG, "server is null");
    auto it = s->kept.find(id);
    NB_CHECK(it != s->kept.end(), NBODY_ERR_INVALID_ARG, "no kept result for request %llu",
             static_cast<unsigned long long>(id));
    NB_CHECK(it->second.size() == 6 * n, NBODY_ERR_INVALID_ARG, "request %llu has %zu particles",
             static_cast<unsigned long long>(id), it->second.size() / 6);
    if (pos) std::memcpy(pos, it->second.data(), 3 * n * sizeof(float));
    if (vel) std::memcpy(vel, it->second.data() + 3 * n, 3 * n * sizeof(float));
    s->kept.erase(it);
    return NBODY_OK;
}

nbody_status nbody_server_get_stats(nbody_server* s, nbody_server_stats* out) {
    clear_error();
    NB_CHECK(s && out, NBODY_ERR_INVALID_ARG, "null argument");
    *out = s->stats;
    out->pending = s->pending.size();
    out->active_direct = 0;
    out->active_bh = 0;
    for (const auto& r : s->active) (r.method == NBODY_METHOD_DIRECT ? out->active_direct : out->active_bh) += 1;
    out->memory_used = s->mem_used;
    out->memory_peak = s->mem_peak;
    out->idle_handles = s->idle.size();
    return NBODY_OK;
}

}  // extern "C"
