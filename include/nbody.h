/*
 * nbody.h - C ABI of the MI355X (gfx950) gravitational N-body step.
 *
 * This library is a drop-in for the simulator behind the reference's
 * `INBodySim` interface (src/Sim/INBodySim.hpp:11-28):
 *
 *     class INBodySim {
 *         virtual void Init(std::vector<Particle>& particles) = 0;
 *         virtual void Update(float dt) = 0;
 *         virtual void RenderDebug(Matrix view, Matrix proj) {}
 *     };
 *     std::unique_ptr<INBodySim> CreateNBodySim(ID3D11DeviceContext*, ENBodySim);
 *     std::string NBodySimGetName(ENBodySim);
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns
 * an `nbody_status` and never throws.  A failed call leaves a message that
 * `nbody_last_error()` returns (per calling thread).  One simulation handle
 * must not be used from two threads at once; distinct handles are
 * independent.  All device work of a handle runs on the handle's own HIP
 * stream; `nbody_step` only enqueues, every read synchronises.
 *
 * The header-only C++ adapter `nbody_sim.hpp` re-exposes the reference's
 * Init/Update/read-in-place contract on top of these calls.
 */
#ifndef NBODY_H_
#define NBODY_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBODY_ABI_VERSION 2

/* ------------------------------------------------------------------------ */
/* Particle record                                                           */
/* ------------------------------------------------------------------------ */

/*
 * Byte-compatible with the reference's `struct Particle`
 * (src/Render/Misc/Particle.hpp:8-18) as laid out by a 64-bit compiler:
 * Vector3 Position @0, Color Colour @12, Color OriginalColour @28,
 * 4 bytes of padding @44, Vec3d Velocity @48, Vec3d Forces @72,
 * double Mass @96; 104 bytes in all.  It is also the record of the
 * reference's `.nbody` snapshot file (SimulationState.cpp:317-331).
 */
typedef struct nbody_particle {
    float  position[3];
    float  colour[4];
    float  original_colour[4];
    float  pad_;
    double velocity[3];
    double forces[3];
    double mass;
} nbody_particle;

/* ------------------------------------------------------------------------ */
/* Status codes                                                              */
/* ------------------------------------------------------------------------ */

typedef enum nbody_status {
    NBODY_OK = 0,
    NBODY_ERR_INVALID_ARG = 1,     /* null pointer, bad size, bad enum value     */
    NBODY_ERR_OUT_OF_MEMORY = 2,   /* device or host allocation failed            */
    NBODY_ERR_DEVICE = 3,          /* a HIP call failed                           */
    NBODY_ERR_NOT_INITIALISED = 4, /* step/read before init                       */
    NBODY_ERR_UNSUPPORTED = 5,     /* valid request the library does not serve    */
    NBODY_ERR_CAPACITY = 6,        /* request exceeds a fixed capacity            */
    NBODY_ERR_IO = 7               /* snapshot file could not be read or written  */
} nbody_status;

/* ------------------------------------------------------------------------ */
/* Simulation parameters                                                     */
/* ------------------------------------------------------------------------ */

/* Replaces `enum class ENBodySim` (src/Sim/INBodySim.hpp:11-17): both of the
 * reference's direct-sum simulators (BruteForceCPU, BruteForceGPU) map to
 * NBODY_METHOD_DIRECT; BarnesHut maps to NBODY_METHOD_BARNES_HUT. */
typedef enum nbody_method {
    NBODY_METHOD_DIRECT = 1,
    NBODY_METHOD_BARNES_HUT = 2
} nbody_method;

/* Pairwise force law.
 * NBODY_LAW_R is the reference's law (src/Sim/Physics.hpp:25-35 with
 * src/Sim/BruteForceCPU.cpp:36-40):  a_i = sum_j G m_j rhat_ij / (r^2 + S),
 * where `softening` is S and rhat is the unit vector from i to j.
 * NBODY_LAW_P is Plummer: a_i = sum_j G m_j r_ij / (r^2 + eps^2)^(3/2), where
 * `softening` is eps^2 (must be > 0). */
typedef enum nbody_law {
    NBODY_LAW_R = 0,
    NBODY_LAW_P = 1
} nbody_law;

/* Barnes-Hut root cell. FIXED uses root_lo/root_size as the reference's
 * fixed cube (src/Sim/BarnesHut.cpp:14-19); DYNAMIC takes the bounding box of
 * the particles at every step and rounds it out to a power-of-two cube. */
typedef enum nbody_root_mode {
    NBODY_ROOT_DYNAMIC = 0,
    NBODY_ROOT_FIXED = 1
} nbody_root_mode;

typedef struct nbody_config {
    int32_t method;      /* nbody_method                                        */
    int32_t law;         /* nbody_law                                           */
    double  G;           /* gravitational constant                              */
    double  softening;   /* S (law R) or eps^2 (law P)                          */
    double  scale;       /* drift divisor: x += v*dt/scale (Physics.hpp:16)     */
    double  theta;       /* Barnes-Hut opening angle (Octree.cpp:5)             */
    int32_t root_mode;   /* nbody_root_mode                                     */
    float   root_lo[3];  /* fixed root: lower corner                            */
    float   root_size;   /* fixed root: edge length                             */
    int32_t device;      /* HIP device ordinal                                  */
    int32_t tune_i_per_lane; /* direct kernel: targets per lane (0 = auto)      */
    int32_t tune_block;      /* direct kernel: threads per workgroup (0 = auto) */
    int32_t tune_split;      /* direct kernel: j-chunks per target (0 = auto)   */
    int32_t reserved[8];
} nbody_config;

/* G = 1, law P, eps^2 = 1e-4, scale = 1, theta = 0.5, dynamic root,
 * device 0, autotuned launch shape. */
void nbody_config_default(nbody_config* cfg);

/* The reference's constants (src/Sim/Physics.hpp:9-16, Octree.cpp:5,
 * BarnesHut.cpp:14-19): law R, G = 6.674e-11, S = 10, scale = 2.3e13,
 * theta = 2.0, fixed root [-4000, 4000)^3. */
void nbody_config_reference(nbody_config* cfg, int32_t method);

/* ------------------------------------------------------------------------ */
/* Simulation handle (INBodySim)                                             */
/* ------------------------------------------------------------------------ */

typedef struct nbody_sim nbody_sim;

/* CreateNBodySim (src/Sim/INBodySim.cpp:7-27). */
nbody_status nbody_create(const nbody_config* cfg, nbody_sim** out);
nbody_status nbody_destroy(nbody_sim* sim);

/* INBodySim::Init (src/Sim/INBodySim.hpp:21): copies n particles to the
 * device (no host pointer is kept).  May be called again with a new n. */
nbody_status nbody_init(nbody_sim* sim, const nbody_particle* particles, size_t n);

/* Same, from structure-of-arrays host buffers: pos[3n], vel[3n] (may be
 * NULL for zero velocity), mass[n]. */
nbody_status nbody_init_soa(nbody_sim* sim, const float* pos, const float* vel,
                            const double* mass, size_t n);

/* INBodySim::Update (src/Sim/INBodySim.hpp:22) applied nsteps times:
 * kick then drift, v += a*dt; x += v*dt/scale (BruteForceCPU.cpp:61-73).
 * Enqueued on the handle's stream; returns before the device finishes. */
nbody_status nbody_step(nbody_sim* sim, float dt, int32_t nsteps);

/* Wait for every step enqueued on the handle. */
nbody_status nbody_synchronize(nbody_sim* sim);

/* Without waiting: *done = 1 when every step enqueued on the handle has
 * completed, else 0.  No counterpart in the reference (its Update is
 * synchronous); for callers that overlap several handles. */
nbody_status nbody_query(nbody_sim* sim, int32_t* done);

#define NBODY_FIELD_POSITION 1u
#define NBODY_FIELD_VELOCITY 2u
#define NBODY_FIELD_FORCES   4u
#define NBODY_FIELD_MASS     8u

/* Write the selected fields of the current state into the caller's records
 * (the reference's in-place update of `std::vector<Particle>`,
 * SimulationState.cpp:52-60).  FORCES is m * a of the last force evaluation
 * in FP64; MASS the FP64 masses (for states generated on the device).  Other
 * fields of the records are left untouched. */
nbody_status nbody_read(nbody_sim* sim, nbody_particle* particles, size_t n, uint32_t fields);

/* Structure-of-arrays read: any pointer may be NULL.  acc is the last
 * acceleration evaluated (3n floats). */
nbody_status nbody_read_soa(nbody_sim* sim, float* pos, float* vel, float* acc, size_t n);

/* Initialise with a Plummer sphere (scale a, total mass M, n equal masses,
 * recentred) generated in device memory from the same counter-based stream as
 * nbody_ic_plummer, so nothing crosses PCIe. */
nbody_status nbody_init_plummer(nbody_sim* sim, uint64_t seed, size_t n, double a, double total_mass);

/* The reference's particle seeders (EParticleSeeder, src/Sim/IParticleSeeder.hpp:12-17). */
typedef enum nbody_seeder {
    NBODY_SEEDER_RANDOM = 0,      /* RandomSeeder.cpp:14-40: cube of half-width 500, radial 1e16 velocities    */
    NBODY_SEEDER_GALAXY = 1,      /* GalaxySeeder.cpp:42-144: two spiral arms and a disc, random orientation  */
    NBODY_SEEDER_STAR_SYSTEM = 2  /* StarSystemSeeder.cpp:17-55: a 1e30 star and bodies at 4-7 AU          */
} nbody_seeder;

/* Initialise with one of the reference's seeders' distributions, drawn in
 * device memory from a counter-based stream keyed by (seed, particle): the
 * same distributions (positions, velocities, masses; positions divided by
 * scale), not the reference's MSVC random sequence.  Masses are copied back
 * once for the FORCES read-back. */
nbody_status nbody_init_seeded(nbody_sim* sim, int32_t seeder, uint64_t seed, size_t n, double scale);

/* Evaluate accelerations at the current positions without stepping; acc
 * receives 3n floats.  Used for parity checks. */
nbody_status nbody_compute_accelerations(nbody_sim* sim, float* acc, size_t n);

/* Direct-sum acceleration at m arbitrary points (points[3m] -> acc[3m]) from
 * the current particles, with the handle's law; a point that coincides with a
 * particle gets no contribution from it.  Used to check Barnes-Hut forces on
 * a subsample. */
nbody_status nbody_field(nbody_sim* sim, const float* points, size_t m, float* acc);

/* nbody_field without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code between
 * graph captures, tight serving loops, harnesses that forbid hipMalloc).  The
 * points and their accelerations are staged in a fixed device array of
 * NBODY_FIELD_NOALLOC_MAX_POINTS records each, which every device the library
 * runs on holds until the process ends (2 MiB); more points give
 * NBODY_ERR_CAPACITY.  All calls on one device share that array: calls that
 * overlap on one device (from different threads or handles) race.  Results
 * equal nbody_field's bit for bit. */
#define NBODY_FIELD_NOALLOC_MAX_POINTS 65536
nbody_status nbody_field_noalloc(nbody_sim* sim, const float* points, size_t m, float* acc);

/* The scaled units the direct sum computes a system in: out4 = {1/lambda,
 * mass factor (law R: c, with h_j = c / (G m_j); law P: 1/c), softening in
 * scaled units, c / lambda^2 (the factor that turns scaled sums back into
 * accelerations)}, all powers of two but the softening, for a system whose
 * positions span `extent` (the largest coordinate range) and whose largest
 * G m is max_gm.  Computed on the host; needs no device.  For checking the
 * pair arithmetic's range. */
nbody_status nbody_direct_scale(int32_t law, double softening, double extent, double max_gm, float* out4);

/* Number of particles of the last init. */
nbody_status nbody_size(nbody_sim* sim, size_t* n);

/* ------------------------------------------------------------------------ */
/* Diagnostics and timing                                                    */
/* ------------------------------------------------------------------------ */

typedef struct nbody_diag {
    double kinetic;              /* sum 1/2 m v^2                                    */
    double potential;            /* sum over pairs of the law's pair potential        */
    double total;                /* kinetic + potential                              */
    double momentum[3];          /* sum m v                                          */
    double angular_momentum[3];  /* sum m x cross v                                   */
    double mass;                 /* sum m                                            */
} nbody_diag;

/* FP64 energy and momenta of the current state, computed on the device. */
nbody_status nbody_diagnostics(nbody_sim* sim, nbody_diag* out);

typedef struct nbody_timing {
    int32_t steps;              /* steps in the last nbody_step call               */
    int32_t force_launches;     /* force-kernel launches timed                     */
    double  force_ms;           /* summed force-kernel time (HIP events)           */
    double  tree_ms;            /* summed tree-build time (Barnes-Hut only)        */
    double  step_ms;            /* wall time of the call on the device (events)   */
    uint64_t interactions;      /* pair or node interactions evaluated             */
    int32_t i_per_lane;         /* direct kernel launch shape in use               */
    int32_t block;
    int32_t split;
    int32_t reserved[5];
} nbody_timing;

/* Replay nbody_step from hipGraphs: each distinct (nsteps, dt) sequence is
 * captured on first use and then launched as one graph (a Barnes-Hut step is
 * about forty kernels).  Results are bit-identical to plain launches.  Not
 * used while timing is on; a new initial state drops the graphs.  No
 * counterpart in the reference; for callers that step many small handles. */
nbody_status nbody_set_graph(nbody_sim* sim, int32_t enable);

/* Octree::Theta of a live simulation, the reference's BHThetaChanged handler
 * (src/Sim/BarnesHut.cpp:29-31 sets the global Octree::Theta, Octree.cpp:5,
 * from the UI's slider, UI.cpp:105-108,193).  Waits for the steps already
 * enqueued on the handle (they keep the old theta), then every later step
 * opens cells with the new one; captured step graphs are dropped.  theta must
 * be finite and >= 0.  A direct-sum handle keeps the value and ignores it. */
nbody_status nbody_set_theta(nbody_sim* sim, double theta);

/* Turn per-kernel HIP-event timing of nbody_step on or off (default off). */
nbody_status nbody_set_timing(nbody_sim* sim, int32_t enable);

/* Timing of the last nbody_step call; synchronises the handle. */
nbody_status nbody_get_timing(nbody_sim* sim, nbody_timing* out);

/* ------------------------------------------------------------------------ */
/* Barnes-Hut introspection                                                  */
/* ------------------------------------------------------------------------ */

typedef struct nbody_bh_stats {
    uint64_t interactions;    /* cell + particle interactions since the last nbody_step began   */
    uint32_t outside_root;    /* particles clamped into a fixed root (the reference drops them) */
    uint32_t stack_overflow;  /* nonzero if a traversal stack overflowed (forces invalid)        */
    uint64_t list_entries;    /* sources evaluated by a whole wavefront (64 lanes, some masked)  */
    uint64_t cells_opened;    /* cells whose children a wavefront examined                       */
} nbody_bh_stats;

/* Counters of the Barnes-Hut steps since the last nbody_step began, or of the
 * force evaluation of the last nbody_compute_accelerations if that came later. */
nbody_status nbody_bh_stats_get(nbody_sim* sim, nbody_bh_stats* out);

/* Sorted 63-bit Morton keys and the particle order (stable) of the last tree
 * build, for checking against a CPU computation.  Either pointer may be NULL. */
nbody_status nbody_bh_tree_order(nbody_sim* sim, uint64_t* keys, uint32_t* order, size_t n);

/* One octree cell of the last tree build. */
typedef struct nbody_bh_cell {
    float    lo[3];     /* corner of the cube: root corner + index * side        */
    float    side;      /* root size / 2^level                                   */
    float    com[3];    /* centre of mass as the traversal uses it (FP32)        */
    uint32_t level;     /* 0 = root; 21 = a cell of identical keys               */
    double   mass;      /* FP64 sum of its particles' masses                     */
    uint32_t first;     /* its particles: sorted positions [first, first+count)  */
    uint32_t count;     /* (see nbody_bh_tree_order for the order)                */
} nbody_bh_cell;

/* The cells of the last tree build, for drawing or checking the octree (the
 * reference's Octree::RenderDebug draws its cubes, src/Sim/Octree.cpp:147-175,
 * called from BarnesHut::RenderDebug, src/Sim/BarnesHut.cpp:98-101).  The
 * octree is compressed: a cell is listed where its particles split among two
 * or more children (or at the root); the reference's one-particle leaves and
 * the chains of cells with a single non-empty child above a split are not
 * listed.  Cells come in depth-first order of the binary radix tree.  *count
 * receives the number of cells; at most capacity are written (out may be NULL
 * with capacity 0 to ask for the count). */
nbody_status nbody_bh_cells(nbody_sim* sim, nbody_bh_cell* out, size_t capacity, size_t* count);

/* ------------------------------------------------------------------------ */
/* Batched direct sum: many independent systems in one device arena          */
/* ------------------------------------------------------------------------ */

/* The reference advances one particle set per INBodySim; a batch advances many
 * independent systems (galaxies) with one persistent kernel per step that
 * drains a queue of (system, target block) work items, largest first, over a
 * shared device arena, and runs a captured hipGraph of the steps.  A system's
 * forces are bit-identical whether it runs alone or in any mix.  Systems may
 * be added and removed between steps (the server does so); a batch is not
 * thread-safe. */
typedef struct nbody_batch nbody_batch;

typedef struct nbody_batch_tuning {
    int32_t i_per_lane;   /* targets per lane: 1 or 2                              */
    int32_t block;        /* threads per workgroup: 128 or 256                     */
    int32_t wg_per_cu;    /* persistent workgroups launched per CU (1..64)         */
    int32_t reserved[5];
    double  ms_per_step;  /* the tuner's time for one step of this shape (0: none) */
} nbody_batch_tuning;

typedef struct nbody_batch_info {
    int32_t  systems;
    int32_t  work_items;
    uint64_t particles;
    uint64_t interactions_per_step;  /* sum of n^2 over the systems                 */
    uint64_t capacity;               /* arena records                               */
    uint64_t free_records;
    uint64_t largest_free;           /* largest system (padded) that still fits      */
    nbody_batch_tuning tuning;
} nbody_batch_info;

/* cfg->method must be NBODY_METHOD_DIRECT; capacity = arena size in particles
 * (each system takes its size rounded up to 256). */
nbody_status nbody_batch_create(const nbody_config* cfg, size_t capacity, nbody_batch** out);
nbody_status nbody_batch_destroy(nbody_batch* batch);

/* Add a system (pos[3n], vel[3n] or NULL, mass[n]); NBODY_ERR_CAPACITY when
 * the arena has no room.  *id names it until it is removed. */
nbody_status nbody_batch_add(nbody_batch* batch, const float* pos, const float* vel, const double* mass,
                             size_t n, int32_t* id);
nbody_status nbody_batch_remove(nbody_batch* batch, int32_t id);

/* Add a Plummer sphere (scale a, total mass M, n equal masses) generated in
 * the arena by the device from the counter-based stream of nbody_ic_plummer. */
nbody_status nbody_batch_add_plummer(nbody_batch* batch, uint64_t seed, size_t n, double a, double total_mass,
                                     int32_t* id);

/* Advance every system nsteps kick-drift steps (asynchronous, one graph launch). */
nbody_status nbody_batch_step(nbody_batch* batch, float dt, int32_t nsteps);

/* The baseline: each system alone, one single-system launch per step, waiting
 * for each system before the next.  split_small = 0: the plain launch (one
 * workgroup per 256 targets, all sources); 1: the single-system default,
 * which splits a small system's sources over more workgroups. */
nbody_status nbody_batch_step_sequential(nbody_batch* batch, float dt, int32_t nsteps, int32_t split_small);

nbody_status nbody_batch_synchronize(nbody_batch* batch);
/* Without waiting: *done = 1 when every step enqueued on the batch has completed. */
nbody_status nbody_batch_query(nbody_batch* batch, int32_t* done);
nbody_status nbody_batch_read(nbody_batch* batch, int32_t id, float* pos, float* vel, float* acc, size_t n);
nbody_status nbody_batch_get_info(nbody_batch* batch, nbody_batch_info* out);
nbody_status nbody_batch_set_tuning(nbody_batch* batch, const nbody_batch_tuning* tuning);

/* Time one step of each candidate (targets per lane, block, workgroups per
 * CU) on the live systems, keep the fastest, and leave the state as it was. */
nbody_status nbody_batch_tune(nbody_batch* batch, float dt, nbody_batch_tuning* out);

/* Run steps through a captured hipGraph (default) or as plain launches. */
nbody_status nbody_batch_set_graph(nbody_batch* batch, int32_t enable);

/* Device time of the last step call (events around it) and its interactions. */
nbody_status nbody_batch_get_timing(nbody_batch* batch, nbody_timing* out);

/* ------------------------------------------------------------------------ */
/* Request server: continuous batching of galaxy simulations on one GPU       */
/* ------------------------------------------------------------------------ */

/* Requests (a galaxy's size, seed and step count) are admitted in arrival
 * order as their device-memory footprint fits the budget; their states are
 * generated on the device and stay there.  Requests below bh_threshold join
 * the batch arena (one persistent graph launch per tick); larger ones run
 * Barnes-Hut on handles of their own, on their own streams, overlapping the
 * batch on the device.  Each tick advances every active request one step;
 * requests that finish leave and free their memory, and requests that arrived
 * meanwhile are admitted on the next tick.  A request's result is the same
 * alone and in the mix.  Not thread-safe. */
typedef struct nbody_server nbody_server;

typedef struct nbody_server_config {
    nbody_config sim;          /* law, G, softening, scale, theta, device (method ignored) */
    uint64_t memory_budget;    /* device bytes the request states may hold                  */
    uint64_t direct_capacity;  /* particles in the direct-sum arena                         */
    uint32_t bh_threshold;     /* requests with n >= bh_threshold use Barnes-Hut            */
    int32_t  max_active_bh;    /* concurrent Barnes-Hut requests                            */
    float    dt;               /* step of every request                                    */
    int32_t  keep_results;     /* keep final states for nbody_server_fetch                 */
    int32_t  bh_steps_ahead;   /* Barnes-Hut steps enqueued per request between checks (0: 1) */
    int32_t  inject_alloc_failures; /* test hook: the server's next k request allocations fail
                                       as a full device would (0 in production)               */
    int32_t  bh_no_graph;      /* 1: Barnes-Hut requests step with plain launches instead of
                                  replaying captured graphs (for tools that cannot follow
                                  graph capture; the results are the same)                  */
    int32_t  reserved[5];
} nbody_server_config;

typedef struct nbody_request_result {
    uint64_t id;
    uint32_t n;
    int32_t  steps;
    int32_t  method;           /* nbody_method                                    */
    int32_t  pad_;
    double   arrival;          /* server clock (seconds) at arrival, admission... */
    double   admitted;
    double   completed;
    uint64_t interactions;     /* pair (direct) or counted (Barnes-Hut)           */
} nbody_request_result;

typedef struct nbody_server_stats {
    uint64_t submitted;
    uint64_t admitted;
    uint64_t completed;
    uint64_t pending;
    uint64_t active_direct;
    uint64_t active_bh;
    uint64_t idle_handles;
    uint64_t memory_used;      /* bytes held by admitted requests and idle handles */
    uint64_t memory_peak;
    uint64_t interactions;     /* of completed requests                           */
    uint64_t ticks;            /* ticks that completed or enqueued work            */
    double   busy_seconds;     /* time spent in those ticks                       */
    uint64_t alloc_deferrals;  /* admissions put back because a device allocation
                                  failed (retried on later ticks)                  */
} nbody_server_stats;

/* law R, eps = 0.02, theta = 0.6, 64 GiB budget, 64 Mi-particle arena,
 * Barnes-Hut from 32768 particles, at most 16 Barnes-Hut requests with one
 * step enqueued at a time (deeper queues delay the batch's steps: the streams
 * share the device's hardware queues), dt = 1e-3. */
void nbody_server_config_default(nbody_server_config* cfg);
nbody_status nbody_server_create(const nbody_server_config* cfg, nbody_server** out);
nbody_status nbody_server_destroy(nbody_server* server);

/* The server's clock: seconds since nbody_server_create. */
double nbody_server_clock(nbody_server* server);

/* Queue a Plummer-sphere request; arrival is its time on the server clock
 * (a non-finite value means now). */
nbody_status nbody_server_submit_plummer(nbody_server* server, uint64_t seed, size_t n, int32_t steps, double a,
                                         double total_mass, double arrival, uint64_t* id);

/* One pass of the request loop, without waiting for the device: credit the
 * steps that have completed (the batch's last step, each Barnes-Hut handle's
 * enqueued steps), retire finished requests, admit queued ones that fit, and
 * enqueue the next batch step and the next Barnes-Hut steps of idle handles.
 * A request whose device allocation fails (NBODY_ERR_OUT_OF_MEMORY) releases
 * the idle handles and stays queued, in its place, for a later tick; the tick
 * itself does not fail.
 * Each request advances as soon as its own work completes.  *completed =
 * requests that finished in this tick.  Call it in a loop; a tick that finds
 * nothing to do sleeps for about 100 microseconds. */
nbody_status nbody_server_tick(nbody_server* server, int32_t* completed);

/* Take up to max finished-request records (oldest first). */
nbody_status nbody_server_poll(nbody_server* server, nbody_request_result* out, int32_t max, int32_t* count);

/* Final positions and velocities (3n floats each) of a finished request when
 * keep_results is set; the kept copy is released. */
nbody_status nbody_server_fetch(nbody_server* server, uint64_t id, float* pos, float* vel, size_t n);

nbody_status nbody_server_get_stats(nbody_server* server, nbody_server_stats* out);

/* ------------------------------------------------------------------------ */
/* Names, errors, version                                                    */
/* ------------------------------------------------------------------------ */

/* NBodySimGetName (src/Sim/INBodySim.cpp:29-39). */
const char* nbody_method_name(int32_t method);
const char* nbody_status_string(nbody_status status);
const char* nbody_last_error(void);
int32_t     nbody_abi_version(void);
/* Number of HIP devices visible (0 when none); never fails. */
int32_t     nbody_device_count(void);

/* ------------------------------------------------------------------------ */
/* Snapshot files (.nbody)                                                   */
/* ------------------------------------------------------------------------ */

/* The reference's snapshot is a raw array of nbody_particle records
 * (SimulationState.cpp:317-331).  count: records in the file (a trailing
 * partial record is ignored, as the reference's reader does). */
nbody_status nbody_snapshot_count(const char* path, size_t* n);

/* Read up to capacity records; recentre != 0 subtracts the centre of mass from
 * the positions as InitParticlesFromFile does (SimulationState.cpp:252-274):
 * long-double total mass, FP64 weighted sum, centre rounded to FP32. */
nbody_status nbody_snapshot_read(const char* path, nbody_particle* out, size_t capacity, size_t* n,
                                 int32_t recentre);

nbody_status nbody_snapshot_write(const char* path, const nbody_particle* particles, size_t n);

/* ------------------------------------------------------------------------ */
/* Initial conditions (host, counter-based RNG: particle i's draws depend   */
/* only on (seed, i), never on the order of generation)                     */
/* ------------------------------------------------------------------------ */

/* Plummer sphere of scale a and total mass M, isotropic velocities in
 * virial equilibrium, recentred on the centre of mass.  Outputs pos[3n],
 * vel[3n], mass[n] (equal masses M/n). */
nbody_status nbody_ic_plummer(uint64_t seed, size_t n, double a, double total_mass, double G,
                              float* pos, float* vel, double* mass);

/* Exponential disc (scale length Rd, sech^2 height z0) holding
 * (1 - bulge_fraction) of the mass on circular orbits, plus a Hernquist
 * bulge (scale ab, truncated at 50 ab) holding bulge_fraction with
 * isotropic Gaussian velocities.  Recentred on the centre of mass. */
nbody_status nbody_ic_disc_bulge(uint64_t seed, size_t n, double Rd, double z0,
                                 double bulge_fraction, double ab, double total_mass, double G,
                                 float* pos, float* vel, double* mass);

/* Two bulgeless exponential discs of n/2 particles and unit mass each, the
 * second tilted by tilt_deg about the x axis, on a parabolic relative orbit
 * at the given separation (along x) with the given pericentre. */
nbody_status nbody_ic_disc_encounter(uint64_t seed, size_t n, double Rd, double z0, double separation,
                                     double pericentre, double tilt_deg, double G, float* pos, float* vel,
                                     double* mass);

/* Uniform cube [-half, half)^3 with zero velocities and masses uniform in
 * [mass_lo, mass_hi). */
nbody_status nbody_ic_uniform(uint64_t seed, size_t n, double half, double mass_lo, double mass_hi,
                              float* pos, float* vel, double* mass);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* NBODY_H_ */
