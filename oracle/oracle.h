/*
 * oracle.h - CPU restatement of the reference's gravity path, used ONLY as
 * test infrastructure: by tests/, by __graft_entry__.smoke() and by the CPU
 * baseline leg of bench.py, always as the checker and never as the thing
 * measured or shipped.  The product library (libnbody.so) never links it.
 *
 * Each function cites the reference lines it restates.  Positions are FP32,
 * masses FP64, as in the reference's `Particle` (src/Render/Misc/Particle.hpp:8-18).
 * `workers` is the number of std::threads (0 = hardware concurrency); every
 * particle is handled exactly once, which is the reference's intended static
 * partition without its remainder defect (src/Sim/BruteForceCPU.cpp:47-57).
 */
#ifndef NBODY_ORACLE_H_
#define NBODY_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* BruteForceCPU::Exec (src/Sim/BruteForceCPU.cpp:25-43) for the targets
 * rows[0..nrows) (rows == NULL: all n, in order): forces F_i in FP64 with the
 * FP32 Normalize / DistanceSquared of the reference.  out: 3*nrows doubles. */
void oracle_ref_forces(size_t n, const float* pos, const double* mass, double G, double S,
                       const int64_t* rows, size_t nrows, int workers, double* out);

/* BruteForceCPU::Update (src/Sim/BruteForceCPU.cpp:45-73) repeated nsteps
 * times: forces, then v += F/m*dt (FP64), x += (float)(v*dt/scale), F = 0.
 * pos (3n floats) and vel (3n doubles) are updated in place. */
void oracle_ref_step(size_t n, float* pos, double* vel, const double* mass, double G, double S, double scale,
                     float dt, int nsteps, int workers);

/* FP64 direct-sum oracle (accelerations, long-double accumulation) for the
 * targets rows[0..nrows) (NULL: all).  law 0: a = G m_j d / (|d| (|d|^2 + soft));
 * law 1: a = G m_j d / (|d|^2 + soft)^(3/2); d = x_j - x_i; coincident pairs
 * contribute 0.  out: 3*nrows doubles. */
void oracle_fp64_accel(int law, size_t n, const float* pos, const double* mass, double G, double soft,
                       const int64_t* rows, size_t nrows, int workers, double* out);

/* FP64 energies and momenta: out[0] kinetic (sum m v^2 / (2 scale)),
 * out[1] potential (pair potential of the law), out[2..4] momentum,
 * out[5..7] angular momentum, out[8] total mass. */
void oracle_fp64_energy(int law, size_t n, const float* pos, const float* vel, const double* mass, double G,
                        double soft, double scale, int workers, double* out);

/* ---- Barnes-Hut (src/Sim/Octree.cpp, src/Sim/BarnesHut.cpp) ---------- */

/* Build the reference's pointer octree over the root cube [lo, lo+size)^3 by
 * inserting particles in index order (Octree::Add/Split, Octree.cpp:16-84),
 * compute node masses (Octree::CalculateMass, :86-105) and every particle's
 * force (Octree::CalculateForce, :107-145) with opening angle theta.
 * out_forces: 3n doubles (F, not a).  stats (may be NULL, else 5 entries):
 * [0] nodes, [1] leaves holding one particle, [2] maximum depth, [3] particles
 * dropped (outside every child cube), [4] interactions summed over the targets
 * with nodes that hold mass (a single-particle leaf other than the target's
 * own, or an accepted cell; the empty cubes the reference also evaluates add
 * zero force and are not counted). */
void oracle_bh_forces(size_t n, const float* pos, const double* mass, double G, double S, double theta,
                      const float* root_lo, float root_size, int workers, double* out_forces, int64_t* stats);

/* As oracle_bh_forces; com_fp64 != 0 sums the cells' mass-weighted positions
 * in FP64 (the device tree's arithmetic) instead of the reference's FP32
 * accumulation (Octree.cpp:99,103), with the same cells, masses and opening test. */
void oracle_bh_forces_ex(size_t n, const float* pos, const double* mass, double G, double S, double theta,
                         const float* root_lo, float root_size, int com_fp64, int workers, double* out_forces,
                         int64_t* stats);

/* BarnesHut::Update (src/Sim/BarnesHut.cpp:44-96) repeated nsteps times:
 * rebuild, forces, v += F/m*dt, x += (float)(v*dt/scale). */
void oracle_bh_step(size_t n, float* pos, double* vel, const double* mass, double G, double S, double scale,
                    double theta, const float* root_lo, float root_size, float dt, int nsteps, int workers);

/* Leaves of the reference octree that hold one particle: for each particle
 * its leaf depth (root = 0), or -1 if it was dropped.  out_depth: n ints. */
void oracle_bh_leaf_depths(size_t n, const float* pos, const double* mass, const float* root_lo, float root_size,
                           int32_t* out_depth);

/* ---- Morton keys and order (the device tree's input) -------------------- */

/* 63-bit keys over the root cube [lo, lo+size)^3 at 21 levels: the cell index
 * along each axis is floor((x - lo) / (size / 2^21)) computed exactly, clamped
 * to [0, 2^21), and the three indices are interleaved with x in the lowest bit
 * of each 3-bit group, the child order x + 2y + 4z of Octree::Split
 * (Octree.cpp:16-51).  lo must be a multiple of size / 2^21.  Returns the
 * number of particles outside the root (clamped). */
int64_t oracle_morton_keys(size_t n, const float* pos, const float* root_lo, float root_size, uint64_t* keys);

/* Stable ascending order of keys (std::stable_sort of indices). */
void oracle_sort_order(size_t n, const uint64_t* keys, uint32_t* order);

/* Dynamic root as the device computes it: the bounding box of pos grown to a
 * power-of-two cube whose corner is a multiple of its level-21 cell.
 * out_lo: 3 floats, returns the size. */
float oracle_dynamic_root(size_t n, const float* pos, float* out_lo);

/* Hardware threads the oracle's pool would use for workers = 0. */
int oracle_default_workers(void);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_ORACLE_H_ */
