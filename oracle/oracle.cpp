// CPU restatement of the reference's gravity path - TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py
// load this library, and only as the checker (or, for the baseline, as the
// timed reference CPU path).  The product library never links or calls it.
//
// What is restated, with the reference lines followed:
//   * Phys::Gravity                  src/Sim/Physics.hpp:25-47
//   * BruteForceCPU::Exec / Update   src/Sim/BruteForceCPU.cpp:25-73
//   * Octree Split/Add/CalculateMass/CalculateForce, BoundingCube::Contains
//                                    src/Sim/Octree.cpp:16-145, Octree.hpp:13-23
//   * BarnesHut::Update / Exec       src/Sim/BarnesHut.cpp:44-96, 103-110
//   * Vec3<double> += * /            src/Core/Vec3.hpp:51-81
// DirectXTK SimpleMath (an empty submodule in the reference, version unknown)
// is restated from its documented semantics as plain IEEE FP32 operations:
// Vector3 subtraction per component, Length = sqrtf(x*x + y*y + z*z),
// Normalize = v / Length (a zero vector stays zero), DistanceSquared =
// (a-b).(a-b).  The reference's thread pool partition leaves the last
// N mod W particles without a force (BruteForceCPU.cpp:56-57,
// BarnesHut.cpp:70-77); the restatement computes every particle exactly once.
//
// Built with -ffp-contract=off so that no multiply-add is fused.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <thread>
#include <vector>

namespace {

// ---- small FP32 vector with SimpleMath::Vector3 semantics ----------------
struct V3 {
    float x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) {
    const float l = length(a);
    if (!(l > 0.f)) return {0.f, 0.f, 0.f};
    return {a.x / l, a.y / l, a.z / l};
}
inline float distance_squared(V3 a, V3 b) {
    const V3 d = sub(a, b);
    return dot(d, d);
}
inline V3 load(const float* p, size_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

int resolve_workers(int w) {
    if (w > 0) return w;
    unsigned h = std::thread::hardware_concurrency();
    return h ? static_cast<int>(h) : 1;
}

// Static block partition of [0, count): every index exactly once.
template <class F>
void parallel_for(size_t count, int workers, F&& body) {
    workers = resolve_workers(workers);
    if (workers <= 1 || count < 2) {
        for (size_t i = 0; i < count; ++i) body(i);
        return;
    }
    const size_t w = std::min<size_t>(static_cast<size_t>(workers), count);
    std::vector<std::thread> threads;
    threads.reserve(w);
    for (size_t t = 0; t < w; ++t) {
        const size_t b = count * t / w, e = count * (t + 1) / w;
        threads.emplace_back([b, e, &body] {
            for (size_t i = b; i < e; ++i) body(i);
        });
    }
    for (auto& th : threads) th.join();
}

// Phys::Gravity(a, b) = -(G * a.Mass * b.Mass) / (DistanceSquared + S)  (Physics.hpp:25-35)
inline double gravity(double G, double S, V3 pa, double ma, V3 pb, double mb) {
    const double d = static_cast<double>(distance_squared(pa, pb));
    return -(G * ma * mb) / (d + S);
}

// BruteForceCPU::Exec body for one target i (BruteForceCPU.cpp:27-42).
void ref_force_one(size_t n, const float* pos, const double* mass, double G, double S, size_t i, double* f) {
    const V3 pi = load(pos, i);
    double fx = 0, fy = 0, fz = 0;
    for (size_t j = 0; j < n; ++j) {
        if (i == j) continue;
        const V3 pj = load(pos, j);            // a = particle j, b = particle i
        const V3 diff = normalize(sub(pi, pj));  // (b.Position - a.Position).Normalize()
        const double g = gravity(G, S, pj, mass[j], pi, mass[i]);
        fx += g * static_cast<double>(diff.x);
        fy += g * static_cast<double>(diff.y);
        fz += g * static_cast<double>(diff.z);
    }
    f[0] = fx;
    f[1] = fy;
    f[2] = fz;
}

// Kick then drift of one particle (BruteForceCPU.cpp:61-73, BarnesHut.cpp:81-95).
inline void kick_drift(float* p, double* v, const double* F, double m, float dt, double scale) {
    const double dtd = static_cast<double>(dt);
    for (int c = 0; c < 3; ++c) {
        const double a = F[c] / m;
        v[c] += a * dtd;
        const double dl = (v[c] * dtd) / scale;
        p[c] += static_cast<float>(dl);
    }
}

// ---- Octree (Octree.hpp:13-50, Octree.cpp) --------------------------------
struct Cube {
    float lo[3];  // TopLeft
    float hi[3];  // BottomRight
    bool contains(V3 p) const {  // Octree.hpp:18-22 (half-open)
        return p.x >= lo[0] && p.y >= lo[1] && p.z >= lo[2] && p.x < hi[0] && p.y < hi[1] && p.z < hi[2];
    }
};

struct Node {
    Cube bounds;
    float size;  // BottomRight.x - TopLeft.x at construction (Octree.cpp:7-10)
    int depth;
    int num = 0;          // NumParticles
    bool leaf = true;     // IsLeaf
    int32_t p = -1;       // P (particle index)
    double mass = 0.0;    // TotalMass
    V3 com{0.f, 0.f, 0.f};  // CentreOfMass
    double mx = 0.0, my = 0.0, mz = 0.0;  // FP64 mass-weighted position (com_fp64 mode)
    int32_t child[8];
};

struct Tree {
    std::vector<Node> nodes;
    const float* pos;
    const double* mass;
    bool com_fp64 = false;  // FP64 centre of mass instead of the reference's FP32 one
    int64_t dropped = 0;
    int max_depth = 0;

    int make(const Cube& c, int depth) {
        Node nd;
        nd.bounds = c;
        nd.size = c.hi[0] - c.lo[0];
        nd.depth = depth;
        for (int k = 0; k < 8; ++k) nd.child[k] = -1;
        nodes.push_back(nd);
        max_depth = std::max(max_depth, depth);
        return static_cast<int>(nodes.size() - 1);
    }

    // Octree::Split (Octree.cpp:16-51): children in order i = x + 2y + 4z,
    // corners advanced by repeated FP32 additions of size/2.
    void split(int id) {
        const float size = nodes[id].size / 2;
        float cur[3] = {nodes[id].bounds.lo[0], nodes[id].bounds.lo[1], nodes[id].bounds.lo[2]};
        int i = 0;
        for (int z = 0; z < 2; ++z) {
            cur[1] = nodes[id].bounds.lo[1];
            for (int y = 0; y < 2; ++y) {
                cur[0] = nodes[id].bounds.lo[0];
                for (int x = 0; x < 2; ++x, ++i) {
                    Cube c;
                    for (int a = 0; a < 3; ++a) {
                        c.lo[a] = cur[a];
                        c.hi[a] = cur[a] + size;
                    }
                    const int ch = make(c, nodes[id].depth + 1);
                    nodes[id].child[i] = ch;
                    cur[0] += size;
                }
                cur[1] += size;
            }
            cur[2] += size;
        }
        nodes[id].leaf = false;
    }

    // Octree::Add (Octree.cpp:53-84), recursion unrolled into the same decisions.
    void add(int id, int32_t p) {
        const V3 pp = load(pos, static_cast<size_t>(p));
        if (nodes[id].num > 1) {
            bool any = false;
            for (int k = 0; k < 8; ++k) {
                const int ch = nodes[id].child[k];
                if (nodes[ch].bounds.contains(pp)) {
                    add(ch, p);
                    any = true;
                }
            }
            if (!any) ++dropped;
        } else if (nodes[id].num == 1) {
            if (nodes[id].leaf) split(id);
            const int32_t P = nodes[id].p;
            const V3 pP = load(pos, static_cast<size_t>(P));
            bool any_p = false, any_P = false;
            for (int k = 0; k < 8; ++k) {
                const int ch = nodes[id].child[k];
                // A cube one FP32 ulp wide splits into children as wide as
                // itself (lo + size/2 rounds up to the next float), so two
                // particles that share such a child are never separated: the
                // reference recurses until either a child cube comes out
                // empty (hi == lo: both dropped) or, when none does, until
                // its stack overflows.  The oracle drops both at the first
                // split that makes no progress, which is the reference's
                // outcome whenever the reference ends at all.
                if (nodes[ch].size >= nodes[id].size && nodes[ch].bounds.contains(pp) &&
                    nodes[ch].bounds.contains(pP))
                    continue;
                if (nodes[ch].bounds.contains(pp)) {
                    add(ch, p);
                    any_p = true;
                }
                if (nodes[ch].bounds.contains(pP)) {
                    add(ch, P);
                    any_P = true;
                }
            }
            if (!any_p) ++dropped;
            if (!any_P) ++dropped;
            nodes[id].p = -1;
        } else {
            nodes[id].p = p;
        }
        ++nodes[id].num;
    }

    // Octree::CalculateMass (Octree.cpp:86-105): FP64 mass, FP32 centre of mass.
    // With com_fp64 the mass-weighted position is summed in FP64 and divided
    // once (the device tree's arithmetic); the cells and the masses are the same.
    void calculate_mass(int id) {
        Node& nd = nodes[id];
        if (com_fp64) {
            if (nd.num == 1) {
                const V3 q = load(pos, static_cast<size_t>(nd.p));
                nd.mass = mass[nd.p];
                nd.mx = nd.mass * q.x, nd.my = nd.mass * q.y, nd.mz = nd.mass * q.z;
                nd.com = q;
            } else if (!nd.leaf) {
                double M = 0, X = 0, Y = 0, Z = 0;
                for (int k = 0; k < 8; ++k) {
                    const int ch = nodes[id].child[k];
                    calculate_mass(ch);
                    const Node& c = nodes[ch];
                    M += c.mass, X += c.mx, Y += c.my, Z += c.mz;
                }
                Node& n2 = nodes[id];
                n2.mass = M, n2.mx = X, n2.my = Y, n2.mz = Z;
                if (M > 0.0) n2.com = {static_cast<float>(X / M), static_cast<float>(Y / M), static_cast<float>(Z / M)};
            }
            return;
        }
        if (nd.num == 1) {
            nd.com = load(pos, static_cast<size_t>(nd.p));
            nd.mass = mass[nd.p];
        } else if (!nd.leaf) {
            for (int k = 0; k < 8; ++k) {
                const int ch = nodes[id].child[k];
                calculate_mass(ch);
                Node& n2 = nodes[id];
                const Node& c = nodes[ch];
                n2.mass += c.mass;
                const float cm = static_cast<float>(c.mass);
                n2.com.x += c.com.x * cm;
                n2.com.y += c.com.y * cm;
                n2.com.z += c.com.z * cm;
            }
            Node& n2 = nodes[id];
            if (n2.mass > 0.0) {
                const float tm = static_cast<float>(n2.mass);
                n2.com.x /= tm;
                n2.com.y /= tm;
                n2.com.z /= tm;
            }
        }
    }

    // Octree::CalculateForce (Octree.cpp:107-145).
    // *count gains one per interaction with a node that holds mass (the empty
    // cubes the reference also evaluates add zero force and are not counted).
    void calculate_force(int id, int32_t p, double G, double S, double theta, double* f, int64_t* count) const {
        const Node& nd = nodes[id];
        const V3 pp = load(pos, static_cast<size_t>(p));
        if (nd.num == 1) {
            if (p != nd.p && !nd.bounds.contains(pp)) {
                const V3 pP = load(pos, static_cast<size_t>(nd.p));
                const double g = gravity(G, S, pp, mass[p], pP, mass[nd.p]);
                const V3 diff = normalize(sub(pp, pP));
                f[0] += g * static_cast<double>(diff.x);
                f[1] += g * static_cast<double>(diff.y);
                f[2] += g * static_cast<double>(diff.z);
                ++*count;
            }
        } else {
            const float r = length(sub(pp, nd.com));
            const float d = nd.bounds.hi[0] - nd.bounds.lo[0];
            if (static_cast<double>(d / r) < theta) {
                const double g = gravity(G, S, pp, mass[p], nd.com, nd.mass);
                const V3 diff = normalize(sub(pp, nd.com));
                f[0] += g * static_cast<double>(diff.x);
                f[1] += g * static_cast<double>(diff.y);
                f[2] += g * static_cast<double>(diff.z);
                if (nd.num > 0) ++*count;
            } else if (!nd.leaf) {
                for (int k = 0; k < 8; ++k) calculate_force(nd.child[k], p, G, S, theta, f, count);
            }
        }
    }
};

std::unique_ptr<Tree> build_tree(size_t n, const float* pos, const double* mass, const float* lo, float size,
                                 bool com_fp64 = false) {
    auto t = std::make_unique<Tree>();
    t->pos = pos;
    t->mass = mass;
    t->com_fp64 = com_fp64;
    t->nodes.reserve(4 * n + 16);
    Cube root;
    for (int a = 0; a < 3; ++a) {
        root.lo[a] = lo[a];
        root.hi[a] = lo[a] + size;
    }
    t->make(root, 0);
    for (size_t i = 0; i < n; ++i) t->add(0, static_cast<int32_t>(i));
    t->calculate_mass(0);
    return t;
}

// ---- Morton keys --------------------------------------------------------
inline uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
    v &= 0x1fffff;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

// size = odd * 2^p with odd an odd integer.
void decompose(float size, int64_t& odd, int& p) {
    int e;
    const float m = std::frexp(size, &e);              // size = m * 2^e, 0.5 <= m < 1
    int64_t mi = static_cast<int64_t>(std::ldexp(m, 24));  // exact 24-bit integer
    int pp = e - 24;
    while ((mi & 1) == 0) {
        mi >>= 1;
        ++pp;
    }
    odd = mi;
    p = pp;
}

inline int64_t floor_div(int64_t a, int64_t b) {
    int64_t q = a / b;
    if ((a % b != 0) && (a < 0)) --q;
    return q;
}

// Exact floor((x - lo) / (size / 2^21)) for lo a multiple of the cell.
inline int64_t cell_index(float x, float scale2, int64_t odd, int64_t base) {
    float X = x * scale2;  // exact power-of-two scaling
    X = std::fmax(std::fmin(X, 0x1p40f), -0x1p40f);
    int64_t k = static_cast<int64_t>(std::floor(X));
    if (X == 0.f && x < 0.f) k = -1;
    return floor_div(k, odd) - base;
}

}  // namespace

extern "C" {

int oracle_default_workers(void) { return resolve_workers(0); }

void oracle_ref_forces(size_t n, const float* pos, const double* mass, double G, double S, const int64_t* rows,
                       size_t nrows, int workers, double* out) {
    const size_t count = rows ? nrows : n;
    parallel_for(count, workers, [&](size_t k) {
        const size_t i = rows ? static_cast<size_t>(rows[k]) : k;
        ref_force_one(n, pos, mass, G, S, i, out + 3 * k);
    });
}

void oracle_ref_step(size_t n, float* pos, double* vel, const double* mass, double G, double S, double scale,
                     float dt, int nsteps, int workers) {
    std::vector<double> F(3 * n);
    for (int s = 0; s < nsteps; ++s) {
        oracle_ref_forces(n, pos, mass, G, S, nullptr, n, workers, F.data());
        for (size_t i = 0; i < n; ++i) kick_drift(pos + 3 * i, vel + 3 * i, F.data() + 3 * i, mass[i], dt, scale);
    }
}

void oracle_fp64_accel(int law, size_t n, const float* pos, const double* mass, double G, double soft,
                       const int64_t* rows, size_t nrows, int workers, double* out) {
    const size_t count = rows ? nrows : n;
    parallel_for(count, workers, [&](size_t k) {
        const size_t i = rows ? static_cast<size_t>(rows[k]) : k;
        const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
        long double ax = 0, ay = 0, az = 0;
        for (size_t j = 0; j < n; ++j) {
            if (j == i) continue;
            const double dx = static_cast<double>(pos[3 * j]) - xi;
            const double dy = static_cast<double>(pos[3 * j + 1]) - yi;
            const double dz = static_cast<double>(pos[3 * j + 2]) - zi;
            const double r2 = dx * dx + dy * dy + dz * dz;
            double w;
            if (law == 0) {
                if (!(r2 > 0)) continue;
                w = G * mass[j] / (std::sqrt(r2) * (r2 + soft));
            } else {
                const double q = r2 + soft;
                w = G * mass[j] / (q * std::sqrt(q));
            }
            ax += static_cast<long double>(w * dx);
            ay += static_cast<long double>(w * dy);
            az += static_cast<long double>(w * dz);
        }
        out[3 * k] = static_cast<double>(ax);
        out[3 * k + 1] = static_cast<double>(ay);
        out[3 * k + 2] = static_cast<double>(az);
    });
}

void oracle_fp64_energy(int law, size_t n, const float* pos, const float* vel, const double* mass, double G,
                        double soft, double scale, int workers, double* out) {
    std::vector<long double> phi(n, 0.0L);
    const double rs = soft > 0 ? std::sqrt(soft) : 0.0;
    const double half_pi = 1.5707963267948966;
    parallel_for(n, workers, [&](size_t i) {
        long double acc = 0;
        for (size_t j = 0; j < n; ++j) {
            if (j == i) continue;
            const double dx = static_cast<double>(pos[3 * j]) - pos[3 * i];
            const double dy = static_cast<double>(pos[3 * j + 1]) - pos[3 * i + 1];
            const double dz = static_cast<double>(pos[3 * j + 2]) - pos[3 * i + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            double g;
            if (law == 0) {
                g = rs > 0 ? (half_pi - std::atan(std::sqrt(r2) / rs)) / rs : (r2 > 0 ? 1.0 / std::sqrt(r2) : 0.0);
            } else {
                g = 1.0 / std::sqrt(r2 + soft);
            }
            acc -= static_cast<long double>(G * mass[j] * g);
        }
        phi[i] = acc;
    });
    long double K = 0, U = 0, px = 0, py = 0, pz = 0, lx = 0, ly = 0, lz = 0, M = 0;
    for (size_t i = 0; i < n; ++i) {
        const double m = mass[i];
        const double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        K += 0.5L * m * (vx * vx + vy * vy + vz * vz);
        U += 0.5L * m * phi[i];
        px += m * vx;
        py += m * vy;
        pz += m * vz;
        lx += m * (y * vz - z * vy);
        ly += m * (z * vx - x * vz);
        lz += m * (x * vy - y * vx);
        M += m;
    }
    out[0] = static_cast<double>(K / scale);
    out[1] = static_cast<double>(U);
    out[2] = static_cast<double>(px);
    out[3] = static_cast<double>(py);
    out[4] = static_cast<double>(pz);
    out[5] = static_cast<double>(lx);
    out[6] = static_cast<double>(ly);
    out[7] = static_cast<double>(lz);
    out[8] = static_cast<double>(M);
}

void oracle_bh_forces(size_t n, const float* pos, const double* mass, double G, double S, double theta,
                      const float* root_lo, float root_size, int workers, double* out_forces, int64_t* stats) {
    oracle_bh_forces_ex(n, pos, mass, G, S, theta, root_lo, root_size, 0, workers, out_forces, stats);
}

void oracle_bh_forces_ex(size_t n, const float* pos, const double* mass, double G, double S, double theta,
                         const float* root_lo, float root_size, int com_fp64, int workers, double* out_forces,
                         int64_t* stats) {
    auto t = build_tree(n, pos, mass, root_lo, root_size, com_fp64 != 0);
    std::vector<int64_t> counts(n, 0);
    parallel_for(n, workers, [&](size_t i) {
        double f[3] = {0, 0, 0};
        t->calculate_force(0, static_cast<int32_t>(i), G, S, theta, f, &counts[i]);
        out_forces[3 * i] = f[0];
        out_forces[3 * i + 1] = f[1];
        out_forces[3 * i + 2] = f[2];
    });
    if (stats) {
        int64_t leaves = 0;
        for (const Node& nd : t->nodes) leaves += (nd.num == 1);
        stats[0] = static_cast<int64_t>(t->nodes.size());
        stats[1] = leaves;
        stats[2] = t->max_depth;
        stats[3] = t->dropped;
        int64_t total = 0;
        for (int64_t c : counts) total += c;
        stats[4] = total;
    }
}

void oracle_bh_step(size_t n, float* pos, double* vel, const double* mass, double G, double S, double scale,
                    double theta, const float* root_lo, float root_size, float dt, int nsteps, int workers) {
    std::vector<double> F(3 * n);
    for (int s = 0; s < nsteps; ++s) {
        oracle_bh_forces(n, pos, mass, G, S, theta, root_lo, root_size, workers, F.data(), nullptr);
        for (size_t i = 0; i < n; ++i) kick_drift(pos + 3 * i, vel + 3 * i, F.data() + 3 * i, mass[i], dt, scale);
    }
}

void oracle_bh_leaf_depths(size_t n, const float* pos, const double* mass, const float* root_lo, float root_size,
                           int32_t* out_depth) {
    auto t = build_tree(n, pos, mass, root_lo, root_size);
    for (size_t i = 0; i < n; ++i) out_depth[i] = -1;
    for (const Node& nd : t->nodes)
        if (nd.num == 1 && nd.p >= 0) out_depth[nd.p] = nd.depth;
}

int64_t oracle_morton_keys(size_t n, const float* pos, const float* root_lo, float root_size, uint64_t* keys) {
    int64_t odd;
    int p;
    decompose(root_size, odd, p);
    const float scale2 = std::ldexp(1.0f, 21 - p);
    int64_t base[3];
    for (int a = 0; a < 3; ++a) {
        const int64_t L = static_cast<int64_t>(std::floor(root_lo[a] * scale2));
        base[a] = floor_div(L, odd);
    }
    int64_t outside = 0;
    for (size_t i = 0; i < n; ++i) {
        uint64_t key = 0;
        bool out = false;
        for (int a = 0; a < 3; ++a) {
            int64_t c = cell_index(pos[3 * i + a], scale2, odd, base[a]);
            if (c < 0) {
                c = 0;
                out = true;
            }
            if (c > 0x1fffff) {
                c = 0x1fffff;
                out = true;
            }
            key |= spread3(static_cast<uint64_t>(c)) << a;
        }
        outside += out;
        keys[i] = key;
    }
    return outside;
}

void oracle_sort_order(size_t n, const uint64_t* keys, uint32_t* order) {
    std::vector<uint32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    std::copy(idx.begin(), idx.end(), order);
}

float oracle_dynamic_root(size_t n, const float* pos, float* out_lo) {
    float mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    if (n > 0) {
        for (int a = 0; a < 3; ++a) mn[a] = mx[a] = pos[a];
        for (size_t i = 1; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                mn[a] = std::fmin(mn[a], pos[3 * i + a]);
                mx[a] = std::fmax(mx[a], pos[3 * i + a]);
            }
    }
    float ext = 0.f;
    for (int a = 0; a < 3; ++a) ext = std::fmax(ext, mx[a] - mn[a]);
    // Smallest power of two >= extent (at least 2^-100).
    int e;
    std::frexp(std::fmax(ext, 0x1p-100f), &e);  // ext < 2^e
    float size = std::ldexp(1.0f, e);
    for (int tries = 0; tries < 4; ++tries) {
        const float cell = std::ldexp(size, -21);
        bool fits = true;
        for (int a = 0; a < 3; ++a) {
            const int64_t k0 = static_cast<int64_t>(std::floor(mn[a] / cell));
            const int64_t k1 = static_cast<int64_t>(std::floor(mx[a] / cell));
            out_lo[a] = std::floor(mn[a] / cell) * cell;
            if (k1 - k0 >= (int64_t(1) << 21)) fits = false;
        }
        if (fits) break;
        size *= 2.f;
    }
    return size;
}

}  // extern "C"
