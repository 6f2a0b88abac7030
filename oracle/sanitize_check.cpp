// sanitize_check.cpp - drives every entry point of the oracle (oracle.h) on
// small inputs so that a sanitizer build (AddressSanitizer + UBSan, and
// ThreadSanitizer for the worker pool) sees each code path: empty and
// one-particle systems, coincident particles, particles outside a fixed root,
// every worker count from 1 to 5 and row subsets.  Test infrastructure only,
// like the oracle itself: built by `make -C oracle sanitize` and run by
// tests/test_oracle_sanitize.py.
//
// Besides running clean under the sanitizers it checks what must hold whatever
// the worker count: results independent of the partition (bit for bit), row
// subsets equal to the matching rows of the full evaluation, theta = 0 giving
// the direct sum's forces, and keys sorted by the order.  Exit status 0 when
// everything holds.
#include "oracle.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

namespace {

int failures = 0;

void check(bool ok, const char* what, size_t n, int workers) {
    if (!ok) {
        std::fprintf(stderr, "FAIL %s (n=%zu workers=%d)\n", what, n, workers);
        ++failures;
    }
}

struct System {
    std::vector<float> pos;
    std::vector<double> mass;
    std::vector<double> vel;
    std::vector<float> velf;
};

// n particles in a cube of side 2*half; `twins` of them copied onto earlier
// particles (coincident pairs); `outside` of them moved beyond the cube.
System make_system(size_t n, float half, size_t twins, size_t outside, unsigned seed) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> u(-half, half);
    std::uniform_real_distribution<double> m(0.5, 2.0);
    System s;
    s.pos.resize(3 * n);
    s.mass.resize(n);
    s.vel.resize(3 * n);
    s.velf.resize(3 * n);
    for (size_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) {
            s.pos[3 * i + c] = u(rng);
            s.vel[3 * i + c] = 0.01 * u(rng);
            s.velf[3 * i + c] = static_cast<float>(s.vel[3 * i + c]);
        }
        s.mass[i] = m(rng);
    }
    for (size_t t = 0; t < twins && t + 1 < n; ++t)
        for (int c = 0; c < 3; ++c) s.pos[3 * (n - 1 - t) + c] = s.pos[3 * t + c];
    for (size_t t = 0; t < outside && t < n; ++t) s.pos[3 * (n / 2 + t) % (3 * n)] = 4.0f * half;
    return s;
}

bool same(const std::vector<double>& a, const std::vector<double>& b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(double)) == 0);
}

void run_direct(const System& s, size_t n) {
    const double G = 6.674e-11, S = 10.0;
    std::vector<double> f1(3 * n), fw(3 * n);
    oracle_ref_forces(n, s.pos.data(), s.mass.data(), G, S, nullptr, n, 1, f1.data());
    for (int w = 2; w <= 5; ++w) {
        oracle_ref_forces(n, s.pos.data(), s.mass.data(), G, S, nullptr, n, w, fw.data());
        check(same(f1, fw), "ref_forces independent of workers", n, w);
    }
    if (n > 2) {
        std::vector<int64_t> rows = {static_cast<int64_t>(n - 1), 0, static_cast<int64_t>(n / 2)};
        std::vector<double> fr(3 * rows.size());
        oracle_ref_forces(n, s.pos.data(), s.mass.data(), G, S, rows.data(), rows.size(), 3, fr.data());
        bool ok = true;
        for (size_t r = 0; r < rows.size(); ++r)
            for (int c = 0; c < 3; ++c) ok = ok && fr[3 * r + c] == f1[3 * rows[r] + c];
        check(ok, "ref_forces rows equal the full evaluation", n, 3);
    }
    for (int law = 0; law < 2; ++law) {
        std::vector<double> a1(3 * n), a4(3 * n), e1(9), e4(9);
        oracle_fp64_accel(law, n, s.pos.data(), s.mass.data(), 1.0, 0.01, nullptr, n, 1, a1.data());
        oracle_fp64_accel(law, n, s.pos.data(), s.mass.data(), 1.0, 0.01, nullptr, n, 4, a4.data());
        check(same(a1, a4), "fp64_accel independent of workers", n, 4);
        oracle_fp64_energy(law, n, s.pos.data(), s.velf.data(), s.mass.data(), 1.0, 0.01, 1.0, 1, e1.data());
        oracle_fp64_energy(law, n, s.pos.data(), s.velf.data(), s.mass.data(), 1.0, 0.01, 1.0, 4, e4.data());
        bool ok = true;
        for (int k = 0; k < 9; ++k) ok = ok && std::isfinite(e4[k]);
        check(ok, "fp64_energy finite", n, 4);
        check(std::fabs(e1[1] - e4[1]) <= 1e-12 * std::fabs(e1[1]) + 1e-300, "fp64_energy potential", n, 4);
    }
    std::vector<float> p1 = s.pos, p4 = s.pos;
    std::vector<double> v1 = s.vel, v4 = s.vel;
    oracle_ref_step(n, p1.data(), v1.data(), s.mass.data(), 1.0, 0.01, 1.0, 1e-3f, 3, 1);
    oracle_ref_step(n, p4.data(), v4.data(), s.mass.data(), 1.0, 0.01, 1.0, 1e-3f, 3, 4);
    check(p1 == p4 && same(v1, v4), "ref_step independent of workers", n, 4);
}

void run_tree(const System& s, size_t n, float half) {
    const double G = 1.0, S = 0.01;
    const float lo[3] = {-2.0f * half, -2.0f * half, -2.0f * half};
    const float size = 4.0f * half;
    for (double theta : {0.0, 0.5, 2.0}) {
        std::vector<double> f1(3 * n), f4(3 * n), fd(3 * n);
        int64_t st1[5] = {0}, st4[5] = {0};
        oracle_bh_forces(n, s.pos.data(), s.mass.data(), G, S, theta, lo, size, 1, f1.data(), st1);
        oracle_bh_forces(n, s.pos.data(), s.mass.data(), G, S, theta, lo, size, 4, f4.data(), st4);
        check(same(f1, f4), "bh_forces independent of workers", n, 4);
        check(std::memcmp(st1, st4, sizeof st1) == 0, "bh_forces stats independent of workers", n, 4);
        oracle_bh_forces_ex(n, s.pos.data(), s.mass.data(), G, S, theta, lo, size, 1, 3, fd.data(), nullptr);
        bool finite = true;
        for (double v : fd) finite = finite && std::isfinite(v);
        check(finite, "bh_forces_ex (FP64 centres) finite", n, 3);
        if (theta == 0.0 && st1[3] == 0) {
            std::vector<double> fr(3 * n);
            oracle_ref_forces(n, s.pos.data(), s.mass.data(), G, S, nullptr, n, 2, fr.data());
            double worst = 0.0;
            for (size_t i = 0; i < n; ++i) {
                double num = 0.0, den = 0.0;
                for (int c = 0; c < 3; ++c) {
                    num += (f1[3 * i + c] - fr[3 * i + c]) * (f1[3 * i + c] - fr[3 * i + c]);
                    den += fr[3 * i + c] * fr[3 * i + c];
                }
                if (den > 0.0) worst = std::fmax(worst, std::sqrt(num / den));
            }
            check(worst < 1e-4, "bh_forces at theta 0 equal the direct sum", n, 2);
        }
    }
    std::vector<int32_t> depth(n);
    oracle_bh_leaf_depths(n, s.pos.data(), s.mass.data(), lo, size, depth.data());
    std::vector<float> p = s.pos;
    std::vector<double> v = s.vel;
    oracle_bh_step(n, p.data(), v.data(), s.mass.data(), G, S, 1.0, 0.7, lo, size, 1e-3f, 2, 3);

    std::vector<uint64_t> keys(n);
    std::vector<uint32_t> order(n);
    oracle_morton_keys(n, s.pos.data(), lo, size, keys.data());
    oracle_sort_order(n, keys.data(), order.data());
    bool sorted = true;
    for (size_t i = 1; i < n; ++i)
        sorted = sorted && (keys[order[i - 1]] < keys[order[i]] ||
                            (keys[order[i - 1]] == keys[order[i]] && order[i - 1] < order[i]));
    check(sorted, "sort_order is a stable ascending order", n, 1);
    float root_lo[3];
    if (n > 0) {
        float rs = oracle_dynamic_root(n, s.pos.data(), root_lo);
        check(rs > 0.0f && std::isfinite(rs), "dynamic_root size", n, 1);
    }
}

}  // namespace

int main() {
    check(oracle_default_workers() >= 1, "default_workers", 0, 0);
    const size_t sizes[] = {0, 1, 2, 3, 17, 64, 257, 1000};
    unsigned seed = 1;
    for (size_t n : sizes) {
        for (size_t twins : {size_t(0), size_t(2)}) {
            System s = make_system(n, 10.0f, twins, 0, seed++);
            run_direct(s, n);
            run_tree(s, n, 10.0f);
        }
        if (n >= 3) {
            System s = make_system(n, 10.0f, 0, 2, seed++);
            run_tree(s, n, 10.0f);
        }
    }
    if (failures == 0) std::printf("oracle sanitize check: all entry points clean\n");
    return failures == 0 ? 0 : 1;
}
