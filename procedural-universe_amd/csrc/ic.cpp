// Initial-condition generators with a counter-based random stream.
//
// Particle i's k-th uniform draw is a pure function of (seed, i, k): the
// SplitMix64 finaliser (Steele, Lea & Flood 2014) applied to a mix of the
// three, so a particle's state never depends on how many particles are
// generated, in what order or on how many threads.  The reference's seeders
// (src/Sim/RandomSeeder.cpp:12-40, GalaxySeeder.cpp:42-144) draw from
// std::default_random_engine, whose streams differ between C++ libraries; the
// generators here match distributions, not bits.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "ic.hpp"

namespace nbody {

namespace {

constexpr double kPi = 3.14159265358979323846;

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Uniform draws of one particle, strictly inside (0, 1).
struct Stream {
    uint64_t key;
    uint64_t k = 0;
    Stream(uint64_t seed, uint64_t i) : key(mix64(seed * 0x9E3779B97F4A7C15ULL + mix64(i + 0x632BE59BD9B4E019ULL))) {}
    double u() {
        const uint64_t h = mix64(key + 0xD1B54A32D192ED03ULL * (++k));
        return (static_cast<double>(h >> 11) + 0.5) * 0x1.0p-53;
    }
    double gauss() {  // Box-Muller, one of the pair
        const double a = u(), b = u();
        return std::sqrt(-2.0 * std::log(a)) * std::cos(2.0 * kPi * b);
    }
    void direction(double& x, double& y, double& z) {
        const double c = 2.0 * u() - 1.0;
        const double s = std::sqrt(std::max(0.0, 1.0 - c * c));
        const double p = 2.0 * kPi * u();
        x = s * std::cos(p);
        y = s * std::sin(p);
        z = c;
    }
};

void recentre(size_t n, float* pos, float* vel, const double* mass) {
    double m = 0, cx = 0, cy = 0, cz = 0, vx = 0, vy = 0, vz = 0;
    for (size_t i = 0; i < n; ++i) {
        m += mass[i];
        cx += mass[i] * pos[3 * i];
        cy += mass[i] * pos[3 * i + 1];
        cz += mass[i] * pos[3 * i + 2];
        vx += mass[i] * vel[3 * i];
        vy += mass[i] * vel[3 * i + 1];
        vz += mass[i] * vel[3 * i + 2];
    }
    if (m <= 0) return;
    cx /= m, cy /= m, cz /= m, vx /= m, vy /= m, vz /= m;
    for (size_t i = 0; i < n; ++i) {
        pos[3 * i] = static_cast<float>(pos[3 * i] - cx);
        pos[3 * i + 1] = static_cast<float>(pos[3 * i + 1] - cy);
        pos[3 * i + 2] = static_cast<float>(pos[3 * i + 2] - cz);
        vel[3 * i] = static_cast<float>(vel[3 * i] - vx);
        vel[3 * i + 1] = static_cast<float>(vel[3 * i + 1] - vy);
        vel[3 * i + 2] = static_cast<float>(vel[3 * i + 2] - vz);
    }
}

}  // namespace

void ic_plummer(uint64_t seed, size_t n, double a, double total_mass, double G, float* pos, float* vel,
                double* mass, double x_offset) {
    const double m = total_mass / static_cast<double>(n);
    const double vscale = std::sqrt(G * total_mass / a);
    for (size_t i = 0; i < n; ++i) {
        Stream s(seed, i);
        // Radius from the inverted cumulative mass M(r)/M = r^3 / (r^2 + a^2)^(3/2)
        // (Aarseth, Henon & Wielen 1974); reject the far tail beyond 20 a.
        double r = 0;
        for (int attempt = 0; attempt < 64; ++attempt) {
            const double U = s.u();
            r = a / std::sqrt(std::pow(U, -2.0 / 3.0) - 1.0);
            if (r <= 20.0 * a) break;
            r = 20.0 * a;
        }
        double dx, dy, dz;
        s.direction(dx, dy, dz);
        // Speed as a fraction q of the local escape speed, from
        // g(q) = q^2 (1 - q^2)^(7/2) by von Neumann rejection (max g < 0.1).
        double q = 0;
        for (int attempt = 0; attempt < 1000; ++attempt) {
            q = s.u();
            const double y = 0.1 * s.u();
            if (y < q * q * std::pow(1.0 - q * q, 3.5)) break;
        }
        const double ve = std::sqrt(2.0) * std::pow(1.0 + r * r / (a * a), -0.25) * vscale;
        double ux, uy, uz;
        s.direction(ux, uy, uz);
        pos[3 * i] = static_cast<float>(r * dx + x_offset);
        pos[3 * i + 1] = static_cast<float>(r * dy);
        pos[3 * i + 2] = static_cast<float>(r * dz);
        vel[3 * i] = static_cast<float>(q * ve * ux);
        vel[3 * i + 1] = static_cast<float>(q * ve * uy);
        vel[3 * i + 2] = static_cast<float>(q * ve * uz);
        mass[i] = m;
    }
}

void ic_disc_bulge(uint64_t seed, size_t n, double Rd, double z0, double bulge_fraction, double ab,
                   double total_mass, double G, float* pos, float* vel, double* mass) {
    const size_t nb = static_cast<size_t>(std::llround(bulge_fraction * static_cast<double>(n)));
    const size_t nd = n - nb;
    const double md = nd ? total_mass * (1.0 - bulge_fraction) / static_cast<double>(nd) : 0.0;
    const double mb = nb ? total_mass * bulge_fraction / static_cast<double>(nb) : 0.0;
    std::vector<double> phi(n), rad(n);
    for (size_t i = 0; i < n; ++i) {
        Stream s(seed, i);
        if (i < nd) {
            // Exponential surface density: R ~ Gamma(2, Rd) = -Rd ln(U1 U2);
            // sech^2 vertical profile: z = z0 atanh(2U - 1).
            double R = 0;
            for (int attempt = 0; attempt < 64; ++attempt) {
                R = -Rd * std::log(s.u() * s.u());
                if (R < 15.0 * Rd) break;
                R = 15.0 * Rd;
            }
            double z = 0;
            for (int attempt = 0; attempt < 64; ++attempt) {
                z = z0 * std::atanh(2.0 * s.u() - 1.0);
                if (std::fabs(z) < 10.0 * z0) break;
                z = 0;
            }
            const double p = 2.0 * kPi * s.u();
            pos[3 * i] = static_cast<float>(R * std::cos(p));
            pos[3 * i + 1] = static_cast<float>(R * std::sin(p));
            pos[3 * i + 2] = static_cast<float>(z);
            phi[i] = p;
            mass[i] = md;
        } else {
            // Hernquist: M(<r)/M = r^2 / (r + ab)^2  =>  r = ab sqrt(U) / (1 - sqrt(U)).
            double r = 0;
            for (int attempt = 0; attempt < 64; ++attempt) {
                const double su = std::sqrt(s.u());
                r = ab * su / (1.0 - su);
                if (r < 50.0 * ab) break;
                r = 50.0 * ab;
            }
            double dx, dy, dz;
            s.direction(dx, dy, dz);
            pos[3 * i] = static_cast<float>(r * dx);
            pos[3 * i + 1] = static_cast<float>(r * dy);
            pos[3 * i + 2] = static_cast<float>(r * dz);
            phi[i] = 0;
            mass[i] = mb;
        }
        const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        rad[i] = std::sqrt(x * x + y * y + z * z);
    }
    // Enclosed mass at spherical radius from the sampled particles.
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), size_t{0});
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return rad[a] < rad[b] || (rad[a] == rad[b] && a < b); });
    std::vector<double> rs(n), cum(n);
    double acc = 0;
    for (size_t k = 0; k < n; ++k) {
        rs[k] = rad[order[k]];
        acc += mass[order[k]];
        cum[k] = acc;
    }
    auto enclosed = [&](double r) {
        const size_t k = static_cast<size_t>(std::upper_bound(rs.begin(), rs.end(), r) - rs.begin());
        return k ? cum[k - 1] : 0.0;
    };
    for (size_t i = 0; i < n; ++i) {
        Stream s(seed ^ 0xA5A5A5A5A5A5A5A5ULL, i);
        if (i < nd) {
            const double x = pos[3 * i], y = pos[3 * i + 1];
            const double R = std::sqrt(x * x + y * y);
            const double vc = R > 0 ? std::sqrt(G * enclosed(R) / R) : 0.0;
            vel[3 * i] = static_cast<float>(-vc * std::sin(phi[i]));
            vel[3 * i + 1] = static_cast<float>(vc * std::cos(phi[i]));
            vel[3 * i + 2] = 0.f;
        } else {
            const double r = rad[i];
            const double sigma = r > 0 ? std::sqrt(G * enclosed(r) / (3.0 * r)) : 0.0;
            vel[3 * i] = static_cast<float>(sigma * s.gauss());
            vel[3 * i + 1] = static_cast<float>(sigma * s.gauss());
            vel[3 * i + 2] = static_cast<float>(sigma * s.gauss());
        }
    }
    recentre(n, pos, vel, mass);
}

void ic_uniform(uint64_t seed, size_t n, double half, double mass_lo, double mass_hi, float* pos, float* vel,
                double* mass) {
    for (size_t i = 0; i < n; ++i) {
        Stream s(seed, i);
        pos[3 * i] = static_cast<float>((2.0 * s.u() - 1.0) * half);
        pos[3 * i + 1] = static_cast<float>((2.0 * s.u() - 1.0) * half);
        pos[3 * i + 2] = static_cast<float>((2.0 * s.u() - 1.0) * half);
        vel[3 * i] = vel[3 * i + 1] = vel[3 * i + 2] = 0.f;
        mass[i] = mass_lo + (mass_hi - mass_lo) * s.u();
    }
}

void ic_recentre(size_t n, float* pos, float* vel, const double* mass) { recentre(n, pos, vel, mass); }

void ic_disc_encounter(uint64_t seed, size_t n, double Rd, double z0, double separation, double pericentre,
                       double tilt_deg, double G, float* pos, float* vel, double* mass) {
    const size_t n1 = n / 2, n2 = n - n1;
    // Two bulgeless discs of unit mass, each centred on its own centre of mass.
    ic_disc_bulge(seed, n1, Rd, z0, 0.0, 1.0, 1.0, G, pos, vel, mass);
    ic_disc_bulge(seed ^ 0x5A5A5A5A00000000ULL, n2, Rd, z0, 0.0, 1.0, 1.0, G, pos + 3 * n1, vel + 3 * n1, mass + n1);
    // Tilt the second disc's spin by rotating it about the x axis.
    const double a = tilt_deg * kPi / 180.0, c = std::cos(a), sn = std::sin(a);
    for (size_t i = n1; i < n; ++i) {
        const double y = pos[3 * i + 1], z = pos[3 * i + 2], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        pos[3 * i + 1] = static_cast<float>(c * y - sn * z);
        pos[3 * i + 2] = static_cast<float>(sn * y + c * z);
        vel[3 * i + 1] = static_cast<float>(c * vy - sn * vz);
        vel[3 * i + 2] = static_cast<float>(sn * vy + c * vz);
    }
    // Parabolic relative orbit of total mass M = 2: speed sqrt(2 G M / r0) at
    // separation r0, and angular momentum sqrt(2 G M q) for pericentre q, so the
    // velocity makes an angle with the separation whose sine is sqrt(q / r0).
    const double M = 2.0, r0 = separation;
    const double v0 = std::sqrt(2.0 * G * M / r0);
    const double sa = std::sqrt(std::min(1.0, pericentre / r0)), ca = std::sqrt(1.0 - sa * sa);
    const double rx = r0, vx = -v0 * ca, vy = v0 * sa;  // galaxy 2 relative to galaxy 1
    for (size_t i = 0; i < n; ++i) {
        const double h = (i < n1) ? -0.5 : 0.5;  // equal masses: each moves half the relative orbit
        pos[3 * i] = static_cast<float>(pos[3 * i] + h * rx);
        vel[3 * i] = static_cast<float>(vel[3 * i] + h * vx);
        vel[3 * i + 1] = static_cast<float>(vel[3 * i + 1] + h * vy);
    }
}

}  // namespace nbody
