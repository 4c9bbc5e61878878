// The drop-in in a reference-shaped program: a Plummer sphere integrated with
// kick-drift-kick velocity Verlet through the ForceSolver3D interface only
// (the reference's VelocityVerlet3D, Integrator3D.hpp:19-40), with the solver
// swapped for AdaptiveFMM3D_HIP. Prints one JSON line with the relative energy
// change, computed by O(N^2) FP64 summation (Diagnostics3D.hpp:20-38).
//   usage: drop_in_demo [N=4000] [steps=20] [p=6]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "galaxy3d/AdaptiveFMM3D_HIP.hpp"

using namespace galaxy3d;

namespace {
Bodies3D plummer(std::size_t n, unsigned seed) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    Bodies3D b;
    b.resize(n);
    auto dir = [&](double s) {
        const double ct = 2 * u(rng) - 1, st = std::sqrt(std::max(0.0, 1 - ct * ct)), ph = 2 * M_PI * u(rng);
        return Vec3{s * st * std::cos(ph), s * st * std::sin(ph), s * ct};
    };
    for (std::size_t i = 0; i < n; ++i) {
        const double r = 1.0 / std::sqrt(std::pow(std::max(u(rng), 1e-12), -2.0 / 3.0) - 1.0);
        double q, g;
        do {
            q = u(rng);
            g = 0.1 * u(rng);
        } while (g > q * q * std::pow(1 - q * q, 3.5));
        b.mass[i] = 1.0 / n;
        b.pos[i] = dir(r);
        b.vel[i] = dir(q * std::sqrt(2.0) * std::pow(r * r + 1, -0.25));
    }
    return b;
}
double energy(const Bodies3D& b, const KernelParams3D& k) {
    double e = 0;
    for (std::size_t i = 0; i < b.size(); ++i) {
        const Vec3& v = b.vel[i];
        e += 0.5 * b.mass[i] * (v.x * v.x + v.y * v.y + v.z * v.z);
        for (std::size_t j = i + 1; j < b.size(); ++j) {
            const double dx = b.pos[j].x - b.pos[i].x, dy = b.pos[j].y - b.pos[i].y, dz = b.pos[j].z - b.pos[i].z;
            e -= k.G * b.mass[i] * b.mass[j] / std::sqrt(dx * dx + dy * dy + dz * dz + k.eps * k.eps);
        }
    }
    return e;
}
}  // namespace

int main(int argc, char** argv) {
    const std::size_t n = argc > 1 ? std::strtoul(argv[1], nullptr, 10) : 4000;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 20;
    const int p = argc > 3 ? std::atoi(argv[3]) : 6;
    const KernelParams3D k{1.0, 0.02};
    Bodies3D b = plummer(n, 4);
    AdaptiveFMM3D_HIP solver(k, p, 64);
    const ForceSolver3D& s = solver;
    const double e0 = energy(b, k);
    Accelerations3D acc = s.computeAccelerations(b);
    const double dt = 0.004;
    for (int step = 0; step < steps; ++step) {
        for (std::size_t i = 0; i < n; ++i) {
            b.pos[i].x += b.vel[i].x * dt + 0.5 * dt * dt * acc[i].x;
            b.pos[i].y += b.vel[i].y * dt + 0.5 * dt * dt * acc[i].y;
            b.pos[i].z += b.vel[i].z * dt + 0.5 * dt * dt * acc[i].z;
        }
        Accelerations3D a2 = s.computeAccelerations(b);
        for (std::size_t i = 0; i < n; ++i) {
            b.vel[i].x += 0.5 * (acc[i].x + a2[i].x) * dt;
            b.vel[i].y += 0.5 * (acc[i].y + a2[i].y) * dt;
            b.vel[i].z += 0.5 * (acc[i].z + a2[i].z) * dt;
        }
        acc.swap(a2);
    }
    const double e1 = energy(b, k);
    std::printf("{\"solver\": \"%s\", \"n\": %zu, \"steps\": %d, \"p\": %d, \"dE_rel\": %.6e, \"max_leaf_occ\": %d, "
                "\"boxes\": %d}\n",
                s.name(), n, steps, p, std::fabs(e1 - e0) / std::fabs(e0), solver.lastMaxLeafOcc, solver.lastNumBoxes);
    return 0;
}
