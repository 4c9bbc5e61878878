// ============================================================================
// ics.cpp -- host-side initial conditions of the C ABI (afmm_make_plummer).
//
// makePlummerSphere (InitialConditions3D.hpp:100-152) draws from libstdc++'s
// std::mt19937 through std::uniform_real_distribution<double>, whose doubles
// are implementation-defined (generate_canonical over two 32-bit outputs).
// Generating on the host with the same library types, the same draw order and
// the same floating-point expression order gives the reference's bodies bit
// for bit for a given seed; a numpy stream cannot. Host code only: there is
// nothing here worth a device pass, and the bodies are uploaded by the
// evaluation entry points like any others. Built with g++ like the reference
// (on-galaxy_amd/Makefile): g++ turns the sin and cos of one angle into one
// sincos call, and the reference's bits are those of that call.
// ============================================================================
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>

#include "../../include/afmm.h"


extern "C" int afmm_make_plummer(int64_t n, double a, double total_mass, double G, uint32_t seed, double* pos_xyz,
                                 double* vel_xyz, double* mass) {
    if (n <= 0 || !pos_xyz || !vel_xyz || !mass) return AFMM_ERR_ARG;
    if (!(a > 0) || !std::isfinite(a) || !std::isfinite(total_mass) || !std::isfinite(G)) return AFMM_ERR_ARG;
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double mi = total_mass / static_cast<double>(n);
    const double vfac = std::sqrt(2.0 * G * total_mass);
    // isotropic unit vector (cos theta first, then the azimuth), scaled by s
    auto direction = [&](double s, double* out) {
        const double cz = 2.0 * U(gen) - 1.0;
        const double sz = std::sqrt(std::max(0.0, 1.0 - cz * cz));
        const double phi = 2.0 * M_PI * U(gen);
        out[0] = s * (sz * std::cos(phi));
        out[1] = s * (sz * std::sin(phi));
        out[2] = s * cz;
    };
    for (int64_t i = 0; i < n; ++i) {
        // enclosed-mass fraction -> radius (inverse CDF, floored away from r = 0)
        double u = U(gen);
        if (u < 1e-12) u = 1e-12;
        const double r = a / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
        // speed fraction by rejection on q^2 (1 - q^2)^{7/2}, bound 0.1
        double q, g;
        do {
            q = U(gen);
            g = 0.1 * U(gen);
        } while (g > q * q * std::pow(1.0 - q * q, 3.5));
        const double vesc = vfac * std::pow(r * r + a * a, -0.25);
        mass[i] = mi;
        direction(r, pos_xyz + 3 * i);
        direction(q * vesc, vel_xyz + 3 * i);
    }
    // centre of mass and net momentum removed, sums in body order
    double cx[3] = {0, 0, 0}, cv[3] = {0, 0, 0}, mt = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            cx[k] += mass[i] * pos_xyz[3 * i + k];
            cv[k] += mass[i] * vel_xyz[3 * i + k];
        }
        mt += mass[i];
    }
    const double inv = 1.0 / mt;
    for (int k = 0; k < 3; ++k) {
        cx[k] = inv * cx[k];
        cv[k] = inv * cv[k];
    }
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            pos_xyz[3 * i + k] -= cx[k];
            vel_xyz[3 * i + k] -= cv[k];
        }
    return AFMM_OK;
}
