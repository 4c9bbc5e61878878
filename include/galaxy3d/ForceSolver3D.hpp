// Re-declaration of the reference's 3D force-solver seam with identical names
// and signatures, so code written against it compiles unchanged:
//   Vec3           src3d/Vec3.hpp:11-20
//   Bodies3D       src3d/Types3D.hpp:12-24, Accelerations3D :26
//   KernelParams3D src3d/Kernel3D.hpp:24-27
//   ForceSolver3D  src3d/ForceSolver3D.hpp:20-28
// Only the members the solver seam uses are declared here.
#pragma once

#include <cstddef>
#include <vector>

namespace galaxy3d {

struct Vec3 {
    double x = 0.0, y = 0.0, z = 0.0;
    Vec3() = default;
    Vec3(double x_, double y_, double z_) : x(x_), y(y_), z(z_) {}
};

struct Bodies3D {
    std::vector<Vec3> pos;
    std::vector<Vec3> vel;
    std::vector<double> mass;
    std::size_t size() const { return mass.size(); }
    void resize(std::size_t n) {
        pos.assign(n, Vec3{});
        vel.assign(n, Vec3{});
        mass.assign(n, 0.0);
    }
};

using Accelerations3D = std::vector<Vec3>;

struct KernelParams3D {
    double G = 1.0;
    double eps = 0.05;
};

class ForceSolver3D {
public:
    virtual ~ForceSolver3D() = default;
    virtual Accelerations3D computeAccelerations(const Bodies3D& bodies) const = 0;
    virtual const char* name() const = 0;
};

}  // namespace galaxy3d
