// ============================================================================
// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY. A C-ABI driver around the
// reference's own header-only CPU path, compiled by oracle/Makefile with
// -I <reference>/src3d into oracle/_ref/libref.so. Nothing of the reference is
// copied: this file only calls AdaptiveOctree, AdaptiveFMM3D, DirectSolver3D,
// makePlummerSphere, computeDiagnostics, VelocityVerlet3D and the fmm:: operators through their public interface.
// Exports mirror oracle_capi.cpp with the ref_ prefix.
// ============================================================================
#include <cstdint>
#include <vector>

#include "AdaptiveFMM3D.hpp"
#include "AdaptiveOctree.hpp"
#include "Diagnostics3D.hpp"
#include "DirectSolver3D.hpp"
#include "FmmExpansions3D.hpp"
#include "InitialConditions3D.hpp"
#include "Integrator3D.hpp"

#include "tree_export.hpp"

using namespace galaxy3d;

namespace {
Bodies3D toBodies(int64_t n, const double* pos, const double* mass) {
    Bodies3D b;
    b.resize((std::size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        b.pos[i] = Vec3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        b.mass[i] = mass[i];
    }
    return b;
}
void fromAcc(const Accelerations3D& a, double* out) {
    for (std::size_t i = 0; i < a.size(); ++i) {
        out[3 * i] = a[i].x;
        out[3 * i + 1] = a[i].y;
        out[3 * i + 2] = a[i].z;
    }
}
Vec3 v3(const double* a) { return Vec3{a[0], a[1], a[2]}; }
template <class E>
void unpack(const double* in, std::vector<fmm::Complex>& dst) {
    for (std::size_t s = 0; s < dst.size(); ++s) dst[s] = fmm::Complex(in[2 * s], in[2 * s + 1]);
}
void pack(const std::vector<fmm::Complex>& src, double* out) {
    for (std::size_t s = 0; s < src.size(); ++s) {
        out[2 * s] = src[s].real();
        out[2 * s + 1] = src[s].imag();
    }
}
}  // namespace

extern "C" {

int ref_fmm(int64_t n, const double* pos, const double* mass, double G, double eps, int p, int cap,
            double* acc, int32_t* stats) {
    const Bodies3D b = toBodies(n, pos, mass);
    AdaptiveFMM3D solver(KernelParams3D{G, eps}, p, cap);
    fromAcc(solver.computeAccelerations(b), acc);
    if (stats) {
        stats[0] = solver.lastMaxLeafOcc;
        stats[1] = solver.lastNumBoxes;
        stats[2] = -1;
        stats[3] = -1;
    }
    return 0;
}

int ref_leapfrog(int64_t n, double* pos, double* vel, const double* mass, double G, double eps, int p, int cap,
                 double dt, int nsteps, double* acc) {
    Bodies3D b = toBodies(n, pos, mass);
    for (int64_t i = 0; i < n; ++i) b.vel[i] = Vec3{vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
    AdaptiveFMM3D solver(KernelParams3D{G, eps}, p, cap);
    VelocityVerlet3D integ(solver);
    integ.initialize(b);
    for (int k = 0; k < nsteps; ++k) integ.step(b, dt);
    for (int64_t i = 0; i < n; ++i) {
        pos[3 * i] = b.pos[i].x, pos[3 * i + 1] = b.pos[i].y, pos[3 * i + 2] = b.pos[i].z;
        vel[3 * i] = b.vel[i].x, vel[3 * i + 1] = b.vel[i].y, vel[3 * i + 2] = b.vel[i].z;
    }
    fromAcc(integ.accelerations(), acc);
    return 0;
}

int ref_direct(int64_t n, const double* pos, const double* mass, double G, double eps, double* acc) {
    fromAcc(DirectSolver3D(KernelParams3D{G, eps}).computeAccelerations(toBodies(n, pos, mass)), acc);
    return 0;
}

int ref_diagnostics(int64_t n, const double* pos, const double* vel, const double* mass, double G, double eps,
                    double* out) {
    Bodies3D b = toBodies(n, pos, mass);
    for (int64_t i = 0; i < n; ++i) b.vel[i] = Vec3{vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
    const Diagnostics3D d = computeDiagnostics(b, KernelParams3D{G, eps});
    const double v[9] = {d.kinetic, d.potential, d.total, d.momentum.x, d.momentum.y, d.momentum.z,
                         d.angular.x, d.angular.y, d.angular.z};
    for (int k = 0; k < 9; ++k) out[k] = v[k];
    return 0;
}

void* ref_tree_new(int64_t n, const double* pos, const double* mass, int cap, double epsFloor) {
    AdaptiveOctree* t = new AdaptiveOctree();
    t->build(toBodies(n, pos, mass), cap, epsFloor);
    t->buildLists();
    return t;
}
void ref_tree_free(void* h) { delete static_cast<AdaptiveOctree*>(h); }
void ref_tree_sizes(void* h, int64_t* out) {
    const AdaptiveOctree& t = *static_cast<AdaptiveOctree*>(h);
    oracle_export::sizes(t.box, t.px.size(), t.leaves.size(), t.colleague, t.U, t.V, t.W, t.X, out);
}
void ref_tree_export(void* h, afmm_tree* out) {
    const AdaptiveOctree& t = *static_cast<AdaptiveOctree*>(h);
    oracle_export::boxes(t.box, out);
    oracle_export::bodies(t.px, t.py, t.pz, t.mass, t.origIndex, out);
    if (out->leaves)
        for (std::size_t k = 0; k < t.leaves.size(); ++k) out->leaves[k] = t.leaves[k];
    oracle_export::csr(t.colleague, out->colleague_off, out->colleague_idx);
    oracle_export::csr(t.U, out->u_off, out->u_idx);
    oracle_export::csr(t.V, out->v_off, out->v_idx);
    oracle_export::csr(t.W, out->w_off, out->w_idx);
    oracle_export::csr(t.X, out->x_off, out->x_idx);
    out->tol = t.tol;
}

double ref_harmonics(double x, double y, double z, int L, double* out) {
    std::vector<fmm::Complex> Y;
    const double r = fmm::computeY(Vec3{x, y, z}, L, Y);
    pack(Y, out);
    return r;
}
double ref_acoef(int n, int m) { return fmm::Acoef(n, m); }

void ref_p2m(const double* c, int ns, const double* spos, const double* q, int p, double* out) {
    std::vector<Vec3> sp(ns);
    std::vector<double> sq(q, q + ns);
    for (int i = 0; i < ns; ++i) sp[i] = v3(spos + 3 * i);
    pack(fmm::P2M(v3(c), sp, sq, p).M, out);
}
void ref_m2m(const double* in, int p, const double* oldc, const double* newc, double* out) {
    fmm::Multipole m(v3(oldc), p);
    unpack<fmm::Multipole>(in, m.M);
    pack(fmm::M2M(m, v3(newc)).M, out);
}
void ref_m2l(const double* in, int p, const double* srcc, const double* locc, double* out) {
    fmm::Multipole m(v3(srcc), p);
    unpack<fmm::Multipole>(in, m.M);
    pack(fmm::M2L(m, v3(locc)).L, out);
}
void ref_l2l(const double* in, int p, const double* oldc, const double* newc, double* out) {
    fmm::Local l(v3(oldc), p);
    unpack<fmm::Local>(in, l.L);
    pack(fmm::L2L(l, v3(newc)).L, out);
}
void ref_local_eval(const double* in, int p, const double* c, const double* t, double* pot, double* grad) {
    fmm::Local l(v3(c), p);
    unpack<fmm::Local>(in, l.L);
    const fmm::Complex ph = fmm::localPotential(l, v3(t));
    const Vec3 g = fmm::localField(l, v3(t));
    pot[0] = ph.real(), pot[1] = ph.imag();
    grad[0] = g.x, grad[1] = g.y, grad[2] = g.z;
}
void ref_multipole_eval(const double* in, int p, const double* c, const double* t, double* pot,
                        double* grad) {
    fmm::Multipole m(v3(c), p);
    unpack<fmm::Multipole>(in, m.M);
    const fmm::Complex ph = fmm::multipolePotential(m, v3(t));
    const Vec3 g = fmm::multipoleField(m, v3(t));
    pot[0] = ph.real(), pot[1] = ph.imag();
    grad[0] = g.x, grad[1] = g.y, grad[2] = g.z;
}

void ref_plummer(int64_t n, double a, double M, double G, unsigned seed, double* pos, double* vel,
                 double* mass) {
    PlummerConfig pc;
    pc.N = (std::size_t)n;
    pc.a = a;
    pc.total_mass = M;
    pc.G = G;
    pc.seed = seed;
    const Bodies3D b = makePlummerSphere(pc);
    for (int64_t i = 0; i < n; ++i) {
        pos[3 * i] = b.pos[i].x, pos[3 * i + 1] = b.pos[i].y, pos[3 * i + 2] = b.pos[i].z;
        vel[3 * i] = b.vel[i].x, vel[3 * i + 1] = b.vel[i].y, vel[3 * i + 2] = b.vel[i].z;
        mass[i] = b.mass[i];
    }
}

}  // extern "C"
