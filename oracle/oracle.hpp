// ============================================================================
// oracle.hpp -- TEST INFRASTRUCTURE ONLY (never linked into the product).
//
// A from-scratch CPU restatement of the reference adaptive FMM force path
// (adam1sm/on-galaxy, src3d/). It is the parity checker for the HIP library:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it. Every function cites the reference file:line whose behaviour it
// restates. Arithmetic expressions keep the reference's operand order, so the
// same compiler flags give the same IEEE results (checked bit-for-bit against
// oracle/_ref, the reference compiled from its own headers).
//
// Parity pins (see tests/test_oracle_*.py):
//   tree + lists      bit-exact vs oracle/_ref and tests/golden fixtures
//   harmonics         closed forms n<=2 (fmm3d_tests.cpp:142-149)
//   gradients         central differences (fmm3d_tests.cpp:159-188)
//   P2M/M2M/M2L/L2L   reference golden error table
//                     (outputs/fmm3d_operator_convergence.csv)
//   whole solver      bit-identical forces vs oracle/_ref; N<=cap equals
//                     direct summation; error falls monotonically in p
// ============================================================================
#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace oracle {

using cplx = std::complex<double>;

struct V3 {
    double x = 0, y = 0, z = 0;
};
inline V3 sub(const V3& a, const V3& b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot3(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// One octree node (AdaptiveOctree.hpp:31-38).
struct Box {
    int parent = -1;
    int child[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    int level = 0;
    double cx = 0, cy = 0, cz = 0, half = 0;
    bool leaf = true;
    int pfirst = 0, pcount = 0;
};

// Adaptive octree + CGR interaction lists (AdaptiveOctree.hpp:40-57).
struct Tree {
    std::vector<Box> box;
    std::vector<double> px, py, pz, m;
    std::vector<int> orig;   // sorted slot -> input index
    std::vector<int> leaves; // ascending ids of non-empty leaves
    std::vector<std::vector<int>> coll, U, V, W, X;
    double tol = 1e-9;
    double epsFloor = 0.0;
    static constexpr int kMaxDepth = 24;

    void build(const double* pos, const double* mass, std::size_t n, int cap, double epsFloor);
    void buildLists();
    bool adjacent(int a, int b) const;
    int maxLeafOcc() const;
};

// ---- spherical-harmonic machinery (FmmExpansions3D.hpp:40-145) ------------
inline int sidx(int n, int m) { return n * n + n + m; }
double A(int n, int m);         // A_n^m = (-1)^n / sqrt((n-m)!(n+m)!)
double ipow_even(int e);        // i^e for even e
double harmonics(const V3& v, int L, std::vector<cplx>& Y);  // returns |v|

struct Expansion {
    V3 c;
    int p = 0;
    std::vector<cplx> a;  // (p+1)^2 coefficients, index sidx(n,m)
    Expansion() = default;
    Expansion(V3 cc, int pp) : c(cc), p(pp), a((pp + 1) * (pp + 1), cplx(0, 0)) {}
};

Expansion p2m(V3 c, const std::vector<V3>& pos, const std::vector<double>& q, int p);
Expansion m2m(const Expansion& in, V3 newc);
Expansion m2l(const Expansion& in, V3 localc);
Expansion l2l(const Expansion& in, V3 newc);
cplx multipole_potential(const Expansion& mp, V3 t);
V3 multipole_field(const Expansion& mp, V3 t);
cplx local_potential(const Expansion& lp, V3 t);
V3 local_field(const Expansion& lp, V3 t);
void p2l_accumulate(Expansion& L, const Tree& t, const Box& src, V3 cb);

// ---- solvers ---------------------------------------------------------------
struct FmmStats {
    int maxLeafOcc = 0, nBoxes = 0, nLeaves = 0, maxLevel = 0;
};
// AdaptiveFMM3D::computeAccelerations (AdaptiveFMM3D.hpp:29-110).
void adaptive_fmm(const double* pos, const double* mass, std::size_t n, double G, double eps,
                  int p, int cap, double* acc, FmmStats* st);
// DirectSolver3D::computeAccelerations (DirectSolver3D.hpp:15-34).
void direct(const double* pos, const double* mass, std::size_t n, double G, double eps, double* acc);

// ---- initial conditions ----------------------------------------------------
// makePlummerSphere (InitialConditions3D.hpp:100-152).
void plummer(std::size_t n, double a, double mtot, double G, unsigned seed, double* pos,
             double* vel, double* mass);

}  // namespace oracle
